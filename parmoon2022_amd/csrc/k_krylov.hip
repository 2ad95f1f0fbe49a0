// Conjugate gradients with every scalar kept on the device: one iteration is
// two kernels. Each kernel writes one partial sum per block; every block of
// the next kernel reads all of them and sums them in block order, so every
// block derives the same scalar (alpha in the update kernel; the stopping
// test, rho and beta in the direction kernel) without atomics, fences or a
// separate reduction launch. The host launches iterations in chunks and only
// reads the state between chunks; iterations launched after the stop are
// no-ops, so the iterate, the residual and the iteration count are those of
// the loop that tests after every iteration.
//
// Reference: Iteration_cg::iterate (src/Solver/Iteration_cg.C:27-147) with
// the stopping test of IterativeMethod::converged
// (include/Solver/IterativeMethod.h:83-162).
#include "kernels.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace pmcd {

namespace {

// block-wide sum, returned to every thread
__device__ inline double cg_block_sum(double v) {
  __shared__ double s_w[4];
  for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if(l == 0) s_w[w] = v;
  __syncthreads();
  return (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

// sum of p[0..m) in a fixed order (the same in every block)
__device__ inline double cg_partials(const double* __restrict__ p, int m) {
  double s = 0;
  for(int b = threadIdx.x; b < m; b += 256) s += p[b];
  return cg_block_sum(s);
}

// Matrix row, z and p_old entries one W-lane group needs for one row; loaded
// before the scalars of the iteration are known, so that the load latency
// overlaps the reduction of the previous kernel's partials.
template <int U>
struct CgRow {
  int b, e;
  double zr, pr;     // z[row], p_old[row]
  double v[U], zv[U], pv[U];
  int cc[U];
};

template <int W, int U>
__device__ inline void cg_load_row(CgRow<U>& R, int row, int l, int n, int n_active, bool first,
                                   const int* __restrict__ rp, const int* __restrict__ col,
                                   const double* __restrict__ val, const double* __restrict__ z,
                                   const double* __restrict__ p_old) {
  if(row >= n) return;
  R.zr = z[row];
  R.pr = first ? 0.0 : p_old[row];
  R.b = R.e = 0;
  if(row < n_active) {
    R.b = rp[row];
    R.e = rp[row + 1];
  }
#pragma unroll
  for(int u = 0; u < U; ++u) {
    const int k = R.b + l + u * W;
    R.cc[u] = k < R.e ? col[k] : -1;
    R.v[u] = k < R.e ? val[k] : 0.0;
  }
#pragma unroll
  for(int u = 0; u < U; ++u) {
    const int c = R.cc[u];
    R.zv[u] = c < 0 ? 0.0 : z[c];
    R.pv[u] = (c < 0 || first) ? 0.0 : p_old[c];
  }
}

template <int W, int U>
__global__ __launch_bounds__(256) void k_cg_direction(int it, int n, int n_active, const int* __restrict__ rp,
                                                      const int* __restrict__ col,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ z,
                                                      const double* __restrict__ p_old,
                                                      double* __restrict__ p_new, double* __restrict__ q,
                                                      const double* __restrict__ pu, int nbu,
                                                      double* __restrict__ pd, CgState* st) {
  const bool first = it == 1;
  const int l = threadIdx.x % W;
  constexpr int RPB = 256 / W;
  const int stride = gridDim.x * RPB;
  int row = blockIdx.x * RPB + (int)threadIdx.x / W;
  CgRow<U> R;
  cg_load_row<W, U>(R, row, l, n, n_active, first, rp, col, val, z, p_old);
  if(st->done) return;
  double beta = 0.0;
  if(!first) {
    // stopping test of iteration k = it - 1 (IterativeMethod::converged)
    const int k = it - 1;
    const double rr = cg_partials(pu, nbu);
    const double rz = cg_partials(pu + nbu, nbu);
    const double res = sqrt(rr) / st->normb;
    const double init = st->initial;
    const bool small = res < st->res_tol;
    const bool reduced = res < st->res_red * init;
    const bool enough = k >= st->min_iter;
    const bool stagnation = res > 1e10 * st->res[(k - 1) & 1];
    const bool diverged = res > st->div_factor * init;
    // divergence ends the solve with an error whatever the iteration count
    const bool stop = ((small || reduced || stagnation) && enough) || k == st->max_iter || diverged;
    if(stop) {
      if(blockIdx.x == 0 && threadIdx.x == 0) {
        st->iter = k;
        st->final_res = res;
        st->done = diverged ? 2 : 1;
      }
      return;
    }
    beta = rz / st->rho[k & 1];
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->res[k & 1] = res;
      st->rho[it & 1] = rz;
    }
  }
  double acc = 0;
  while(row < n) {
    // p_new = beta p_old + z (z on the first iteration), formed where used
    const double pr = first ? R.zr : beta * R.pr + R.zr;
    double qr = pr;
    if(row < n_active) {
      double s = 0;
#pragma unroll
      for(int u = 0; u < U; ++u) s += R.v[u] * (first ? R.zv[u] : beta * R.pv[u] + R.zv[u]);
      for(int k = R.b + l + U * W; k < R.e; k += W) {
        const int c = col[k];
        s += val[k] * (first ? z[c] : beta * p_old[c] + z[c]);
      }
#pragma unroll
      for(int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
      qr = s;
    }
    if(l == 0) {
      p_new[row] = pr;
      q[row] = qr;
      acc += pr * qr;
    }
    row += stride;
    cg_load_row<W, U>(R, row, l, n, n_active, first, rp, col, val, z, p_old);
  }
  const double part = cg_block_sum(acc);
  if(threadIdx.x == 0) pd[blockIdx.x] = part;
}

__global__ __launch_bounds__(256) void k_cg_update(int it, int n, const double* __restrict__ p,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ diag, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ z,
                                                   const double* __restrict__ pd, int nbd,
                                                   double* __restrict__ pu, CgState* st, int jacobi) {
  const int stride = gridDim.x * 256;
  int i = blockIdx.x * 256 + threadIdx.x;
  // the first row's operands are loaded before alpha is known
  double xi = 0, pi = 0, ri = 0, qi = 0, di = 1;
  auto load = [&](int j) {
    if(j >= n) return;
    xi = x[j];
    pi = p[j];
    ri = r[j];
    qi = q[j];
    if(jacobi) di = diag[j];
  };
  load(i);
  if(st->done) return;
  const double alpha = st->rho[it & 1] / cg_partials(pd, nbd);
  double rr = 0, rz = 0;
  while(i < n) {
    x[i] = xi + alpha * pi;
    const double rn = ri - alpha * qi;
    r[i] = rn;
    const double zn = jacobi ? rn / di : rn;
    z[i] = zn;
    rr += rn * rn;
    rz += rn * zn;
    i += stride;
    load(i);
  }
  rr = cg_block_sum(rr);
  rz = cg_block_sum(rz);
  if(threadIdx.x == 0) {
    pu[blockIdx.x] = rr;
    pu[gridDim.x + blockIdx.x] = rz;
  }
}

template <int W, int U>
void direction(int blocks, int it, int n, int n_active, const int* rp, const int* col, const double* val,
               const double* z, const double* p_old, double* p_new, double* q, const double* pu, int nbu,
               double* pd, CgState* st, hipStream_t s) {
  k_cg_direction<W, U><<<blocks, 256, 0, s>>>(it, n, n_active, rp, col, val, z, p_old, p_new, q, pu, nbu, pd, st);
}

}  // namespace

int cg_direction_blocks(int n, int max_row) {
  const int W = max_row <= 16 ? 4 : (max_row <= 32 ? 8 : (max_row <= 128 ? 16 : 64));
  const long b = ((long)n * W + 255) / 256;
  return (int)std::max(1L, std::min(b, (long)cg_block_cap()));
}

int cg_update_blocks(int n) { return std::max(1, std::min((n + 255) / 256, kCgUpdateBlocks)); }

void launch_cg_direction(int it, int n, int n_active, const int* rp, const int* col, const double* val,
                         int max_row, const double* z, const double* p_old, double* p_new, double* q,
                         const double* pu, int nbu, double* pd, CgState* st, hipStream_t s) {
  if(n == 0) return;
  PMCD_CHECK(nbu == cg_update_blocks(n), "partials of the update kernel: ", nbu, " blocks");
  const int b = cg_direction_blocks(n, max_row);
#define PMCD_CG_DIR(W, U) direction<W, U>(b, it, n, n_active, rp, col, val, z, p_old, p_new, q, pu, nbu, pd, st, s)
  if(max_row <= 8) PMCD_CG_DIR(4, 2);
  else if(max_row <= 16) PMCD_CG_DIR(4, 4);
  else if(max_row <= 32) PMCD_CG_DIR(8, 4);
  else if(max_row <= 64) PMCD_CG_DIR(16, 4);
  else if(max_row <= 128) PMCD_CG_DIR(16, 8);
  else PMCD_CG_DIR(64, 4);
#undef PMCD_CG_DIR
  PMCD_HIP(hipGetLastError());
}

void launch_cg_update(int it, int n, const double* p, const double* q, const double* diag, double* x, double* r,
                      double* z, const double* pd, int nbd, double* pu, CgState* st, hipStream_t s) {
  if(n == 0) return;
  PMCD_CHECK(nbd >= 1 && nbd <= kCgBlocks, "partials of the direction kernel: ", nbd, " blocks");
  k_cg_update<<<cg_update_blocks(n), 256, 0, s>>>(it, n, p, q, diag, x, r, z, pd, nbd, pu, st, diag != nullptr);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd
