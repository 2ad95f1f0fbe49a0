// Segmented vector kernels of the batched solver: many independent instances
// share one global vector, each instance owning a few contiguous row ranges
// (its active rows and its Dirichlet rows). The ranges are cut into chunks of
// at most kChunkRows rows, one block per chunk, so every launch covers all
// instances at once. Per-instance reductions sum the chunk partials in chunk
// order: the rounding order is fixed and repeatable.
#include "kernels.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace pmcd {

namespace {

__device__ inline double chunk_sum(double v) {
  __shared__ double s_w[4];
  for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if(l == 0) s_w[w] = v;
  __syncthreads();
  double r = 0;
  if(threadIdx.x == 0) r = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
  return r;
}

// partial[c] = sum over chunk c of x*y (0 for an inactive instance)
__global__ __launch_bounds__(256) void k_seg_dot(const SegChunk* __restrict__ ch,
                                                 const uint8_t* __restrict__ active,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 double* __restrict__ partial) {
  const SegChunk c = ch[blockIdx.x];
  if(active && !active[c.inst]) {
    if(threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    return;
  }
  double s = 0;
  for(int i = c.begin + threadIdx.x; i < c.end; i += 256) s += x[i] * y[i];
  s = chunk_sum(s);
  if(threadIdx.x == 0) partial[blockIdx.x] = s;
}

// w -= hprev[inst] * vprev (if vprev), then partial = chunk sum of w * vnext
// (if vnext): one modified Gram-Schmidt step fused with the next projection
__global__ __launch_bounds__(256) void k_seg_mgs(const SegChunk* __restrict__ ch,
                                                 const uint8_t* __restrict__ active,
                                                 const double* __restrict__ hprev,
                                                 const double* __restrict__ vprev,
                                                 double* __restrict__ w,
                                                 const double* __restrict__ vnext,
                                                 double* __restrict__ partial) {
  const SegChunk c = ch[blockIdx.x];
  if(active && !active[c.inst]) {
    if(vnext && threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    return;
  }
  const double h = vprev ? hprev[c.inst] : 0.0;
  double s = 0;
  for(int i = c.begin + threadIdx.x; i < c.end; i += 256) {
    double wi = w[i];
    if(vprev) {
      wi = wi - h * vprev[i];
      w[i] = wi;
    }
    if(vnext) s += wi * vnext[i];
  }
  if(vnext) {
    s = chunk_sum(s);
    if(threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// out[i] = sum of partial[inst_ptr[i] .. inst_ptr[i+1]) in chunk order
__global__ void k_seg_final(int n_inst, const int* __restrict__ inst_ptr,
                            const double* __restrict__ partial, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n_inst) return;
  double s = 0;
  for(int k = inst_ptr[i]; k < inst_ptr[i + 1]; ++k) s += partial[k];
  out[i] = s;
}

// y = a[inst] * x (0 for an inactive instance)
__global__ __launch_bounds__(256) void k_seg_scale(const SegChunk* __restrict__ ch,
                                                   const uint8_t* __restrict__ active,
                                                   const double* __restrict__ a,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ y) {
  const SegChunk c = ch[blockIdx.x];
  if(active && !active[c.inst]) {
    for(int i = c.begin + threadIdx.x; i < c.end; i += 256) y[i] = 0.0;
    return;
  }
  const double s = a[c.inst];
  for(int i = c.begin + threadIdx.x; i < c.end; i += 256) y[i] = s * x[i];
}

// y += a[inst] * x
__global__ __launch_bounds__(256) void k_seg_axpy(const SegChunk* __restrict__ ch,
                                                  const double* __restrict__ a,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y) {
  const SegChunk c = ch[blockIdx.x];
  const double s = a[c.inst];
  if(s == 0.0) return;
  for(int i = c.begin + threadIdx.x; i < c.end; i += 256) y[i] += s * x[i];
}

// Batched SpMV with the non-active rule: rows of inactive (converged)
// instances are written as 0 without reading their matrix rows; the others
// as in k_spmv_unrolled.
template <int W, int U>
__global__ __launch_bounds__(256) void k_spmv_masked(int n, int n_active, const int* __restrict__ rp,
                                                     const int* __restrict__ col,
                                                     const double* __restrict__ val,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ z, double* __restrict__ y,
                                                     double alpha, double beta,
                                                     const int* __restrict__ row_inst,
                                                     const uint8_t* __restrict__ active) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n) return;
  if(!active[row_inst[row]]) {
    if(l == 0) y[row] = 0.0;
    return;
  }
  if(row < n_active) {
    const int b = rp[row], e = rp[row + 1];
    int cc[U];
    double v[U], xv[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int k = b + l + u * W;
      cc[u] = k < e ? col[k] : -1;
    }
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int k = b + l + u * W;
      v[u] = k < e ? val[k] : 0.0;
    }
#pragma unroll
    for(int u = 0; u < U; ++u) xv[u] = cc[u] >= 0 ? x[cc[u]] : 0.0;
    double s = 0;
#pragma unroll
    for(int u = 0; u < U; ++u) s += v[u] * xv[u];
#pragma unroll
    for(int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
    if(l == 0) y[row] = (beta == 0.0 ? 0.0 : beta * z[row]) + alpha * s;
  } else if(l == 0) {
    y[row] = (beta == 0.0 ? 0.0 : beta * z[row]) + alpha * x[row];
  }
}

// ---- device-resident batched FGMRES -----------------------------------------
//
// Every instance runs Iteration_gmres::flexible_gmres (src/Solver/
// Iteration_gmres.C:430-620) on its own segments of the global vectors; the
// per-instance Hessenberg matrix, Givens rotations, convergence test
// (IterativeMethod::converged, include/Solver/IterativeMethod.h:83-162) and
// update x += sum y_k z_k stay on the device, so an iteration is a fixed
// kernel sequence (replayable as a graph) and the host only polls the
// active flags.

// sum of instance inst's chunk partials in chunk order (k_seg_final's order),
// computed by thread 0 and broadcast to the block
__device__ inline double inst_partial_sum(const int* __restrict__ inst_ptr, const double* __restrict__ partial,
                                          int inst) {
  __shared__ double s_t;
  if(threadIdx.x == 0) {
    double t = 0;
    for(int k = inst_ptr[inst]; k < inst_ptr[inst + 1]; ++k) t += partial[k];
    s_t = t;
  }
  __syncthreads();
  return s_t;
}

// modified Gram-Schmidt step k with the projection's reduction folded in: h =
// (w, v) from `prev` (stored in hcol_k by the instance's first chunk),
// w -= h v, then the partials of (w, vnext), or of (w, w) without vnext;
// the arithmetic of k_seg_final + k_seg_mgs (+ k_seg_dot for the norm)
__global__ __launch_bounds__(256) void k_seg_mgs2(const SegChunk* __restrict__ ch, const uint8_t* __restrict__ active,
                                                  const int* __restrict__ inst_ptr, const double* __restrict__ prev,
                                                  double* __restrict__ hcol_k, const double* __restrict__ v,
                                                  double* __restrict__ w, const double* __restrict__ vnext,
                                                  double* __restrict__ partial) {
  const SegChunk c = ch[blockIdx.x];
  if(!active[c.inst]) {
    if(threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    return;
  }
  const double h = inst_partial_sum(inst_ptr, prev, c.inst);
  if(threadIdx.x == 0 && (int)blockIdx.x == inst_ptr[c.inst]) hcol_k[c.inst] = h;
  double s = 0;
  for(int i = c.begin + threadIdx.x; i < c.end; i += 256) {
    const double wi = w[i] - h * v[i];
    w[i] = wi;
    s += wi * (vnext ? vnext[i] : wi);
  }
  s = chunk_sum(s);
  if(threadIdx.x == 0) partial[blockIdx.x] = s;
}

// hn = |w| from `prev` (stored in hcol by the first chunk), vout = (1/hn) w;
// inactive instances get 0 (k_seg_scale's rule)
__global__ __launch_bounds__(256) void k_seg_normscale(const SegChunk* __restrict__ ch,
                                                       const uint8_t* __restrict__ active,
                                                       const int* __restrict__ inst_ptr,
                                                       const double* __restrict__ prev, double* __restrict__ hcol,
                                                       const double* __restrict__ w, double* __restrict__ vout) {
  const SegChunk c = ch[blockIdx.x];
  if(!active[c.inst]) {
    for(int i = c.begin + threadIdx.x; i < c.end; i += 256) vout[i] = 0.0;
    return;
  }
  const double hn = sqrt(inst_partial_sum(inst_ptr, prev, c.inst));
  if(threadIdx.x == 0 && (int)blockIdx.x == inst_ptr[c.inst]) hcol[c.inst] = hn;
  const double a = 1.0 / hn;
  for(int i = c.begin + threadIdx.x; i < c.end; i += 256) vout[i] = a * w[i];
}

__device__ inline void dev_givens_gen(double dx, double dy, double& cs, double& sn) {
  if(dy == 0.0) { cs = 1.0; sn = 0.0; }
  else if(fabs(dy) > fabs(dx)) {
    const double t = dx / dy;
    sn = 1.0 / sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    const double t = dy / dx;
    cs = 1.0 / sqrt(1.0 + t * t);
    sn = t * cs;
  }
}

__device__ inline void dev_givens_app(double& dx, double& dy, double cs, double sn) {
  const double t = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = t;
}

// IterativeMethod::converged for one instance at iteration `it`
__device__ inline bool dev_converged(const BatchSolveParams& p, BatchInstState& st, double res, int it) {
  if(it == 0) { st.initial = res; st.previous = res; }
  const bool small = res < p.res_tol;
  const bool reduced = res < p.res_red * st.initial;
  const bool enough = it >= p.min_iter;
  const bool stagnation = res > 1e10 * st.previous;
  const bool div = res > p.div_factor * st.initial;
  bool conv = (small || reduced || stagnation || div) && enough;
  if(it == p.max_iter) conv = true;
  st.previous = res;
  return conv;
}

// back-solve of the (k+1) x (k+1) upper Hessenberg system (Utilities_gmres.C:
// 90-108) into y
__device__ inline void dev_back_solve(int k, int m, const double* H, const double* s, double* y) {
  for(int i = 0; i <= k; ++i) y[i] = s[i];
  for(int i = k; i >= 0; --i) {
    y[i] /= H[(size_t)i * 1 + (size_t)i * (m + 1)];
    for(int jj = i - 1; jj >= 0; --jj) y[jj] -= H[(size_t)jj + (size_t)i * (m + 1)] * y[i];
  }
}

// column i of every active instance: Hessenberg entries from hcol, the
// previous rotations, a new rotation, the residual estimate and the
// convergence test; a converged instance gets its update coefficients and
// leaves the active set
__global__ void k_batch_givens(int ni, int i, BatchSolveParams p, const double* __restrict__ hcol,
                               double* __restrict__ H, double* __restrict__ s, double* __restrict__ cs,
                               double* __restrict__ sn, BatchInstState* __restrict__ st, double* __restrict__ y,
                               uint8_t* __restrict__ active, int* __restrict__ breakdown) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if(inst >= ni) return;
  BatchInstState& S = st[inst];
  S.update = 0;
  if(!active[inst]) return;
  const int m = p.m;
  double* Hi = H + (size_t)inst * m * (m + 1);
  double* h = Hi + (size_t)i * (m + 1);  // column i
  for(int k = 0; k <= i + 1; ++k) h[k] = hcol[(size_t)k * ni + inst];
  if(!(h[i + 1] != 0.0)) {
    *breakdown = 1;
    active[inst] = 0;
    return;
  }
  double* si = s + (size_t)inst * (m + 1);
  double* ci = cs + (size_t)inst * (m + 1);
  double* ni_ = sn + (size_t)inst * (m + 1);
  for(int k = 0; k < i; ++k) dev_givens_app(h[k], h[k + 1], ci[k], ni_[k]);
  dev_givens_gen(h[i], h[i + 1], ci[i], ni_[i]);
  dev_givens_app(h[i], h[i + 1], ci[i], ni_[i]);
  dev_givens_app(si[i], si[i + 1], ci[i], ni_[i]);
  const double res = fabs(si[i + 1]);
  S.resid = res;
  S.iterations += 1;
  if(dev_converged(p, S, res, S.iterations)) {
    dev_back_solve(i, m, Hi, si, y + (size_t)inst * m);
    S.n_coef = i + 1;
    S.update = 1;
    S.converged = res < p.res_tol || res < p.res_red * S.initial;
    active[inst] = 0;
  }
}

// end of a cycle: every active instance updates with its m coefficients
__global__ void k_batch_restart_coef(int ni, BatchSolveParams p, const double* __restrict__ H,
                                     const double* __restrict__ s, BatchInstState* __restrict__ st,
                                     double* __restrict__ y, const uint8_t* __restrict__ active) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if(inst >= ni) return;
  BatchInstState& S = st[inst];
  S.update = 0;
  if(!active[inst]) return;
  const int m = p.m;
  dev_back_solve(m - 1, m, H + (size_t)inst * m * (m + 1), s + (size_t)inst * (m + 1), y + (size_t)inst * m);
  S.n_coef = m;
  S.update = 1;
}

// x += sum_{k < n_coef} y_k z_k on the instances flagged for an update: the
// terms in k order into a zero accumulator, then the sum into x (the
// arithmetic of a zero fill, k segmented axpys and one more axpy)
__global__ __launch_bounds__(1024) void k_seg_update(const SegChunk* __restrict__ ch,
                                                     const BatchInstState* __restrict__ st, int m,
                                                     const double* __restrict__ y, const double* const* __restrict__ z,
                                                     double* __restrict__ x) {
  const SegChunk c = ch[blockIdx.x];
  const BatchInstState& S = st[c.inst];
  if(!S.update) return;
  const double* yi = y + (size_t)c.inst * m;
  for(int i = c.begin + threadIdx.x; i < c.end; i += 1024) {
    double acc = 0.0;
    for(int k = 0; k < S.n_coef; ++k)
      if(yi[k] != 0.0) acc += yi[k] * z[k][i];
    x[i] += 1.0 * acc;
  }
}

// after a true residual: beta = |r| per instance (from the chunk partials);
// mode 0 (start of the solve): the test at iteration 0 on beta, every
// instance active before it; mode 1 (restart): the test on the last estimate
// (the reference tests it once more at the restart); then the start of the
// next cycle: s = (beta, 0, ...), scale = 1/beta on active instances
__global__ void k_batch_restart_state(int ni, BatchSolveParams p, const int* __restrict__ inst_ptr,
                                      const double* __restrict__ partial, BatchInstState* __restrict__ st,
                                      double* __restrict__ s, double* __restrict__ scale,
                                      uint8_t* __restrict__ active, int mode) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if(inst >= ni) return;
  BatchInstState& S = st[inst];
  double t = 0;
  for(int k = inst_ptr[inst]; k < inst_ptr[inst + 1]; ++k) t += partial[k];
  S.beta = sqrt(t);
  if(mode == 0) {
    S.iterations = 0;
    S.update = 0;
    S.n_coef = 0;
    S.resid = S.beta;
    S.converged = 0;
    active[inst] = 1;
    if(dev_converged(p, S, S.beta, 0)) {
      S.converged = 1;
      active[inst] = 0;
    }
  } else if(active[inst] && dev_converged(p, S, S.resid, S.iterations + 1)) {
    // (the reference's loop counter has passed the cycle's last iteration)
    S.converged = S.resid < p.res_tol || S.resid < p.res_red * S.initial;
    active[inst] = 0;
  }
  double* si = s + (size_t)inst * (p.m + 1);
  for(int k = 0; k <= p.m; ++k) si[k] = 0.0;
  si[0] = S.beta;
  scale[inst] = active[inst] ? 1.0 / S.beta : 0.0;
}

}  // namespace

void launch_seg_mgs2(int n_chunks, const SegChunk* ch, const uint8_t* active, const int* inst_ptr,
                     const double* prev, double* hcol_k, const double* v, double* w, const double* vnext,
                     double* partial, hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_mgs2<<<n_chunks, 256, 0, s>>>(ch, active, inst_ptr, prev, hcol_k, v, w, vnext, partial);
  PMCD_HIP(hipGetLastError());
}
void launch_seg_normscale(int n_chunks, const SegChunk* ch, const uint8_t* active, const int* inst_ptr,
                          const double* prev, double* hcol, const double* w, double* vout, hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_normscale<<<n_chunks, 256, 0, s>>>(ch, active, inst_ptr, prev, hcol, w, vout);
  PMCD_HIP(hipGetLastError());
}
void launch_seg_dot_partial(int n_chunks, const SegChunk* ch, const uint8_t* active, const double* x,
                            const double* y, double* partial, hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_dot<<<n_chunks, 256, 0, s>>>(ch, active, x, y, partial);
  PMCD_HIP(hipGetLastError());
}
void launch_batch_givens(int ni, int i, const BatchSolveParams& p, const double* hcol, double* H, double* sv,
                         double* cs, double* sn, BatchInstState* st, double* y, uint8_t* active, int* breakdown,
                         hipStream_t s) {
  k_batch_givens<<<(ni + 63) / 64, 64, 0, s>>>(ni, i, p, hcol, H, sv, cs, sn, st, y, active, breakdown);
  PMCD_HIP(hipGetLastError());
}
void launch_batch_restart_coef(int ni, const BatchSolveParams& p, const double* H, const double* sv,
                               BatchInstState* st, double* y, const uint8_t* active, hipStream_t s) {
  k_batch_restart_coef<<<(ni + 63) / 64, 64, 0, s>>>(ni, p, H, sv, st, y, active);
  PMCD_HIP(hipGetLastError());
}
void launch_seg_update(int n_chunks, const SegChunk* ch, const BatchInstState* st, int m, const double* y,
                       const double* const* z, double* x, hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_update<<<n_chunks, 1024, 0, s>>>(ch, st, m, y, z, x);
  PMCD_HIP(hipGetLastError());
}
void launch_batch_restart_state(int ni, const BatchSolveParams& p, const int* inst_ptr, const double* partial,
                                BatchInstState* st, double* sv, double* scale, uint8_t* active, int mode,
                                hipStream_t s) {
  k_batch_restart_state<<<(ni + 63) / 64, 64, 0, s>>>(ni, p, inst_ptr, partial, st, sv, scale, active, mode);
  PMCD_HIP(hipGetLastError());
}

void launch_spmv_masked(int n, int n_active, const int* rp, const int* col, const double* val,
                        const double* x, const double* z, double* y, double alpha, double beta,
                        int max_row, const int* row_inst, const uint8_t* active, hipStream_t s) {
  if(n == 0) return;
  constexpr int W = 8;
  PMCD_CHECK(max_row <= 8 * W, "batched rows longer than ", 8 * W, " entries");
  const int blocks = (int)(((long)n * W + 255) / 256);
  if(max_row <= 2 * W) k_spmv_masked<W, 2><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta, row_inst, active);
  else if(max_row <= 4 * W) k_spmv_masked<W, 4><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta, row_inst, active);
  else k_spmv_masked<W, 8><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta, row_inst, active);
  PMCD_HIP(hipGetLastError());
}

void launch_seg_dot(int n_chunks, const SegChunk* ch, int n_inst, const int* inst_ptr,
                    const double* x, const double* y, double* partial, double* out, hipStream_t s,
                    const uint8_t* active) {
  if(n_chunks == 0) return;
  k_seg_dot<<<n_chunks, 256, 0, s>>>(ch, active, x, y, partial);
  k_seg_final<<<(n_inst + 63) / 64, 64, 0, s>>>(n_inst, inst_ptr, partial, out);
  PMCD_HIP(hipGetLastError());
}

void launch_seg_mgs(int n_chunks, const SegChunk* ch, int n_inst, const int* inst_ptr,
                    const double* hprev, const double* vprev, double* w, const double* vnext,
                    double* partial, double* out, hipStream_t s, const uint8_t* active) {
  if(n_chunks == 0) return;
  k_seg_mgs<<<n_chunks, 256, 0, s>>>(ch, active, hprev, vprev, w, vnext, partial);
  if(vnext) k_seg_final<<<(n_inst + 63) / 64, 64, 0, s>>>(n_inst, inst_ptr, partial, out);
  PMCD_HIP(hipGetLastError());
}

void launch_seg_scale(int n_chunks, const SegChunk* ch, const double* a, const double* x, double* y,
                      hipStream_t s, const uint8_t* active) {
  if(n_chunks == 0) return;
  k_seg_scale<<<n_chunks, 256, 0, s>>>(ch, active, a, x, y);
  PMCD_HIP(hipGetLastError());
}

void launch_seg_axpy(int n_chunks, const SegChunk* ch, const double* a, const double* x, double* y,
                     hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_axpy<<<n_chunks, 256, 0, s>>>(ch, a, x, y);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd
