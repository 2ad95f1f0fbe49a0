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

}  // namespace

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
