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

// partial[c] = sum over chunk c of x*y
__global__ __launch_bounds__(256) void k_seg_dot(const SegChunk* __restrict__ ch,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 double* __restrict__ partial) {
  const SegChunk c = ch[blockIdx.x];
  double s = 0;
  for(int i = c.begin + threadIdx.x; i < c.end; i += 256) s += x[i] * y[i];
  s = chunk_sum(s);
  if(threadIdx.x == 0) partial[blockIdx.x] = s;
}

// w -= hprev[inst] * vprev (if vprev), then partial = chunk sum of w * vnext
// (if vnext): one modified Gram-Schmidt step fused with the next projection
__global__ __launch_bounds__(256) void k_seg_mgs(const SegChunk* __restrict__ ch,
                                                 const double* __restrict__ hprev,
                                                 const double* __restrict__ vprev,
                                                 double* __restrict__ w,
                                                 const double* __restrict__ vnext,
                                                 double* __restrict__ partial) {
  const SegChunk c = ch[blockIdx.x];
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

// y = a[inst] * x
__global__ __launch_bounds__(256) void k_seg_scale(const SegChunk* __restrict__ ch,
                                                   const double* __restrict__ a,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ y) {
  const SegChunk c = ch[blockIdx.x];
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

}  // namespace

void launch_seg_dot(int n_chunks, const SegChunk* ch, int n_inst, const int* inst_ptr,
                    const double* x, const double* y, double* partial, double* out, hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_dot<<<n_chunks, 256, 0, s>>>(ch, x, y, partial);
  k_seg_final<<<(n_inst + 63) / 64, 64, 0, s>>>(n_inst, inst_ptr, partial, out);
  PMCD_HIP(hipGetLastError());
}

void launch_seg_mgs(int n_chunks, const SegChunk* ch, int n_inst, const int* inst_ptr,
                    const double* hprev, const double* vprev, double* w, const double* vnext,
                    double* partial, double* out, hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_mgs<<<n_chunks, 256, 0, s>>>(ch, hprev, vprev, w, vnext, partial);
  if(vnext) k_seg_final<<<(n_inst + 63) / 64, 64, 0, s>>>(n_inst, inst_ptr, partial, out);
  PMCD_HIP(hipGetLastError());
}

void launch_seg_scale(int n_chunks, const SegChunk* ch, const double* a, const double* x, double* y,
                      hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_scale<<<n_chunks, 256, 0, s>>>(ch, a, x, y);
  PMCD_HIP(hipGetLastError());
}

void launch_seg_axpy(int n_chunks, const SegChunk* ch, const double* a, const double* x, double* y,
                     hipStream_t s) {
  if(n_chunks == 0) return;
  k_seg_axpy<<<n_chunks, 256, 0, s>>>(ch, a, x, y);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd
