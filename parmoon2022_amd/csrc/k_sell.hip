// Sliced ELL matrix kernels: one wavefront per slice of 64 active rows, one
// lane per row. A load instruction reads 64 consecutive values (and column
// indices) of the slice, and every row is summed serially in its CSR column
// order, the order of TMatrix::multiply (src/Matrix/Matrix.C:490-512): the
// row sums are bitwise those of a serial CSR loop. Non-active rows (the
// Dirichlet tail) take extra threads after the slices and follow the
// non-active rule of BlockFEMatrix::apply_scaled_add (BlockFEMatrix.C:1008-1056).
#include "kernels.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace pmcd {

namespace {

// sum over the row in `slot`: U entries of the lane loaded per step, then
// added in order
template <int U, class T>
__device__ inline T sell_row_sum(const SellView<T>& m, int slot, int len, const T* __restrict__ x) {
  const long long base = m.off[slot / kSellRows] + (slot % kSellRows);
  T s = 0;
  for(int k0 = 0; k0 < len; k0 += U) {
    int c[U];
    T v[U], xv[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int k = k0 + u;
      c[u] = k < len ? m.col[base + (long long)k * kSellRows] : -1;
      v[u] = k < len ? m.val[base + (long long)k * kSellRows] : T(0);
    }
#pragma unroll
    for(int u = 0; u < U; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : T(0);
#pragma unroll
    for(int u = 0; u < U; ++u)
      if(k0 + u < len) s += v[u] * xv[u];
  }
  return s;
}

// y[r] = beta z[r] + alpha (A x)[r] on active rows, beta z[r] + alpha x[r] on the tail
template <int U, class T>
__global__ __launch_bounds__(256) void k_spmv_sell(int n, int n_active, SellView<T> m, const T* __restrict__ x,
                                                   const T* __restrict__ z, T* __restrict__ y, T alpha, T beta) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  if(gt < m.n_slots) {
    const int row = m.perm[gt];
    if(row < 0) return;
    const int len = m.len[gt];
    const T zr = beta == T(0) ? T(0) : z[row];
    const T s = sell_row_sum<U, T>(m, gt, len, x);
    y[row] = (beta == T(0) ? T(0) : beta * zr) + alpha * s;
  } else {
    const int r = n_active + (gt - m.n_slots);
    if(r >= n) return;
    y[r] = (beta == T(0) ? T(0) : beta * z[r]) + alpha * x[r];
  }
}

// x_out = x + omega D^{-1} (b - A x), the non-active rows with A = I
template <int U, class T>
__global__ __launch_bounds__(256) void k_jacobi_sell(int n, int n_active, SellView<T> m,
                                                     const T* __restrict__ diag, const T* __restrict__ b,
                                                     const T* __restrict__ x, T* __restrict__ x_out, T omega) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  int row;
  T s;
  if(gt < m.n_slots) {
    row = m.perm[gt];
    if(row < 0) return;
    const int len = m.len[gt];
    s = sell_row_sum<U, T>(m, gt, len, x);
  } else {
    row = n_active + (gt - m.n_slots);
    if(row >= n) return;
    s = x[row];
  }
  T d = b[row] + T(-1) * s;
  d = d / diag[row];
  x_out[row] = x[row] + omega * d;
}

// copy CSR entries (values or column indices) of the active rows into the
// sliced layout; padding positions are left as they are (never read)
template <class T>
__global__ __launch_bounds__(256) void k_sell_fill(int n_slots, const long long* __restrict__ off,
                                                   const int* __restrict__ perm, const int* __restrict__ len,
                                                   const int* __restrict__ rp, const T* __restrict__ src,
                                                   T* __restrict__ dst) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  if(gt >= n_slots) return;
  const int row = perm[gt];
  if(row < 0) return;
  const long long base = off[gt / kSellRows] + (gt % kSellRows);
  const int b = rp[row], l = len[gt];
  for(int k = 0; k < l; ++k) dst[base + (long long)k * kSellRows] = src[b + k];
}

template <class T>
int sell_blocks(int n, int n_active, const SellView<T>& m) {
  return (int)(((long)m.n_slots + (n - n_active) + 255) / 256);
}

}  // namespace

static int g_sell_unroll = 16;
static int g_sell_window = 4;
static int g_sell_min_rows = 0;
int sell_min_rows() { return g_sell_min_rows; }
void set_sell_min_rows(int n) {
  PMCD_CHECK(n >= 0, "sliced ELL row threshold must be non-negative");
  g_sell_min_rows = n;
}
int sell_window() { return g_sell_window * kSellRows; }
void set_sell_window(int slices) {
  PMCD_CHECK(slices >= 1 && slices <= 1024, "sliced ELL window must be 1..1024 slices");
  g_sell_window = slices;
}
int get_sell_window_slices() { return g_sell_window; }
int sell_unroll() { return g_sell_unroll; }
void set_sell_unroll(int u) {
  PMCD_CHECK(u == 4 || u == 8 || u == 16, "sliced ELL unroll must be 4, 8 or 16");
  g_sell_unroll = u;
}

template <class T>
void launch_spmv_sell(int n, int n_active, const SellView<T>& m, const T* x, const T* z, T* y, T alpha, T beta,
                      hipStream_t s) {
  if(n == 0) return;
  const int blocks = sell_blocks(n, n_active, m);
  switch(g_sell_unroll) {
    case 4: k_spmv_sell<4, T><<<blocks, 256, 0, s>>>(n, n_active, m, x, z, y, alpha, beta); break;
    case 16: k_spmv_sell<16, T><<<blocks, 256, 0, s>>>(n, n_active, m, x, z, y, alpha, beta); break;
    default: k_spmv_sell<8, T><<<blocks, 256, 0, s>>>(n, n_active, m, x, z, y, alpha, beta); break;
  }
  PMCD_HIP(hipGetLastError());
}

template <class T>
void launch_jacobi_sell(int n, int n_active, const SellView<T>& m, const T* diag, const T* b, const T* x, T* x_out,
                        T omega, hipStream_t s) {
  if(n == 0) return;
  const int blocks = sell_blocks(n, n_active, m);
  switch(g_sell_unroll) {
    case 4: k_jacobi_sell<4, T><<<blocks, 256, 0, s>>>(n, n_active, m, diag, b, x, x_out, omega); break;
    case 16: k_jacobi_sell<16, T><<<blocks, 256, 0, s>>>(n, n_active, m, diag, b, x, x_out, omega); break;
    default: k_jacobi_sell<8, T><<<blocks, 256, 0, s>>>(n, n_active, m, diag, b, x, x_out, omega); break;
  }
  PMCD_HIP(hipGetLastError());
}

template <class T>
void launch_sell_fill(int n_slots, const long long* off, const int* perm, const int* len,
                      const int* rp, const T* src, T* dst, hipStream_t s) {
  if(n_slots == 0) return;
  k_sell_fill<T><<<(n_slots + 255) / 256, 256, 0, s>>>(n_slots, off, perm, len, rp, src, dst);
  PMCD_HIP(hipGetLastError());
}

template void launch_spmv_sell<double>(int, int, const SellView<double>&, const double*, const double*, double*,
                                       double, double, hipStream_t);
template void launch_spmv_sell<float>(int, int, const SellView<float>&, const float*, const float*, float*, float,
                                      float, hipStream_t);
template void launch_jacobi_sell<double>(int, int, const SellView<double>&, const double*, const double*,
                                         const double*, double*, double, hipStream_t);
template void launch_jacobi_sell<float>(int, int, const SellView<float>&, const float*, const float*, const float*,
                                        float*, float, hipStream_t);
template void launch_sell_fill<double>(int, const long long*, const int*, const int*, const int*,
                                       const double*, double*, hipStream_t);
template void launch_sell_fill<int>(int, const long long*, const int*, const int*, const int*, const int*,
                                    int*, hipStream_t);

}  // namespace pmcd
