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

// Launch layout: blocks of 256 threads (4 slices); the first slot_blocks
// blocks cover the slots, the rest the non-active tail rows in order.
//
// mode bit 0 (XCD-contiguous slices): the hardware hands blocks to the 8 XCDs
// round-robin; the slot blocks are renumbered so that XCD k works through one
// contiguous range of slices, keeping the x entries its rows share in its own
// L2 instead of every XCD fetching most of x.
//
// mode bit 1 (row-ordered output): a block's 256 slots are the rows of one
// 256-row window, sorted by length; the row sums go to LDS at the row's
// position in the window, and the row-wise reads (b, diag, z, the row's own
// x) and the output are done by thread t for row w0 + t, coalesced, instead
// of at the sorted positions (partial cache lines, written by several
// waves). The arithmetic per row is the same, so the bits are too.
__device__ inline int sell_block(int bid, int slot_blocks, int mode) {
  if(!(mode & 1) || bid >= slot_blocks) return bid;
  const int per = slot_blocks / 8;
  if(bid >= per * 8) return bid;
  return (bid % 8) * per + bid / 8;
}

// y[r] = beta z[r] + alpha (A x)[r] on active rows, beta z[r] + alpha x[r] on the tail
template <int U, class T>
__global__ __launch_bounds__(256) void k_spmv_sell(int n, int n_active, SellView<T> m, int slot_blocks, int mode,
                                                   const T* __restrict__ x, const T* __restrict__ z,
                                                   T* __restrict__ y, T alpha, T beta) {
  const int blk = sell_block(blockIdx.x, slot_blocks, mode);
  if(blk < slot_blocks) {
    const int gt = blk * 256 + threadIdx.x;
    int row = -1;
    T s = 0;
    if(gt < m.n_slots) {
      row = m.perm[gt];
      if(row >= 0) s = sell_row_sum<U, T>(m, gt, m.len[gt], x);
    }
    if(mode & 2) {
      __shared__ T sh[256];
      const int w0 = blk * 256;
      if(row >= 0) sh[row - w0] = s;
      __syncthreads();
      row = w0 + threadIdx.x;
      if(row >= n_active) return;
      s = sh[threadIdx.x];
    } else if(row < 0) {
      return;
    }
    y[row] = (beta == T(0) ? T(0) : beta * z[row]) + alpha * s;
  } else {
    const int r = n_active + (blk - slot_blocks) * 256 + threadIdx.x;
    if(r >= n) return;
    y[r] = (beta == T(0) ? T(0) : beta * z[r]) + alpha * x[r];
  }
}

// x_out = x + omega D^{-1} (b - A x), the non-active rows with A = I
template <int U, class T>
__global__ __launch_bounds__(256) void k_jacobi_sell(int n, int n_active, SellView<T> m, int slot_blocks, int mode,
                                                     const T* __restrict__ diag, const T* __restrict__ b,
                                                     const T* __restrict__ x, T* __restrict__ x_out, T omega) {
  const int blk = sell_block(blockIdx.x, slot_blocks, mode);
  int row = -1;
  T s = 0;
  if(blk < slot_blocks) {
    const int gt = blk * 256 + threadIdx.x;
    if(gt < m.n_slots) {
      row = m.perm[gt];
      if(row >= 0) s = sell_row_sum<U, T>(m, gt, m.len[gt], x);
    }
    if(mode & 2) {
      __shared__ T sh[256];
      const int w0 = blk * 256;
      if(row >= 0) sh[row - w0] = s;
      __syncthreads();
      row = w0 + threadIdx.x;
      if(row >= n_active) return;
      s = sh[threadIdx.x];
    } else if(row < 0) {
      return;
    }
  } else {
    row = n_active + (blk - slot_blocks) * 256 + threadIdx.x;
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
int sell_slot_blocks(const SellView<T>& m) {
  return (m.n_slots + 255) / 256;
}

template <class T>
int sell_blocks(int n, int n_active, const SellView<T>& m) {
  return sell_slot_blocks(m) + (n - n_active + 255) / 256;
}

}  // namespace

static int g_sell_unroll = 16;
static int g_sell_mode = 3;
int sell_mode() { return g_sell_mode; }
void set_sell_mode(int mode) {
  PMCD_CHECK(mode >= 0 && mode <= 3, "sliced ELL mode must be 0..3");
  g_sell_mode = mode;
}
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

// row-ordered output needs each block's 256 slots to be one whole window of
// the layout (the window it was built with, not the current setting)
template <class T>
static int sell_launch_mode(const SellView<T>& m) {
  return m.window == 256 ? g_sell_mode : (g_sell_mode & 1);
}

template <class T>
void launch_spmv_sell(int n, int n_active, const SellView<T>& m, const T* x, const T* z, T* y, T alpha, T beta,
                      hipStream_t s) {
  if(n == 0) return;
  const int blocks = sell_blocks(n, n_active, m), sb = sell_slot_blocks(m), mode = sell_launch_mode(m);
  switch(g_sell_unroll) {
    case 4: k_spmv_sell<4, T><<<blocks, 256, 0, s>>>(n, n_active, m, sb, mode, x, z, y, alpha, beta); break;
    case 16: k_spmv_sell<16, T><<<blocks, 256, 0, s>>>(n, n_active, m, sb, mode, x, z, y, alpha, beta); break;
    default: k_spmv_sell<8, T><<<blocks, 256, 0, s>>>(n, n_active, m, sb, mode, x, z, y, alpha, beta); break;
  }
  PMCD_HIP(hipGetLastError());
}

template <class T>
void launch_jacobi_sell(int n, int n_active, const SellView<T>& m, const T* diag, const T* b, const T* x, T* x_out,
                        T omega, hipStream_t s) {
  if(n == 0) return;
  const int blocks = sell_blocks(n, n_active, m), sb = sell_slot_blocks(m), mode = sell_launch_mode(m);
  switch(g_sell_unroll) {
    case 4: k_jacobi_sell<4, T><<<blocks, 256, 0, s>>>(n, n_active, m, sb, mode, diag, b, x, x_out, omega); break;
    case 16: k_jacobi_sell<16, T><<<blocks, 256, 0, s>>>(n, n_active, m, sb, mode, diag, b, x, x_out, omega); break;
    default: k_jacobi_sell<8, T><<<blocks, 256, 0, s>>>(n, n_active, m, sb, mode, diag, b, x, x_out, omega); break;
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
