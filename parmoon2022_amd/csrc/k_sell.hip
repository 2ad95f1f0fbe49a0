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

// matrix entries are read once per sweep: with NT the loads are
// non-temporal, so that they do not displace the gathered x entries, which
// neighbouring rows read again, from the caches
template <bool NT, class V>
__device__ inline V stream_load(const V* p) {
  if constexpr(NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// sum over the row in `slot`: U entries of the lane loaded per step, then
// added in order. CI: the column array's type (global int columns, or 16-bit
// positions in the block's footprint with xs the footprint's x entries)
template <int U, bool NT, class T, class CI>
__device__ inline T sell_row_sum(const SellView<T>& m, const CI* __restrict__ cols, int slot, int len,
                                 const T* __restrict__ x) {
  const long long base = m.off[slot / kSellRows] + (slot % kSellRows);
  T s = 0;
  for(int k0 = 0; k0 < len; k0 += U) {
    int c[U];
    T v[U], xv[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int k = k0 + u;
      c[u] = k < len ? (int)stream_load<NT>(&cols[base + (long long)k * kSellRows]) : -1;
      v[u] = k < len ? stream_load<NT>(&m.val[base + (long long)k * kSellRows]) : T(0);
    }
#pragma unroll
    for(int u = 0; u < U; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : T(0);
#pragma unroll
    for(int u = 0; u < U; ++u)
      if(k0 + u < len) s += v[u] * xv[u];
  }
  return s;
}

// The same sum with the matrix loads of the next U entries issued before the
// products of the current ones: the HBM latency of the matrix stream overlaps
// the gathers of x and the arithmetic (same products, same order).
template <int U, bool NT, class T, class CI>
__device__ inline T sell_row_sum_pipe(const SellView<T>& m, const CI* __restrict__ cols, int slot, int len,
                                      const T* __restrict__ x) {
  const long long base = m.off[slot / kSellRows] + (slot % kSellRows);
  T s = 0;
  int c[U];
  T v[U];
#pragma unroll
  for(int u = 0; u < U; ++u) {
    c[u] = u < len ? (int)stream_load<NT>(&cols[base + (long long)u * kSellRows]) : -1;
    v[u] = u < len ? stream_load<NT>(&m.val[base + (long long)u * kSellRows]) : T(0);
  }
  for(int k0 = 0; k0 < len; k0 += U) {
    T xv[U];
#pragma unroll
    for(int u = 0; u < U; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : T(0);
    int cn[U];
    T vn[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int k = k0 + U + u;
      cn[u] = k < len ? (int)stream_load<NT>(&cols[base + (long long)k * kSellRows]) : -1;
      vn[u] = k < len ? stream_load<NT>(&m.val[base + (long long)k * kSellRows]) : T(0);
    }
#pragma unroll
    for(int u = 0; u < U; ++u)
      if(k0 + u < len) s += v[u] * xv[u];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      c[u] = cn[u];
      v[u] = vn[u];
    }
  }
  return s;
}

// PIPE: the pipelined sum (a kernel of its own, since it needs more registers)
template <int U, bool PIPE, class T, class CI>
__device__ inline T sell_row(const SellView<T>& m, const CI* cols, int slot, int len, const T* __restrict__ x,
                             int mode) {
  if constexpr(PIPE)
    return (mode & 4) ? sell_row_sum_pipe<U, true, T, CI>(m, cols, slot, len, x)
                      : sell_row_sum_pipe<U, false, T, CI>(m, cols, slot, len, x);
  else
    return (mode & 4) ? sell_row_sum<U, true, T, CI>(m, cols, slot, len, x)
                      : sell_row_sum<U, false, T, CI>(m, cols, slot, len, x);
}

// Row sum of the slot's row. FP (mode bit 4): the block first gathers the x
// entries of its footprint (the distinct columns of its 256 rows, about 16
// entries per distinct column on Q2 hexahedra) into LDS with one load each,
// then the rows read x there by their 16-bit footprint positions, so that a
// gather instruction no longer costs the vector L1 one access per lane. The
// products and their order are the same, so are the bits. Every thread of
// the block must call it (it synchronizes the block).
template <int U, bool PIPE, bool FP, class T>
__device__ inline T sell_block_row(const SellView<T>& m, int blk, int gt, int row, const T* __restrict__ x,
                                   int mode) {
  if constexpr(FP) {
    extern __shared__ __align__(16) unsigned char sell_dyn[];
    T* xs = reinterpret_cast<T*>(sell_dyn);
    const int f0 = m.fp_off[blk], nf = m.fp_off[blk + 1] - f0;
    for(int i = threadIdx.x; i < nf; i += 256) xs[i] = x[m.fp[f0 + i]];
    __syncthreads();
    return row >= 0 ? sell_row<U, PIPE, T, unsigned short>(m, m.lcol, gt, m.len[gt], xs, mode) : T(0);
  } else {
    return row >= 0 ? sell_row<U, PIPE, T, int>(m, m.col, gt, m.len[gt], x, mode) : T(0);
  }
}

// Launch layout: blocks of 256 threads (4 slices); the first slot_blocks
// blocks cover the slots, the rest the non-active tail rows in order.
//
// mode bit 0 (XCD-contiguous slices): the hardware hands blocks to the 8 XCDs
// round-robin; the slot blocks are renumbered so that XCD k works through one
// contiguous range of slices, keeping the x entries its rows share in its own
// L2 instead of every XCD fetching most of x.
//
// mode bit 2 (non-temporal matrix loads): see stream_load. mode bit 3
// (pipelined matrix loads): see sell_row_sum_pipe.
//
// mode bit 1 (row-ordered output): a block's 256 slots are the rows of one
// 256-row window, sorted by length; the row sums go to LDS at the row's
// position in the window, and the row-wise reads (b, diag, z, the row's own
// x) and the output are done by thread t for row w0 + t, coalesced, instead
// of at the sorted positions (partial cache lines, written by several
// waves). The arithmetic per row is the same, so the bits are too.
//
// mode bit 4 (x staged in LDS, with bit 1 and column footprints): see
// sell_block_row; the launch then takes the footprint's x entries as
// dynamic LDS.
__device__ inline int sell_block(int bid, int slot_blocks, int mode) {
  if(!(mode & 1) || bid >= slot_blocks) return bid;
  const int per = slot_blocks / 8;
  if(bid >= per * 8) return bid;
  return (bid % 8) * per + bid / 8;
}

// y[r] = beta z[r] + alpha (A x)[r] on active rows, beta z[r] + alpha x[r] on the tail
template <int U, bool PIPE, bool FP, class T>
__global__ __launch_bounds__(256) void k_spmv_sell(int n, int n_active, SellView<T> m, int slot_blocks, int mode,
                                                   const T* __restrict__ x, const T* __restrict__ z,
                                                   T* __restrict__ y, T alpha, T beta) {
  const int blk = sell_block(blockIdx.x, slot_blocks, mode);
  if(blk < slot_blocks) {
    const int gt = blk * 256 + threadIdx.x;
    const int row = gt < m.n_slots ? m.perm[gt] : -1;
    T s = sell_block_row<U, PIPE, FP, T>(m, blk, gt, row, x, mode);
    if(mode & 2) {
      __shared__ T sh[256];
      const int w0 = blk * 256;
      if(row >= 0) sh[row - w0] = s;
      __syncthreads();
      const int r = w0 + threadIdx.x;
      if(r >= n_active) return;
      y[r] = (beta == T(0) ? T(0) : beta * z[r]) + alpha * sh[threadIdx.x];
      return;
    }
    if(row < 0) return;
    y[row] = (beta == T(0) ? T(0) : beta * z[row]) + alpha * s;
  } else {

    const int r = n_active + (blk - slot_blocks) * 256 + threadIdx.x;
    if(r >= n) return;
    y[r] = (beta == T(0) ? T(0) : beta * z[r]) + alpha * x[r];
  }
}

// x_out = x + omega D^{-1} (b - A x), the non-active rows with A = I
template <int U, bool PIPE, bool FP, class T>
__global__ __launch_bounds__(256) void k_jacobi_sell(int n, int n_active, SellView<T> m, int slot_blocks, int mode,
                                                     const T* __restrict__ diag, const T* __restrict__ b,
                                                     const T* __restrict__ x, T* __restrict__ x_out, T omega) {
  const int blk = sell_block(blockIdx.x, slot_blocks, mode);
  int row = -1;
  T s = 0;
  if(blk < slot_blocks) {
    const int gt = blk * 256 + threadIdx.x;
    row = gt < m.n_slots ? m.perm[gt] : -1;
    s = sell_block_row<U, PIPE, FP, T>(m, blk, gt, row, x, mode);
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
static int g_sell_mode = 23;
int sell_mode() { return g_sell_mode; }
void set_sell_mode(int mode) {
  PMCD_CHECK(mode >= 0 && mode <= 31, "sliced ELL mode must be 0..31");
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
// blocks of the sliced ELL kernels resident per CU at most (0: no cap): the
// cap is set through unused dynamic LDS (160 KiB per CU shared by the
// resident blocks; one 256-thread block is one wave on each of the 4 SIMDs)
static int g_sell_blocks_per_cu = 0;
int sell_blocks_per_cu() { return g_sell_blocks_per_cu; }
void set_sell_blocks_per_cu(int b) {
  PMCD_CHECK(b == 0 || (b >= 3 && b <= 16), "sliced ELL blocks per CU must be 0 (no cap) or 3..16");
  g_sell_blocks_per_cu = b;
}
static size_t sell_dynamic_lds() {
  if(!g_sell_blocks_per_cu) return 0;
  const size_t per_block = (160u * 1024u) / (size_t)g_sell_blocks_per_cu + 1;
  const size_t fixed = 256 * sizeof(double);  // the kernels' static row buffer
  return per_block > fixed ? per_block - fixed : 0;
}
int sell_unroll() { return g_sell_unroll; }
void set_sell_unroll(int u) {
  PMCD_CHECK(u == 4 || u == 8 || u == 16, "sliced ELL unroll must be 4, 8 or 16");
  g_sell_unroll = u;
}

// row-ordered output needs each block's 256 slots to be one whole window of
// the layout (the window it was built with, not the current setting); x in
// LDS (bit 4) needs that too, and the footprints
int sell_mode_for(int window, int fp_max) {
  int mode = window == 256 ? g_sell_mode : (g_sell_mode & 13);
  if(!(mode & 2) || fp_max <= 0 || fp_max > kSellMaxFootprint) mode &= 15;
  return mode;
}

template <class T>
static int sell_launch_mode(const SellView<T>& m) {
  return sell_mode_for(m.window, m.lcol ? m.fp_max : 0);
}

// dynamic LDS of a launch: the footprint's x entries under bit 4, at least
// the occupancy cap's share
template <class T>
static size_t sell_launch_lds(const SellView<T>& m, int mode) {
  const size_t cap = sell_dynamic_lds();
  const size_t fp = (mode & 16) ? (size_t)m.fp_max * sizeof(T) : 0;
  return fp > cap ? fp : cap;
}

template <class T>
void launch_spmv_sell(int n, int n_active, const SellView<T>& m, const T* x, const T* z, T* y, T alpha, T beta,
                      hipStream_t s) {
  if(n == 0) return;
  const int blocks = sell_blocks(n, n_active, m), sb = sell_slot_blocks(m), mode = sell_launch_mode(m);
  const size_t lds = sell_launch_lds(m, mode);
#define PMCD_SELL_SPMV(U, P, F) \
  k_spmv_sell<U, P, F, T><<<blocks, 256, lds, s>>>(n, n_active, m, sb, mode, x, z, y, alpha, beta)
#define PMCD_SELL_SPMV_FP(U, P) \
  if(mode & 16) PMCD_SELL_SPMV(U, P, true); else PMCD_SELL_SPMV(U, P, false)
  const bool pipe = (mode & 8) != 0;
  switch(g_sell_unroll) {
    case 4: if(pipe) { PMCD_SELL_SPMV_FP(4, true); } else { PMCD_SELL_SPMV_FP(4, false); } break;
    case 16: if(pipe) { PMCD_SELL_SPMV_FP(16, true); } else { PMCD_SELL_SPMV_FP(16, false); } break;
    default: if(pipe) { PMCD_SELL_SPMV_FP(8, true); } else { PMCD_SELL_SPMV_FP(8, false); } break;
  }
#undef PMCD_SELL_SPMV_FP
#undef PMCD_SELL_SPMV
  PMCD_HIP(hipGetLastError());
}

template <class T>
void launch_jacobi_sell(int n, int n_active, const SellView<T>& m, const T* diag, const T* b, const T* x, T* x_out,
                        T omega, hipStream_t s) {
  if(n == 0) return;
  const int blocks = sell_blocks(n, n_active, m), sb = sell_slot_blocks(m), mode = sell_launch_mode(m);
  const size_t lds = sell_launch_lds(m, mode);
#define PMCD_SELL_JAC(U, P, F) \
  k_jacobi_sell<U, P, F, T><<<blocks, 256, lds, s>>>(n, n_active, m, sb, mode, diag, b, x, x_out, omega)
#define PMCD_SELL_JAC_FP(U, P) \
  if(mode & 16) PMCD_SELL_JAC(U, P, true); else PMCD_SELL_JAC(U, P, false)
  const bool pipe = (mode & 8) != 0;
  switch(g_sell_unroll) {
    case 4: if(pipe) { PMCD_SELL_JAC_FP(4, true); } else { PMCD_SELL_JAC_FP(4, false); } break;
    case 16: if(pipe) { PMCD_SELL_JAC_FP(16, true); } else { PMCD_SELL_JAC_FP(16, false); } break;
    default: if(pipe) { PMCD_SELL_JAC_FP(8, true); } else { PMCD_SELL_JAC_FP(8, false); } break;
  }
#undef PMCD_SELL_JAC_FP
#undef PMCD_SELL_JAC
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
