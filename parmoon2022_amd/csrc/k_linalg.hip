// Bandwidth-bound kernels of the solve: CSR SpMV with the reference's
// non-active-row rule, the damped Jacobi sweep, transfer products, dense
// coarse-grid products, deterministic reductions and vector updates.
//
// Reference: BlockFEMatrix::apply / apply_scaled_add
// (src/Matrix/BlockFEMatrix.C:969-1095) -> FEMatrix::multiplyActive
// (src/Matrix/FEMatrix.C:209-223) + BlockVector::addScaledNonActive;
// JacobiSmoother::smooth (src/Multigrid/JacobiSmoother.C:22-30) with
// Iteration_jacobi::apply (src/Solver/Iteration_jacobi.C:168-182);
// Ddot/Dnorm (src/FE/Blas1.C:21-98).
//
// Reductions use a fixed two-pass tree (fixed grid, fixed block tree), so
// repeated runs give identical bits; the order differs from the reference's
// serial sums, which the parity tests bound by a rounding tolerance.
#include <hip/hip_cooperative_groups.h>

#include "kernels.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace pmcd {

namespace {

constexpr int kRedBlocks = kMaxReduceBlocks;
// tuning key 14: the unrolled CSR kernels renumber their blocks to
// XCD-contiguous row ranges (same bits either way)
static int g_csr_xcd = 0;

// W lanes per row; lanes stride the row, then a butterfly over the W lanes.
template <int W, class T>
__global__ __launch_bounds__(256) void k_spmv(int n, int n_active, const int* __restrict__ rp,
                                              const int* __restrict__ col,
                                              const T* __restrict__ val, const T* __restrict__ x,
                                              const T* __restrict__ z, T* __restrict__ y,
                                              T alpha, T beta) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n) return;
  if(row < n_active) {
    T s = 0;
    const int e = rp[row + 1];
    for(int k = rp[row] + l; k < e; k += W) s += val[k] * x[col[k]];
#pragma unroll
    for(int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
    if(l == 0) y[row] = (beta == T(0) ? T(0) : beta * z[row]) + alpha * s;
  } else if(l == 0) {
    y[row] = (beta == T(0) ? T(0) : beta * z[row]) + alpha * x[row];
  }
}

template <int W, class T>
__global__ __launch_bounds__(256) void k_jacobi(int n, int n_active, const int* __restrict__ rp,
                                                const int* __restrict__ col,
                                                const T* __restrict__ val,
                                                const T* __restrict__ diag,
                                                const T* __restrict__ b, const T* __restrict__ x,
                                                T* __restrict__ x_out, T omega) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n) return;
  T s;
  if(row < n_active) {
    s = 0;
    const int e = rp[row + 1];
    for(int k = rp[row] + l; k < e; k += W) s += val[k] * x[col[k]];
#pragma unroll
    for(int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
  } else {
    s = x[row];
  }
  if(l == 0) {
    T d = b[row] + T(-1) * s;
    d = d / diag[row];
    x_out[row] = x[row] + omega * d;
  }
}

// Streaming variant: blocks are remapped so that each of the 8 XCDs (which
// receive blocks round-robin) works on one contiguous range of rows, keeping
// the x entries those rows share in the XCD's own L2; the matrix streams are
// read with non-temporal loads so they do not evict x; each lane keeps two
// independent partial sums to keep two loads in flight.
__device__ inline int xcd_block(int bid, int nb) {
  const int per = nb / 8;
  if(bid >= per * 8) return bid;
  return (bid % 8) * per + bid / 8;
}

template <int W, class T>
__device__ inline T row_sum_stream(const int* __restrict__ rp, const int* __restrict__ col,
                                   const T* __restrict__ val, const T* __restrict__ x, int row,
                                   int l) {
  const int e = rp[row + 1];
  int k = rp[row] + l;
  T s0 = 0, s1 = 0;
  for(; k + W < e; k += 2 * W) {
    const int c0 = __builtin_nontemporal_load(&col[k]);
    const int c1 = __builtin_nontemporal_load(&col[k + W]);
    const T v0 = __builtin_nontemporal_load(&val[k]);
    const T v1 = __builtin_nontemporal_load(&val[k + W]);
    s0 += v0 * x[c0];
    s1 += v1 * x[c1];
  }
  if(k < e) s0 += __builtin_nontemporal_load(&val[k]) * x[__builtin_nontemporal_load(&col[k])];
  T s = s0 + s1;
#pragma unroll
  for(int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
  return s;
}

template <int W, class T>
__global__ __launch_bounds__(256) void k_spmv_stream(int n, int n_active, const int* __restrict__ rp,
                                                     const int* __restrict__ col,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ x,
                                                     const T* __restrict__ z, T* __restrict__ y,
                                                     T alpha, T beta) {
  const int gt = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n) return;
  if(row < n_active) {
    const T s = row_sum_stream<W, T>(rp, col, val, x, row, l);
    if(l == 0) y[row] = (beta == T(0) ? T(0) : beta * z[row]) + alpha * s;
  } else if(l == 0) {
    y[row] = (beta == T(0) ? T(0) : beta * z[row]) + alpha * x[row];
  }
}

template <int W, class T>
__global__ __launch_bounds__(256) void k_jacobi_stream(int n, int n_active,
                                                       const int* __restrict__ rp,
                                                       const int* __restrict__ col,
                                                       const T* __restrict__ val,
                                                       const T* __restrict__ diag,
                                                       const T* __restrict__ b,
                                                       const T* __restrict__ x,
                                                       T* __restrict__ x_out, T omega) {
  const int gt = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n) return;
  T s;
  if(row < n_active) s = row_sum_stream<W, T>(rp, col, val, x, row, l);
  else s = x[row];
  if(l == 0) {
    T d = b[row] + T(-1) * s;
    d = d / diag[row];
    x_out[row] = x[row] + omega * d;
  }
}

// Unrolled variant: every lane owns at most U entries of its row (U W >= the
// longest row); it issues all its column and value loads, then all the
// dependent x gathers, then the products, maximising the loads in flight.
template <int W, int U, class T>
__device__ inline T row_sum_unrolled(const int* __restrict__ rp, const int* __restrict__ col,
                                     const T* __restrict__ val, const T* __restrict__ x, int row,
                                     int l) {
  const int b = rp[row], e = rp[row + 1];
  int c[U];
  T v[U], xv[U];
#pragma unroll
  for(int u = 0; u < U; ++u) {
    const int k = b + l + u * W;
    c[u] = k < e ? col[k] : -1;
  }
#pragma unroll
  for(int u = 0; u < U; ++u) {
    const int k = b + l + u * W;
    v[u] = k < e ? val[k] : T(0);
  }
#pragma unroll
  for(int u = 0; u < U; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : T(0);
  T s = 0;
#pragma unroll
  for(int u = 0; u < U; ++u) s += v[u] * xv[u];
#pragma unroll
  for(int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
  return s;
}

// Same with non-temporal matrix loads (the CSR arrays stream through once
// per pass; x keeps the cache)
template <int W, int U, class T>
__device__ inline T row_sum_unrolled_nt(const int* __restrict__ rp, const int* __restrict__ col,
                                        const T* __restrict__ val, const T* __restrict__ x, int row,
                                        int l) {
  const int b = rp[row], e = rp[row + 1];
  int c[U];
  T v[U], xv[U];
#pragma unroll
  for(int u = 0; u < U; ++u) {
    const int k = b + l + u * W;
    c[u] = k < e ? __builtin_nontemporal_load(&col[k]) : -1;
  }
#pragma unroll
  for(int u = 0; u < U; ++u) {
    const int k = b + l + u * W;
    v[u] = k < e ? __builtin_nontemporal_load(&val[k]) : T(0);
  }
#pragma unroll
  for(int u = 0; u < U; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : T(0);
  T s = 0;
#pragma unroll
  for(int u = 0; u < U; ++u) s += v[u] * xv[u];
#pragma unroll
  for(int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
  return s;
}

template <int W, int U, class T, bool NT>
__device__ inline T row_sum_u(const int* __restrict__ rp, const int* __restrict__ col,
                              const T* __restrict__ val, const T* __restrict__ x, int row, int l) {
  if constexpr(NT) return row_sum_unrolled_nt<W, U, T>(rp, col, val, x, row, l);
  else return row_sum_unrolled<W, U, T>(rp, col, val, x, row, l);
}

// xcd != 0: blocks renumbered to XCD-contiguous row ranges (xcd_block)
template <int W, int U, class T, bool NT = false>
__global__ __launch_bounds__(256) void k_spmv_unrolled(int n, int n_active, const int* __restrict__ rp,
                                                       const int* __restrict__ col,
                                                       const T* __restrict__ val,
                                                       const T* __restrict__ x,
                                                       const T* __restrict__ z, T* __restrict__ y,
                                                       T alpha, T beta, int xcd) {
  const int gt = (xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x) * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n) return;
  if(row < n_active) {
    const T s = row_sum_u<W, U, T, NT>(rp, col, val, x, row, l);
    if(l == 0) y[row] = (beta == T(0) ? T(0) : beta * z[row]) + alpha * s;
  } else if(l == 0) {
    y[row] = (beta == T(0) ? T(0) : beta * z[row]) + alpha * x[row];
  }
}

template <int W, int U, class T, bool NT = false>
__global__ __launch_bounds__(256) void k_jacobi_unrolled(int n, int n_active,
                                                         const int* __restrict__ rp,
                                                         const int* __restrict__ col,
                                                         const T* __restrict__ val,
                                                         const T* __restrict__ diag,
                                                         const T* __restrict__ b,
                                                         const T* __restrict__ x,
                                                         T* __restrict__ x_out, T omega, int xcd) {
  const int gt = (xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x) * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n) return;
  // the row's own b, diag and x are loaded before the row sum so that their
  // latency overlaps the matrix loads
  T bb = 0, dd = 1, xr = 0;
  if(l == 0) {
    bb = b[row];
    dd = diag[row];
    xr = x[row];
  }
  T s;
  if(row < n_active) s = row_sum_u<W, U, T, NT>(rp, col, val, x, row, l);
  else s = xr;
  if(l == 0) {
    T d = bb + T(-1) * s;
    d = d / dd;
    x_out[row] = xr + omega * d;
  }
}

// First Jacobi sweep from x = 0 with b = 0 on the non-active rows: A x = 0,
// so the sweep reduces to x_out = 0 + omega * ((b + (-1)*0) / diag) on active
// rows (the same floating-point result as the general sweep) and x_out = b on
// non-active rows.
template <class T>
__global__ void k_jacobi_from_zero(int n, int n_active, const T* __restrict__ diag,
                                   const T* __restrict__ b, T* __restrict__ x_out, T omega) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if(i < n_active) {
      T d = b[i] / diag[i];
      x_out[i] = T(0) + omega * d;
    } else {
      x_out[i] = b[i];
    }
  }
}

// Grid transfer with the reference's non-active rules fused in (Multigrid.C:
// 310-410): restriction writes (R d)[r] on active coarse rows and 0 on the
// Dirichlet tail; prolongation adds a (P e)[r] to the fine iterate on active
// rows only. U = 0 walks the row in a loop; U > 0 is the unrolled row sum
// (U W >= the longest row). Same per-lane order and butterfly as k_spmv.
template <int W, int U, class T>
__device__ inline T transfer_row(const int* __restrict__ rp, const int* __restrict__ col,
                                 const T* __restrict__ val, const T* __restrict__ x, int row, int l) {
  if constexpr(U > 0) {
    return row_sum_unrolled<W, U, T>(rp, col, val, x, row, l);
  } else {
    T s = 0;
    const int e = rp[row + 1];
    for(int k = rp[row] + l; k < e; k += W) s += val[k] * x[col[k]];
#pragma unroll
    for(int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
    return s;
  }
}

template <int W, int U, class T>
__global__ __launch_bounds__(256) void k_restrict(int n, int n_active, const int* __restrict__ rp,
                                                  const int* __restrict__ col,
                                                  const T* __restrict__ val,
                                                  const T* __restrict__ x, T* __restrict__ y) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n) return;
  if(row < n_active) {
    const T s = transfer_row<W, U, T>(rp, col, val, x, row, l);
    if(l == 0) y[row] = s;
  } else if(l == 0) {
    y[row] = T(0);
  }
}

template <int W, int U, class T>
__global__ __launch_bounds__(256) void k_prolong_add(int n_active, const int* __restrict__ rp,
                                                     const int* __restrict__ col,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     T a) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int row = gt / W, l = gt % W;
  if(row >= n_active) return;
  const T s = transfer_row<W, U, T>(rp, col, val, x, row, l);
  if(l == 0) y[row] += a * s;
}

// Rows of one class of a successive over-relaxation sweep
// (TMatrix::sor_sweep, src/Matrix/Matrix.C:1308-1420): one thread per row, the
// row's sum taken in column order starting from b, then
// x[row] += omega * s / diag[row], in place. Rows of one class are never
// adjacent, so each reads only x entries that no other thread of the launch
// writes.
// A row whose owner (row_owner[r]) is flagged inactive is skipped (batched
// solves: converged instances); row_owner null means every row is active.
template <class T>
__global__ __launch_bounds__(256) void k_sor_rows(int count, const int* __restrict__ rows,
                                                  const int* __restrict__ rp,
                                                  const int* __restrict__ col,
                                                  const T* __restrict__ val,
                                                  const T* __restrict__ diag,
                                                  const T* __restrict__ b, T* x, T omega,
                                                  const int* __restrict__ row_owner,
                                                  const uint8_t* __restrict__ owner_active) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if(i >= count) return;
  const int r = rows[i];
  if(row_owner && !owner_active[row_owner[r]]) return;
  T s = b[r];
  const int e = rp[r + 1];
  for(int k = rp[r]; k < e; ++k) s -= val[k] * x[col[k]];
  x[r] = x[r] + omega * s / diag[r];
}

// A whole exact (level-scheduled) sweep in one workgroup: the classes in
// order (forward), in reverse order (backward) or both, a barrier between
// classes. Each row is updated exactly as k_sor_rows updates it, so the
// result is the same bit for bit; on small levels this replaces one launch
// per class by one launch per sweep.
template <class T>
__global__ __launch_bounds__(1024) void k_sor_sweep_block(int K, const int* __restrict__ class_ptr,
                                                          const int* __restrict__ rows,
                                                          const int* __restrict__ rp,
                                                          const int* __restrict__ col,
                                                          const T* __restrict__ val,
                                                          const T* __restrict__ diag,
                                                          const T* __restrict__ b, T* x, T omega, int flag,
                                                          const int* __restrict__ row_owner,
                                                          const uint8_t* __restrict__ owner_active) {
  const int n_pass = flag == 2 ? 2 * K : K;
  for(int p = 0; p < n_pass; ++p) {
    const int k = flag == 1 ? K - 1 - p : (p < K ? p : 2 * K - 1 - p);
    const int e = class_ptr[k + 1];
    for(int i = class_ptr[k] + (int)threadIdx.x; i < e; i += 1024) {
      const int r = rows[i];
      if(row_owner && !owner_active[row_owner[r]]) continue;
      T s = b[r];
      const int re = rp[r + 1];
      for(int q = rp[r]; q < re; ++q) s -= val[q] * x[col[q]];
      x[r] = x[r] + omega * s / diag[r];
    }
    __syncthreads();
  }
}

// Multicolour variant: W lanes share a row (unrolled loads, butterfly sum),
// then x[row] += omega * (b - sum) / diag. The row sum is rounded in a
// different order than the serial sweep's b - a_1 x_1 - a_2 x_2 ...
template <int W, int U, class T>
__global__ __launch_bounds__(256) void k_sor_rows_lanes(int count, const int* __restrict__ rows,
                                                        const int* __restrict__ rp,
                                                        const int* __restrict__ col,
                                                        const T* __restrict__ val,
                                                        const T* __restrict__ diag,
                                                        const T* __restrict__ b, T* x, T omega,
                                                        const int* __restrict__ row_owner,
                                                        const uint8_t* __restrict__ owner_active) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int i = gt / W, l = gt % W;
  if(i >= count) return;
  const int r = rows[i];
  if(row_owner && !owner_active[row_owner[r]]) return;
  const T s = transfer_row<W, U, T>(rp, col, val, x, r, l);
  if(l == 0) x[r] = x[r] + omega * (b[r] - s) / diag[r];
}

// Multicolour class over the class-ordered copy of the matrix: permuted row i
// (original row rows[i]) has its entries, in their original order, at
// sw_rp[i] .. sw_rp[i+1]; a class is a contiguous range of permuted rows, so
// its matrix data streams like a plain SpMV.
template <int W, int U, class T>
__global__ __launch_bounds__(256) void k_sor_rows_perm(int count, int i0, const int* __restrict__ rows,
                                                       const int* __restrict__ sw_rp,
                                                       const int* __restrict__ sw_col,
                                                       const T* __restrict__ sw_val,
                                                       const T* __restrict__ diag,
                                                       const T* __restrict__ b, T* x, T omega,
                                                       const int* __restrict__ row_owner,
                                                       const uint8_t* __restrict__ owner_active) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int li = gt / W, l = gt % W;
  if(li >= count) return;
  const int i = i0 + li;
  const int r = rows[i];
  if(row_owner && !owner_active[row_owner[r]]) return;
  const T s = transfer_row<W, U, T>(sw_rp, sw_col, sw_val, x, i, l);
  if(l == 0) x[r] = x[r] + omega * (b[r] - s) / diag[r];
}

// A whole multicolour sweep in one workgroup: the class kernel above run
// class after class with a barrier between classes (W lanes per row, the
// same row sums), so the result equals one launch per class bit for bit.
template <int W, int U, class T>
__global__ __launch_bounds__(1024) void k_sor_sweep_block_perm(int K, const int* __restrict__ class_ptr,
                                                               const int* __restrict__ rows,
                                                               const int* __restrict__ sw_rp,
                                                               const int* __restrict__ sw_col,
                                                               const T* __restrict__ sw_val,
                                                               const T* __restrict__ diag,
                                                               const T* __restrict__ b, T* x, T omega, int flag,
                                                               const int* __restrict__ row_owner,
                                                               const uint8_t* __restrict__ owner_active) {
  const int n_pass = flag == 2 ? 2 * K : K;
  const int l = (int)threadIdx.x % W;
  for(int p = 0; p < n_pass; ++p) {
    const int k = flag == 1 ? K - 1 - p : (p < K ? p : 2 * K - 1 - p);
    const int i0 = class_ptr[k], count = class_ptr[k + 1] - i0;
    for(int li = (int)threadIdx.x / W; li < count; li += 1024 / W) {
      const int i = i0 + li;
      const int r = rows[i];
      if(row_owner && !owner_active[row_owner[r]]) continue;
      const T s = transfer_row<W, U, T>(sw_rp, sw_col, sw_val, x, i, l);
      if(l == 0) x[r] = x[r] + omega * (b[r] - s) / diag[r];
    }
    __syncthreads();
  }
}

// A whole multicolour sweep in one cooperative launch over the whole grid:
// the class kernel above run class after class with a grid-wide barrier
// between classes (the runtime guarantees every block of a cooperative
// launch is resident; every block runs the same passes and reaches every
// barrier). Same row sums, so the same bits as one launch per class.
template <int W, int U, class T>
__global__ __launch_bounds__(256) void k_sor_sweep_coop(int K, const int* __restrict__ class_ptr,
                                                        const int* __restrict__ rows, const int* __restrict__ sw_rp,
                                                        const int* __restrict__ sw_col,
                                                        const T* __restrict__ sw_val, const T* __restrict__ diag,
                                                        const T* __restrict__ b, T* x, T omega, int flag,
                                                        const int* __restrict__ row_owner,
                                                        const uint8_t* __restrict__ owner_active) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int n_pass = flag == 2 ? 2 * K : K;
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int l = gt % W, step = (int)(gridDim.x * 256 / W);
  for(int p = 0; p < n_pass; ++p) {
    const int k = flag == 1 ? K - 1 - p : (p < K ? p : 2 * K - 1 - p);
    const int i0 = class_ptr[k], count = class_ptr[k + 1] - i0;
    for(int li = gt / W; li < count; li += step) {
      const int i = i0 + li;
      const int r = rows[i];
      if(row_owner && !owner_active[row_owner[r]]) continue;
      const T s = transfer_row<W, U, T>(sw_rp, sw_col, sw_val, x, i, l);
      if(l == 0) x[r] = x[r] + omega * (b[r] - s) / diag[r];
    }
    grid.sync();
  }
}

// val in class order: permuted row i <- original row rows[i]
template <class T>
__global__ __launch_bounds__(256) void k_permute_rows(int n, const int* __restrict__ rows,
                                                      const int* __restrict__ rp,
                                                      const int* __restrict__ sw_rp,
                                                      const T* __restrict__ val, T* __restrict__ sw_val) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int i = gt / 8, l = gt % 8;
  if(i >= n) return;
  const int r = rows[i];
  const int src = rp[r], dst = sw_rp[i], len = rp[r + 1] - src;
  for(int t = l; t < len; t += 8) sw_val[dst + t] = val[src + t];
}

__device__ inline double block_sum(double v) {
  __shared__ double s_w[4];
  for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if(l == 0) s_w[w] = v;
  __syncthreads();
  double r = 0;
  if(threadIdx.x == 0) r = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_dot_partial(int n, const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     double* __restrict__ partial) {
  double s = 0;
  const int stride = gridDim.x * 256;
  int i = blockIdx.x * 256 + threadIdx.x;
  // four elements per step, loads first, products added in element order
  for(; i + 3 * stride < n; i += 4 * stride) {
    double xx[4], yy[4];
#pragma unroll
    for(int u = 0; u < 4; ++u) {
      xx[u] = x[i + u * stride];
      yy[u] = y[i + u * stride];
    }
#pragma unroll
    for(int u = 0; u < 4; ++u) s += xx[u] * yy[u];
  }
  for(; i < n; i += stride) s += x[i] * y[i];
  s = block_sum(s);
  if(threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_final_sum(int m, const double* __restrict__ partial,
                                                   double* __restrict__ out) {
  double s = 0;
  for(int i = threadIdx.x; i < m; i += 256) s += partial[i];
  s = block_sum(s);
  if(threadIdx.x == 0) out[0] = s;
}

// per-cell error records: sum of fields 0..3, max of field 4
__global__ __launch_bounds__(256) void k_sum_max_partial(int n, int stride,
                                                         const double* __restrict__ v,
                                                         double* __restrict__ partial) {
  double s[4] = {0, 0, 0, 0};
  double mx = 0;
  for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    for(int k = 0; k < 4; ++k) s[k] += v[(size_t)i * stride + k];
    mx = fmax(mx, v[(size_t)i * stride + 4]);
  }
  for(int k = 0; k < 4; ++k) {
    double t = block_sum(s[k]);
    if(threadIdx.x == 0) partial[blockIdx.x * 5 + k] = t;
  }
  __shared__ double s_m[256];
  s_m[threadIdx.x] = mx;
  __syncthreads();
  if(threadIdx.x == 0) {
    double m = 0;
    for(int i = 0; i < 256; ++i) m = fmax(m, s_m[i]);
    partial[blockIdx.x * 5 + 4] = m;
  }
}

__global__ void k_sum_max_final(int m, const double* __restrict__ partial, double* __restrict__ out) {
  if(threadIdx.x == 0 && blockIdx.x == 0) {
    double s[4] = {0, 0, 0, 0}, mx = 0;
    for(int i = 0; i < m; ++i) {
      for(int k = 0; k < 4; ++k) s[k] += partial[i * 5 + k];
      mx = fmax(mx, partial[i * 5 + 4]);
    }
    for(int k = 0; k < 4; ++k) out[k] = s[k];
    out[4] = mx;
  }
}

__global__ void k_axpy(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

template <class T>
__global__ void k_fill(int n, T v, T* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = v;
}

// Plain streaming copy of 16-byte vectors: the HBM calibration kernel (known
// bytes read and written per dispatch) and the practical copy ceiling.
__global__ __launch_bounds__(256) void k_stream_copy(long n2, const double2* __restrict__ x,
                                                     double2* __restrict__ y) {
  for(long i = (long)blockIdx.x * 256 + threadIdx.x; i < n2; i += (long)gridDim.x * 256) y[i] = x[i];
}

typedef double f64x2 __attribute__((ext_vector_type(2)));

// four independent 16-byte loads in flight per thread, non-temporal stores
// (the fastest copy shape measured on MI355X: 8192 blocks)
__global__ __launch_bounds__(256) void k_stream_copy4(long n2, const double2* __restrict__ x,
                                                      double2* __restrict__ y) {
  const long stride = (long)gridDim.x * 256;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  const f64x2* xv = (const f64x2*)x;
  f64x2* yv = (f64x2*)y;
  for(; i + 3 * stride < n2; i += 4 * stride) {
    const f64x2 a = xv[i], b = xv[i + stride], c = xv[i + 2 * stride], d = xv[i + 3 * stride];
    __builtin_nontemporal_store(a, &yv[i]);
    __builtin_nontemporal_store(b, &yv[i + stride]);
    __builtin_nontemporal_store(c, &yv[i + 2 * stride]);
    __builtin_nontemporal_store(d, &yv[i + 3 * stride]);
  }
  for(; i < n2; i += stride) y[i] = x[i];
}

// read-only stream (sum of 16-byte vectors, four non-temporal loads in flight
// per thread): the HBM read ceiling for kernels that mostly read
__global__ __launch_bounds__(256) void k_stream_read(long n2, const double2* __restrict__ x,
                                                     double* __restrict__ out) {
  const long stride = (long)gridDim.x * 256;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  const f64x2* xv = (const f64x2*)x;
  double s0 = 0, s1 = 0;
  for(; i + 3 * stride < n2; i += 4 * stride) {
    const f64x2 a = __builtin_nontemporal_load(&xv[i]), b = __builtin_nontemporal_load(&xv[i + stride]);
    const f64x2 c = __builtin_nontemporal_load(&xv[i + 2 * stride]), d = __builtin_nontemporal_load(&xv[i + 3 * stride]);
    s0 += (a.x + b.x) + (c.x + d.x);
    s1 += (a.y + b.y) + (c.y + d.y);
  }
  for(; i < n2; i += stride) { s0 += x[i].x; s1 += x[i].y; }
  s0 += s1;
  for(int off = 32; off > 0; off >>= 1) s0 += __shfl_xor(s0, off, 64);
  if((threadIdx.x & 63) == 0) out[(blockIdx.x * 256 + threadIdx.x) / 64] = s0;
}

// y = sum_j c_j V_j, the terms added in order j = 0, 1, ...
__global__ void k_lincomb(int n, LinComb lc, double* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for(int j = 0; j < lc.k; ++j) s = s + lc.v[j][i] * lc.c[j];
    y[i] = s;
  }
}

__global__ void k_copy(int n, const double* __restrict__ x, double* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = x[i];
}

// r += (-h) v ; partial dot of the updated r with v_next
__global__ __launch_bounds__(256) void k_mgs_step(int n, const double* __restrict__ h,
                                                  double* __restrict__ r,
                                                  const double* __restrict__ v,
                                                  const double* __restrict__ vn,
                                                  double* __restrict__ partial) {
  const double mh = -h[0];
  double s = 0;
  for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double ri = r[i] + mh * v[i];
    r[i] = ri;
    if(vn) s += ri * vn[i];
  }
  if(vn) {
    s = block_sum(s);
    if(threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// Sum of the m partials of the previous reduction kernel, in the order of
// k_final_sum, computed by every block of the consuming kernel (the same bits
// in every block) and broadcast to its threads: the consumer replaces the
// one-block k_final_sum launch.
__device__ inline double block_reduce_partials(int m, const double* __restrict__ partial) {
  __shared__ double s_tot;
  double s = 0;
  for(int i = threadIdx.x; i < m; i += 256) s += partial[i];
  s = block_sum(s);
  if(threadIdx.x == 0) s_tot = s;
  __syncthreads();
  return s_tot;
}

// Modified Gram-Schmidt step with the previous dot product's reduction
// folded in: h = (r, v) from the m partials in `prev` (block 0 stores it in
// *h_out), r -= h v, then the partials of (r, wn) -- of (r, r) when wn is
// null -- into `partial`. Same arithmetic as k_final_sum + k_mgs_step.
__global__ __launch_bounds__(256) void k_mgs_fused(int n, int m, const double* __restrict__ prev,
                                                   double* __restrict__ h_out, double* __restrict__ r,
                                                   const double* __restrict__ v,
                                                   const double* __restrict__ wn,
                                                   double* __restrict__ partial) {
  const double h = block_reduce_partials(m, prev);
  if(blockIdx.x == 0 && threadIdx.x == 0) *h_out = h;
  const double mh = -h;
  double s = 0;
  const int stride = gridDim.x * 256;
  int i = blockIdx.x * 256 + threadIdx.x;
  // four of the thread's elements per step, all loads issued first (more
  // bytes in flight per CU); the products are added in the same order as
  // one element per step, so the partials are the same
  for(; i + 3 * stride < n; i += 4 * stride) {
    double rr[4], vv[4], ww[4];
#pragma unroll
    for(int u = 0; u < 4; ++u) {
      rr[u] = r[i + u * stride];
      vv[u] = v[i + u * stride];
      if(wn) ww[u] = wn[i + u * stride];
    }
#pragma unroll
    for(int u = 0; u < 4; ++u) {
      const double ri = rr[u] + mh * vv[u];
      r[i + u * stride] = ri;
      s += ri * (wn ? ww[u] : ri);
    }
  }
  for(; i < n; i += stride) {
    const double ri = r[i] + mh * v[i];
    r[i] = ri;
    s += ri * (wn ? wn[i] : ri);
  }
  s = block_sum(s);
  if(threadIdx.x == 0) partial[blockIdx.x] = s;
}

// v = r / |r| with |r|^2 from the m partials in `prev`; block 0 stores |r| in
// *h_out (the subdiagonal Hessenberg entry).
__global__ __launch_bounds__(256) void k_scale_norm(int n, int m, const double* __restrict__ prev,
                                                    double* __restrict__ h_out, const double* __restrict__ r,
                                                    double* __restrict__ v) {
  const double hn = sqrt(block_reduce_partials(m, prev));
  if(blockIdx.x == 0 && threadIdx.x == 0) *h_out = hn;
  const double a = 1.0 / hn;
  const int stride = gridDim.x * 256;
  int i = blockIdx.x * 256 + threadIdx.x;
  for(; i + 3 * stride < n; i += 4 * stride) {
    double rr[4];
#pragma unroll
    for(int u = 0; u < 4; ++u) rr[u] = r[i + u * stride];
#pragma unroll
    for(int u = 0; u < 4; ++u) v[i + u * stride] = rr[u] * a;
  }
  for(; i < n; i += stride) v[i] = r[i] * a;
}

// x += sum_j y_j z_j (FGMRES update, Utilities_gmres.C:90-108), the terms
// added in j order into a zero accumulator, then the sum into x: the
// arithmetic of a fill, k axpys and one more axpy, in one pass.
__global__ __launch_bounds__(256) void k_fgmres_update(int n, int k, const double* __restrict__ y,
                                                       const double* const* __restrict__ z,
                                                       double* __restrict__ x) {
  for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double acc = 0.0;
    for(int j = 0; j < k; ++j) acc += y[j] * z[j][i];
    x[i] += 1.0 * acc;
  }
}

// The reference's update of a caller's matrix and rhs by an assembly
// (Assemble3D.C:258-300,346-360): entries of the active rows (the first
// row_ptr[n_active] entries) get the assembled sums added, Dirichlet rows get
// 1 on the diagonal; rhs entries of active DOFs get the sums added, the
// others the Dirichlet values.
__global__ void k_add_assembled(int n, int n_active, long long nnz_active, const int* __restrict__ diag_pos,
                                const double* __restrict__ val, const double* __restrict__ f,
                                double* __restrict__ entries, double* __restrict__ rhs) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for(long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnz_active + n; k += stride) {
    if(k < nnz_active) {
      if(entries) entries[k] += val[k];
    } else {
      const int r = (int)(k - nnz_active);
      if(r < n_active) {
        if(rhs) rhs[r] += f[r];
      } else {
        if(entries) entries[diag_pos[r]] = 1.0;
        if(rhs) rhs[r] = f[r];
      }
    }
  }
}

__global__ void k_d2f(int n, const double* __restrict__ x, float* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}
__global__ void k_f2d(int n, const float* __restrict__ x, double* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}
// one wave per output row
template <class T>
__global__ __launch_bounds__(256) void k_gemv(int n, const T* __restrict__ M,
                                              const T* __restrict__ x, T* __restrict__ y) {
  const int row = blockIdx.x * 4 + threadIdx.x / 64, l = threadIdx.x % 64;
  if(row >= n) return;
  T s = 0;
  const T* m = M + (size_t)row * n;
  for(int k = l; k < n; k += 64) s += m[k] * x[k];
  for(int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if(l == 0) y[row] = s;
}

inline int grid_for(int n, int per_block) {
  long g = ((long)n + per_block - 1) / per_block;
  if(g > 65535L * 16) g = 65535L * 16;
  return (int)std::max<long>(g, 1);
}

// per-lane entry count of the unrolled variant for rows of at most max_row
// entries spread over W lanes: the smallest of 2, 4, 8, 16 that covers them,
// 0 when max_row is unknown (< 0) or too long
inline int unroll_for(int max_row, int W) {
  if(max_row < 0) return 0;
  for(int u = 2; u <= 16; u *= 2)
    if(u * W >= max_row) return u;
  return 0;
}

template <int W, class T, bool NT = false>
void spmv_unrolled(int u, int blocks, int n, int n_active, const int* rp, const int* col,
                   const T* val, const T* x, const T* z, T* y, T alpha, T beta, hipStream_t s) {
  switch(u) {
    case 2: k_spmv_unrolled<W, 2, T, NT><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta, g_csr_xcd); break;
    case 4: k_spmv_unrolled<W, 4, T, NT><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta, g_csr_xcd); break;
    case 8: k_spmv_unrolled<W, 8, T, NT><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta, g_csr_xcd); break;
    default: k_spmv_unrolled<W, 16, T, NT><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta, g_csr_xcd); break;
  }
}

template <int W, class T, bool NT = false>
void jacobi_unrolled(int u, int blocks, int n, int n_active, const int* rp, const int* col,
                     const T* val, const T* diag, const T* b, const T* x, T* x_out, T omega,
                     hipStream_t s) {
  switch(u) {
    case 2: k_jacobi_unrolled<W, 2, T, NT><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega, g_csr_xcd); break;
    case 4: k_jacobi_unrolled<W, 4, T, NT><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega, g_csr_xcd); break;
    case 8: k_jacobi_unrolled<W, 8, T, NT><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega, g_csr_xcd); break;
    default: k_jacobi_unrolled<W, 16, T, NT><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega, g_csr_xcd); break;
  }
}

template <class T>
void dispatch_spmv(int width, int variant, int max_row, int n, int n_active, const int* rp,
                   const int* col, const T* val, const T* x, const T* z, T* y, T alpha, T beta,
                   hipStream_t s) {
  const long threads = (long)n * width;
  const int blocks = (int)((threads + 255) / 256);
  const int u = width >= 8 ? unroll_for(max_row, width) : 0;
  // variant 4 (sliced ELL) needs a level's layout; raw CSR calls run unrolled
  if(variant == 4) variant = 2;
  if(variant == 2 && u) {
    switch(width) {
      case 8: spmv_unrolled<8, T>(u, blocks, n, n_active, rp, col, val, x, z, y, alpha, beta, s); break;
      case 16: spmv_unrolled<16, T>(u, blocks, n, n_active, rp, col, val, x, z, y, alpha, beta, s); break;
      default: spmv_unrolled<32, T>(u, blocks, n, n_active, rp, col, val, x, z, y, alpha, beta, s); break;
    }
    return;
  }
  if(variant == 3 && u) {
    switch(width) {
      case 8: spmv_unrolled<8, T, true>(u, blocks, n, n_active, rp, col, val, x, z, y, alpha, beta, s); break;
      case 16: spmv_unrolled<16, T, true>(u, blocks, n, n_active, rp, col, val, x, z, y, alpha, beta, s); break;
      default: spmv_unrolled<32, T, true>(u, blocks, n, n_active, rp, col, val, x, z, y, alpha, beta, s); break;
    }
    return;
  }
  if(variant == 1) {
    switch(width) {
      case 4: k_spmv_stream<4, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta); break;
      case 8: k_spmv_stream<8, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta); break;
      case 16: k_spmv_stream<16, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta); break;
      default: k_spmv_stream<32, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta); break;
    }
    return;
  }
  switch(width) {
    case 4: k_spmv<4, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta); break;
    case 8: k_spmv<8, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta); break;
    case 16: k_spmv<16, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta); break;
    default: k_spmv<32, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, z, y, alpha, beta); break;
  }
}

__global__ void k_scale_copy(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = x[i] * a;
}
__global__ void k_scale_inplace(int n, double a, double* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = y[i] * a;
}
__global__ void k_div(int n, const double* __restrict__ x, const double* __restrict__ d,
                      double* __restrict__ y) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = x[i] / d[i];
}

__global__ void k_gather_idx(int n, const int* __restrict__ idx, const double* __restrict__ src,
                             double* __restrict__ dst) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

// column-major dense copy: M[c * n + r] = A_rc
__global__ void k_csr_to_dense(int n, const int* __restrict__ rp, const int* __restrict__ col,
                               const double* __restrict__ val, double* __restrict__ M) {
  const int r = blockIdx.x;
  for(int k = rp[r] + threadIdx.x; k < rp[r + 1]; k += blockDim.x) M[(size_t)col[k] * n + r] = val[k];
}

// One Gauss-Jordan step k, in place, on a column-major n x n matrix B: one
// wavefront per column. Every wavefront reads the pivot column (col_in: column
// k before the step, written by the previous step), finds the pivot row p
// (largest magnitude in rows k..n-1, lowest row on ties; the same in every
// wavefront), swaps rows k and p of its own column and applies the step:
// column j != k: t = b_kj / piv, b_kj = t, b_ij -= c_i t; column k: b_kk =
// 1 / piv, b_ik = -c_i / piv (c = pivot column after the swap). The owner of
// column k + 1 writes its updated column to col_out for the next step.
// Only rows below m take part (pivot search and update): rows m..n-1 are
// identity rows that no step changes (see dense_inverse).
__global__ __launch_bounds__(256) void k_gj_step(int n, int m, int k, double* __restrict__ B,
                                                 const double* __restrict__ col_in, double* __restrict__ col_out,
                                                 int* __restrict__ perm) {
  const int j = blockIdx.x * 4 + (int)threadIdx.x / 64, l = threadIdx.x % 64;
  if(j >= n) return;
  double best = -1.0;
  int bi = m;
  for(int i = k + l; i < m; i += 64) {
    const double v = fabs(col_in[i]);
    if(v > best) { best = v; bi = i; }
  }
  for(int off = 32; off > 0; off >>= 1) {
    const double v2 = __shfl_xor(best, off, 64);
    const int i2 = __shfl_xor(bi, off, 64);
    if(v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
  }
  const int p = bi < m ? bi : k;  // (a column of NaNs has no pivot; stay in bounds)
  const double ck = col_in[k], cp = col_in[p];
  const double piv = cp;
  double* a = B + (size_t)j * n;
  if(j == k) {
    for(int i = l; i < m; i += 64) {
      const double ci = i == k ? cp : (i == p ? ck : col_in[i]);
      a[i] = i == k ? 1.0 / piv : -ci / piv;
    }
  } else {
    const double akj = a[p], apj = a[k];
    const double t = akj / piv;
    for(int i = l; i < m; i += 64) {
      double ai;
      if(i == k) {
        ai = t;
      } else {
        const double ci = i == p ? ck : col_in[i];
        ai = (i == p ? apj : a[i]) - ci * t;
      }
      a[i] = ai;
      if(j == k + 1) col_out[i] = ai;
    }
  }
  if(j == 0 && l == 0) perm[k] = p;
}

// Gauss-Jordan steps on the pivots m..n-1 of a matrix whose rows m..n-1 are
// identity rows: the pivot row is e_k, so no row interchange and no other
// column changes; column k becomes -c (1 on the diagonal). perm[k] = k.
__global__ void k_gj_identity_steps(int n, int m, double* __restrict__ B, int* __restrict__ perm) {
  const size_t cnt = (size_t)(n - m) * m;
  for(size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < cnt; t += (size_t)gridDim.x * blockDim.x) {
    const size_t j = m + t / m, i = t % m;
    B[j * n + i] = -B[j * n + i];
  }
  for(int k = m + blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) perm[k] = k;
}

// column order of the inverse: the row interchanges undone as column
// interchanges in reverse order (one thread)
__global__ void k_gj_colmap(int n, const int* __restrict__ perm, int* __restrict__ colmap, bool lds) {
  extern __shared__ int gj_map[];
  if(blockIdx.x != 0) return;
  // the interchanges run on a copy in LDS when it fits (one thread: each
  // step depends on the last)
  int* cm = lds ? gj_map : colmap;
  for(int j = threadIdx.x; j < n; j += blockDim.x) cm[j] = j;
  __syncthreads();
  if(threadIdx.x == 0)
    for(int k = n - 1; k >= 0; --k) {
      const int p = perm[k], t = cm[k];
      cm[k] = cm[p];
      cm[p] = t;
    }
  __syncthreads();
  if(lds)
    for(int j = threadIdx.x; j < n; j += blockDim.x) colmap[j] = cm[j];
}

// row-major inverse: inv[i][j] = B[i][colmap[j]] (B column-major)
__global__ void k_gj_store(int n, const double* __restrict__ B, const int* __restrict__ colmap,
                           double* __restrict__ inv) {
  const size_t nn = (size_t)n * n;
  for(size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t / n, j = t % n;
    inv[t] = B[(size_t)colmap[j] * n + i];
  }
}

// S = factor A + M on active rows, S = A on Dirichlet rows; 8 lanes per row.
// With sell_val, an active row also goes to its sliced ELL slot (entry k at
// base + 64 k), as the row gather writes it; the sliced ELL layout holds the
// active rows only (sell_slot has n_active entries).
__global__ __launch_bounds__(256) void k_tcd_system(int n, int n_active, const int* __restrict__ rp,
                                                    const double* __restrict__ A, const double* __restrict__ M,
                                                    double factor, double* __restrict__ S,
                                                    const int* __restrict__ sell_slot,
                                                    const long long* __restrict__ sell_off,
                                                    double* __restrict__ sell_val) {
  const int gt = blockIdx.x * 256 + threadIdx.x;
  const int row = gt / 8, l = gt % 8;
  if(row >= n) return;
  const int b = rp[row], e = rp[row + 1];
  const bool to_sell = sell_val != nullptr && row < n_active;
  long long base = 0;
  if(to_sell) {
    const int t = sell_slot[row];
    base = sell_off[t / kSellRows] + (t % kSellRows);
  }
  for(int k = b + l; k < e; k += 8) {
    const double v = row < n_active ? factor * A[k] + M[k] : A[k];
    S[k] = v;
    if(to_sell) sell_val[base + (long long)(k - b) * kSellRows] = v;
  }
}

}  // namespace

void launch_tcd_system(int n, int n_active, const int* rp, const double* A, const double* M, double factor,
                       double* S, hipStream_t s, const int* sell_slot, const long long* sell_off,
                       double* sell_val) {
  if(n == 0) return;
  k_tcd_system<<<(int)(((long)n * 8 + 255) / 256), 256, 0, s>>>(n, n_active, rp, A, M, factor, S, sell_slot,
                                                                sell_off, sell_val);
  PMCD_HIP(hipGetLastError());
}

void launch_scale_copy(int n, double a, const double* x, double* y, hipStream_t s) {
  if(n) k_scale_copy<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, x, y);
}
void launch_scale_inplace(int n, double a, double* y, hipStream_t s) {
  if(n) k_scale_inplace<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, y);
}
void launch_div(int n, const double* x, const double* d, double* y, hipStream_t s) {
  if(n) k_div<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, x, d, y);
}

void launch_gather_idx(int n, const int* idx, const double* src, double* dst, hipStream_t s) {
  if(n) k_gather_idx<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, idx, src, dst);
}

void launch_csr_to_dense(int n, const int* rp, const int* col, const double* val, double* M,
                         hipStream_t s) {
  if(!n) return;
  const long nn = (long)n * n;
  k_fill<double><<<(int)std::min<long>((nn + 255) / 256, 65535), 256, 0, s>>>((int)nn, 0.0, M);
  k_csr_to_dense<<<n, 64, 0, s>>>(n, rp, col, val, M);
}

void dense_inverse(int n, int m, double* B, double* inv, double* colbuf, int* perm, int* colmap,
                   hipStream_t s) {
  if(!n) return;
  PMCD_CHECK(m >= 0 && m <= n, "dense_inverse: bad leading block ", m, " of ", n);
  // rows m..n-1 are identity rows (Dirichlet rows of a level matrix): their
  // steps come first and only negate the leading part of their columns; the
  // remaining steps 0..m-1 then change rows 0..m-1 only. The inverse is the
  // same as with all n steps in order.
  if(m < n)
    k_gj_identity_steps<<<(int)std::max(1L, std::min(((long)(n - m) * std::max(m, 1) + 255) / 256, 4096L)), 256, 0,
                          s>>>(n, m, B, perm);
  launch_copy(n, B, colbuf, s);  // column 0 = the first pivot column
  const int blocks = (n + 3) / 4;
  for(int k = 0; k < m; ++k)
    k_gj_step<<<blocks, 256, 0, s>>>(n, m, k, B, colbuf + (size_t)(k & 1) * n,
                                     colbuf + (size_t)((k + 1) & 1) * n, perm);
  const bool lds_map = (size_t)n * sizeof(int) <= 64 * 1024;
  k_gj_colmap<<<1, 256, lds_map ? (size_t)n * sizeof(int) : 0, s>>>(n, perm, colmap, lds_map);
  const long nn = (long)n * n;
  k_gj_store<<<(int)std::max(1L, std::min((nn + 255) / 256, 8192L)), 256, 0, s>>>(n, B, colmap, inv);
  PMCD_HIP(hipGetLastError());
}

// launch shape of the CSR kernels (tunable): lanes per row and variant
// (0 plain, 1 streaming with XCD-contiguous rows, 2 unrolled where the
// longest row is known and fits)
// defaults: the unrolled CSR kernels (bit-identical to the plain ones and
// faster on every matrix measured); the autotuner may pick others
static int g_spmv_width = 16, g_spmv_variant = 2;
// prolongations of levels that smooth with the sliced ELL kernels: 1 = the
// sliced ELL copy of P (serial row sums), 0 = the CSR kernel
static int g_transfer_sell = 1;
static int g_jacobi_width = 8, g_jacobi_variant = 2;

static int g_assembly_variant = 2;
int assembly_variant() { return g_assembly_variant; }

// SOR sweeps follow the exact level schedule when it has at most this many
// classes, greedy multicolour order otherwise
static int g_sor_exact_limit = 24;
// exact sweeps whose largest class has at most this many rows run in one
// workgroup (one row per thread and class); larger classes are latency-bound
// there and keep one launch per class
static int g_sweep_block_rows = 1024;
int sor_exact_limit() { return g_sor_exact_limit; }

// V-cycles replayed from captured HIP graphs (1) or launched kernel by kernel
// (0, default: replay measured no faster on these kernel sequences)
static int g_vcycle_graphs = 0;
int vcycle_graphs() { return g_vcycle_graphs; }
static int g_cg_device = 1;
int cg_device() { return g_cg_device; }
static int g_cg_blocks = 1024;
int cg_block_cap() { return g_cg_blocks; }
// changes whenever a launch shape changes, so that captured graphs are rebuilt
static int g_tuning_epoch = 0;
// FGMRES queues the next iteration before reading this one's Hessenberg
// column while the residual exceeds this many times the stopping threshold
// (tuning key 15; 0: never)
static int g_fgmres_lookahead = 4;
double fgmres_lookahead() { return g_fgmres_lookahead; }
// cooperative sweeps off by default: measured on the 256-instance batch, one
// cooperative launch per SSOR sweep (15 grid-wide barriers among ~2,000
// resident blocks) took 4.4 ms against 0.66 ms for one launch per colour class
static int g_batch_graphs = 1, g_batch_coop = 0;
int batch_graphs() { return g_batch_graphs; }
int batch_coop() { return g_batch_coop; }
int tuning_epoch() { return g_tuning_epoch; }

void set_tuning(int key, int value) {
  ++g_tuning_epoch;
  if(key == 6) {
    PMCD_CHECK(value == 0 || value == 1, "graph switch must be 0 or 1");
    g_vcycle_graphs = value;
    return;
  }
  if(key == 7) {
    PMCD_CHECK(value == 0 || value == 1, "CG switch must be 0 or 1");
    g_cg_device = value;
    return;
  }
  if(key == 8) {
    PMCD_CHECK(value >= 1 && value <= kCgBlocks, "CG block cap must be in 1..", kCgBlocks);
    g_cg_blocks = value;
    return;
  }
  if(key == 9) {
    set_sell_unroll(value);
    return;
  }
  if(key == 10) {
    set_sell_window(value);
    return;
  }
  if(key == 11) {
    set_sell_min_rows(value);
    return;
  }
  if(key == 13) {
    set_sell_mode(value);
    return;
  }
  if(key == 18) {
    set_sell_blocks_per_cu(value);
    return;
  }
  if(key == 19) {
    PMCD_CHECK(value == 0 || value == 1, "transfer switch must be 0 or 1");
    g_transfer_sell = value;
    return;
  }
  if(key == 16 || key == 17) {
    PMCD_CHECK(value == 0 || value == 1, "batch switch must be 0 or 1");
    (key == 16 ? g_batch_graphs : g_batch_coop) = value;
    return;
  }
  if(key == 15) {
    PMCD_CHECK(value >= 0, "FGMRES look-ahead factor must be non-negative");
    g_fgmres_lookahead = value;
    return;
  }
  if(key == 14) {
    PMCD_CHECK(value == 0 || value == 1, "CSR XCD switch must be 0 or 1");
    g_csr_xcd = value;
    return;
  }
  if(key == 12) {
    PMCD_CHECK(value >= 0, "the one-launch sweep class-size limit must be non-negative");
    g_sweep_block_rows = value;
    return;
  }
  if(key == 5) {
    PMCD_CHECK(value >= 0, "the exact-sweep class limit must be non-negative");
    g_sor_exact_limit = value;
    return;
  }
  if(key == 4) {
    PMCD_CHECK(value >= 0 && value <= 2, "assembly variant must be 0, 1 or 2");
    g_assembly_variant = value;
    return;
  }
  if(key == 0 || key == 1)
    PMCD_CHECK(value == 4 || value == 8 || value == 16 || value == 32, "lanes per row must be 4, 8, 16 or 32");
  else
    PMCD_CHECK(value >= 0 && value <= 4, "variant must be 0, 1, 2, 3 or 4");
  switch(key) {
    case 0: g_spmv_width = value; break;
    case 1: g_jacobi_width = value; break;
    case 2: g_spmv_variant = value; break;
    case 3: g_jacobi_variant = value; break;
    default: PMCD_FAIL("unknown tuning key ", key);
  }
}

int get_tuning(int key) {
  switch(key) {
    case 0: return g_spmv_width;
    case 1: return g_jacobi_width;
    case 2: return g_spmv_variant;
    case 3: return g_jacobi_variant;
    case 4: return g_assembly_variant;
    case 5: return g_sor_exact_limit;
    case 6: return g_vcycle_graphs;
    case 7: return g_cg_device;
    case 8: return g_cg_blocks;
    case 9: return sell_unroll();
    case 10: return get_sell_window_slices();
    case 11: return sell_min_rows();
    case 12: return sweep_block_rows();
    case 13: return sell_mode();
    case 14: return g_csr_xcd;
    case 15: return g_fgmres_lookahead;
    case 16: return g_batch_graphs;
    case 17: return g_batch_coop;
    case 18: return sell_blocks_per_cu();
    case 19: return g_transfer_sell;
    default: PMCD_FAIL("unknown tuning key ", key);
  }
}

void launch_jacobi_from_zero(int n, int n_active, const double* diag, const double* b, double* x_out,
                             double omega, hipStream_t s) {
  if(n) k_jacobi_from_zero<double><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, n_active, diag, b, x_out, omega);
}
void launch_jacobi_from_zero_f32(int n, int n_active, const float* diag, const float* b, float* x_out,
                                 float omega, hipStream_t s) {
  if(n) k_jacobi_from_zero<float><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, n_active, diag, b, x_out, omega);
}

void launch_spmv(int n, int n_active, const int* rp, const int* col, const double* val,
                 const double* x, const double* z, double* y, double alpha, double beta,
                 hipStream_t s, int max_row) {
  if(n == 0) return;
  dispatch_spmv<double>(g_spmv_width, g_spmv_variant, max_row, n, n_active, rp, col, val, x, z, y, alpha, beta, s);
  PMCD_HIP(hipGetLastError());
}

void launch_spmv_rowlen(int n, int n_active, const int* rp, const int* col, const double* val,
                        const double* x, const double* z, double* y, double alpha, double beta,
                        hipStream_t s, int max_row, double mean_row) {
  if(n == 0) return;
  const int W = mean_row <= 12 ? 4 : (mean_row <= 24 ? 8 : 16);
  // the unrolled variant needs 8 or more lanes per row
  dispatch_spmv<double>(W < 8 ? 8 : W, 2, max_row, n, n_active, rp, col, val, x, z, y, alpha, beta, s);
  PMCD_HIP(hipGetLastError());
}

void launch_spmv_f32(int n, int n_active, const int* rp, const int* col, const float* val,
                     const float* x, const float* z, float* y, float alpha, float beta,
                     hipStream_t s, int max_row) {
  if(n == 0) return;
  dispatch_spmv<float>(g_spmv_width, g_spmv_variant, max_row, n, n_active, rp, col, val, x, z, y, alpha, beta, s);
  PMCD_HIP(hipGetLastError());
}

template <class T>
static void dispatch_jacobi(int width, int max_row, int n, int n_active, const int* rp,
                            const int* col, const T* val, const T* diag, const T* b, const T* x,
                            T* x_out, T omega, hipStream_t s) {
  const int blocks = (int)(((long)n * width + 255) / 256);
  const int u = width >= 8 ? unroll_for(max_row, width) : 0;
  // variant 4 (sliced ELL) needs a level's layout; raw CSR calls run unrolled
  const int variant = g_jacobi_variant == 4 ? 2 : g_jacobi_variant;
  if(variant == 2 && u) {
    switch(width) {
      case 8: jacobi_unrolled<8, T>(u, blocks, n, n_active, rp, col, val, diag, b, x, x_out, omega, s); break;
      case 16: jacobi_unrolled<16, T>(u, blocks, n, n_active, rp, col, val, diag, b, x, x_out, omega, s); break;
      default: jacobi_unrolled<32, T>(u, blocks, n, n_active, rp, col, val, diag, b, x, x_out, omega, s); break;
    }
    return;
  }
  if(variant == 3 && u) {
    switch(width) {
      case 8: jacobi_unrolled<8, T, true>(u, blocks, n, n_active, rp, col, val, diag, b, x, x_out, omega, s); break;
      case 16: jacobi_unrolled<16, T, true>(u, blocks, n, n_active, rp, col, val, diag, b, x, x_out, omega, s); break;
      default: jacobi_unrolled<32, T, true>(u, blocks, n, n_active, rp, col, val, diag, b, x, x_out, omega, s); break;
    }
    return;
  }
  if(variant == 1) {
    switch(width) {
      case 4: k_jacobi_stream<4, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega); break;
      case 8: k_jacobi_stream<8, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega); break;
      case 16: k_jacobi_stream<16, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega); break;
      default: k_jacobi_stream<32, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega); break;
    }
    return;
  }
  switch(width) {
    case 4: k_jacobi<4, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega); break;
    case 8: k_jacobi<8, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega); break;
    case 16: k_jacobi<16, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega); break;
    default: k_jacobi<32, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, diag, b, x, x_out, omega); break;
  }
}

void launch_jacobi(int n, int n_active, const int* rp, const int* col, const double* val,
                   const double*, const double* diag, const double* b, const double* x,
                   double* x_out, double omega, hipStream_t s, int max_row) {
  if(n == 0) return;
  dispatch_jacobi<double>(g_jacobi_width, max_row, n, n_active, rp, col, val, diag, b, x, x_out, omega, s);
  PMCD_HIP(hipGetLastError());
}

void launch_jacobi_f32(int n, int n_active, const int* rp, const int* col, const float* val,
                       const float* diag, const float* b, const float* x, float* x_out,
                       float omega, hipStream_t s, int max_row) {
  if(n == 0) return;
  dispatch_jacobi<float>(g_jacobi_width, max_row, n, n_active, rp, col, val, diag, b, x, x_out, omega, s);
  PMCD_HIP(hipGetLastError());
}

// lanes per row from the mean row length, unroll depth from the longest row
template <class T, bool PROLONG>
static void dispatch_transfer(int n, int n_active, long nnz, int max_row, const int* rp,
                              const int* col, const T* val, const T* x, T* y, T a, hipStream_t s) {
  const int rows = PROLONG ? n_active : n;
  if(rows <= 0) return;
  const double mean = n > 0 ? (double)nnz / n : 0.0;
  const int W = mean <= 12 ? 4 : (mean <= 24 ? 8 : 16);
  const int u = unroll_for(max_row, W);
  const int blocks = (int)(((long)rows * W + 255) / 256);
#define PMCD_TRANSFER(WW, UU)                                                                    \
  do {                                                                                           \
    if constexpr(PROLONG)                                                                        \
      k_prolong_add<WW, UU, T><<<blocks, 256, 0, s>>>(n_active, rp, col, val, x, y, a);         \
    else                                                                                         \
      k_restrict<WW, UU, T><<<blocks, 256, 0, s>>>(n, n_active, rp, col, val, x, y);             \
  } while(0)
#define PMCD_TRANSFER_U(WW)                      \
  switch(u) {                                    \
    case 2: PMCD_TRANSFER(WW, 2); break;         \
    case 4: PMCD_TRANSFER(WW, 4); break;         \
    case 8: PMCD_TRANSFER(WW, 8); break;         \
    case 16: PMCD_TRANSFER(WW, 16); break;       \
    default: PMCD_TRANSFER(WW, 0); break;        \
  }
  switch(W) {
    case 4: PMCD_TRANSFER_U(4); break;
    case 8: PMCD_TRANSFER_U(8); break;
    default: PMCD_TRANSFER_U(16); break;
  }
#undef PMCD_TRANSFER_U
#undef PMCD_TRANSFER
  PMCD_HIP(hipGetLastError());
}

void launch_restrict(int n, int n_active, long nnz, int max_row, const int* rp, const int* col,
                     const double* val, const double* x, double* y, hipStream_t s) {
  dispatch_transfer<double, false>(n, n_active, nnz, max_row, rp, col, val, x, y, 0.0, s);
}
void launch_restrict_f32(int n, int n_active, long nnz, int max_row, const int* rp, const int* col,
                         const float* val, const float* x, float* y, hipStream_t s) {
  dispatch_transfer<float, false>(n, n_active, nnz, max_row, rp, col, val, x, y, 0.0f, s);
}
void launch_prolong_add(int n, int n_active, long nnz, int max_row, const int* rp, const int* col,
                        const double* val, const double* x, double* y, double a, hipStream_t s) {
  dispatch_transfer<double, true>(n, n_active, nnz, max_row, rp, col, val, x, y, a, s);
}
void launch_prolong_add_f32(int n, int n_active, long nnz, int max_row, const int* rp,
                            const int* col, const float* val, const float* x, float* y, float a,
                            hipStream_t s) {
  dispatch_transfer<float, true>(n, n_active, nnz, max_row, rp, col, val, x, y, a, s);
}

template <class T>
static void dispatch_sor_rows(int count, const int* rows, const int* rp, const int* col, const T* val,
                              const T* diag, const T* b, T* x, T omega, bool serial, int max_row,
                              double mean_row, hipStream_t s, const int* row_owner,
                              const uint8_t* owner_active) {
  if(count <= 0) return;
  if(serial) {
    k_sor_rows<T><<<(count + 255) / 256, 256, 0, s>>>(count, rows, rp, col, val, diag, b, x, omega,
                                                      row_owner, owner_active);
  } else {
    // lanes per row from the mean row length of the class
    const double mean = mean_row;
    const int W = mean <= 12 ? 4 : (mean <= 24 ? 8 : 16);
    const int u = unroll_for(max_row, W);
    const int blocks = (int)(((long)count * W + 255) / 256);
#define PMCD_SOR_U(WW)                                                                                     \
  switch(u) {                                                                                              \
    case 2: k_sor_rows_lanes<WW, 2, T><<<blocks, 256, 0, s>>>(count, rows, rp, col, val, diag, b, x, omega, row_owner, owner_active); break; \
    case 4: k_sor_rows_lanes<WW, 4, T><<<blocks, 256, 0, s>>>(count, rows, rp, col, val, diag, b, x, omega, row_owner, owner_active); break; \
    case 8: k_sor_rows_lanes<WW, 8, T><<<blocks, 256, 0, s>>>(count, rows, rp, col, val, diag, b, x, omega, row_owner, owner_active); break; \
    case 16: k_sor_rows_lanes<WW, 16, T><<<blocks, 256, 0, s>>>(count, rows, rp, col, val, diag, b, x, omega, row_owner, owner_active); break; \
    default: k_sor_rows_lanes<WW, 0, T><<<blocks, 256, 0, s>>>(count, rows, rp, col, val, diag, b, x, omega, row_owner, owner_active); break; \
  }
    switch(W) {
      case 4: PMCD_SOR_U(4); break;
      case 8: PMCD_SOR_U(8); break;
      default: PMCD_SOR_U(16); break;
    }
#undef PMCD_SOR_U
  }
  PMCD_HIP(hipGetLastError());
}

template <class T>
static void dispatch_sor_perm(int count, int i0, const int* rows, const int* sw_rp, const int* sw_col,
                              const T* sw_val, const T* diag, const T* b, T* x, T omega, int max_row,
                              double mean_row, hipStream_t s, const int* row_owner,
                              const uint8_t* owner_active) {
  if(count <= 0) return;
  const int W = sweep_lanes(mean_row);
  const int u = unroll_for(max_row, W);
  const int blocks = (int)(((long)count * W + 255) / 256);
#define PMCD_SORP(WW, UU)                                                                      \
  k_sor_rows_perm<WW, UU, T><<<blocks, 256, 0, s>>>(count, i0, rows, sw_rp, sw_col, sw_val, diag, b, x, \
                                                    omega, row_owner, owner_active)
#define PMCD_SORP_U(WW)                       \
  switch(u) {                                 \
    case 2: PMCD_SORP(WW, 2); break;          \
    case 4: PMCD_SORP(WW, 4); break;          \
    case 8: PMCD_SORP(WW, 8); break;          \
    case 16: PMCD_SORP(WW, 16); break;        \
    default: PMCD_SORP(WW, 0); break;         \
  }
  switch(W) {
    case 4: PMCD_SORP_U(4); break;
    case 8: PMCD_SORP_U(8); break;
    default: PMCD_SORP_U(16); break;
  }
#undef PMCD_SORP_U
#undef PMCD_SORP
  PMCD_HIP(hipGetLastError());
}

template <class T>
static void dispatch_sweep_block_perm(int K, const int* class_ptr, const int* rows, const int* sw_rp,
                                      const int* sw_col, const T* sw_val, const T* diag, const T* b, T* x, T omega,
                                      int flag, int max_row, double mean_row, hipStream_t s,
                                      const int* row_owner, const uint8_t* owner_active) {
  if(K <= 0) return;
  const int W = sweep_lanes(mean_row);
  const int u = unroll_for(max_row, W);
#define PMCD_SWB(WW, UU)                                                                                   \
  k_sor_sweep_block_perm<WW, UU, T><<<1, 1024, 0, s>>>(K, class_ptr, rows, sw_rp, sw_col, sw_val, diag, b, x, \
                                                       omega, flag, row_owner, owner_active)
#define PMCD_SWB_U(WW)                \
  switch(u) {                         \
    case 2: PMCD_SWB(WW, 2); break;   \
    case 4: PMCD_SWB(WW, 4); break;   \
    case 8: PMCD_SWB(WW, 8); break;   \
    case 16: PMCD_SWB(WW, 16); break; \
    default: PMCD_SWB(WW, 0); break;  \
  }
  switch(W) {
    case 4: PMCD_SWB_U(4); break;
    case 8: PMCD_SWB_U(8); break;
    default: PMCD_SWB_U(16); break;
  }
#undef PMCD_SWB_U
#undef PMCD_SWB
  PMCD_HIP(hipGetLastError());
}

int sweep_lanes(double mean_row) { return mean_row <= 12 ? 4 : (mean_row <= 24 ? 8 : 16); }

bool cooperative_launch_supported() {
  static const bool ok = []() {
    int dev = 0, coop = 0;
    if(hipGetDevice(&dev) != hipSuccess) return false;
    return hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess && coop != 0;
  }();
  return ok;
}

template <int W, int U, class T>
static bool sor_coop_launch(int K, const int* class_ptr, int max_class, const int* rows, const int* sw_rp,
                            const int* sw_col, const T* sw_val, const T* diag, const T* b, T* x, T omega, int flag,
                            const int* row_owner, const uint8_t* owner_active, hipStream_t s) {
  const void* fn = reinterpret_cast<const void*>(&k_sor_sweep_coop<W, U, T>);
  // resident blocks of this instantiation on the whole device (0: no
  // cooperative launches); thread-safe one-time initialisation
  static const long resident = [fn]() -> long {
    int dev = 0, coop = 0, n_cu = 0, nb = 0;
    if(hipGetDevice(&dev) != hipSuccess) return 0;
    if(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return 0;
    if(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0) != hipSuccess) return 0;
    return (long)nb * n_cu;
  }();
  if(resident <= 0) return false;
  const long need = ((long)max_class * W + 255) / 256;
  const int blocks = (int)std::max<long>(1, std::min<long>(need, resident));
  void* args[] = {&K, &class_ptr, &rows, &sw_rp, &sw_col, &sw_val, &diag, &b, &x, &omega, &flag, &row_owner,
                  &owner_active};
  PMCD_HIP(hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(256), args, 0, s));
  return true;
}

template <class T>
static bool dispatch_sor_coop(int K, const int* class_ptr, int max_class, const int* rows, const int* sw_rp,
                              const int* sw_col, const T* sw_val, const T* diag, const T* b, T* x, T omega,
                              int flag, int max_row, double mean_row, const int* row_owner,
                              const uint8_t* owner_active, hipStream_t s) {
  const int W = sweep_lanes(mean_row);
  const int u = unroll_for(max_row, W);
#define PMCD_COOP(WW, UU)                                                                                 \
  return sor_coop_launch<WW, UU, T>(K, class_ptr, max_class, rows, sw_rp, sw_col, sw_val, diag, b, x, omega, \
                                    flag, row_owner, owner_active, s)
#define PMCD_COOP_U(WW)                  \
  switch(u) {                            \
    case 2: PMCD_COOP(WW, 2);            \
    case 4: PMCD_COOP(WW, 4);            \
    case 8: PMCD_COOP(WW, 8);            \
    case 16: PMCD_COOP(WW, 16);          \
    default: PMCD_COOP(WW, 0);           \
  }
  switch(W) {
    case 4: PMCD_COOP_U(4);
    case 8: PMCD_COOP_U(8);
    default: PMCD_COOP_U(16);
  }
#undef PMCD_COOP_U
#undef PMCD_COOP
}

bool launch_sor_sweep_coop(int K, const int* class_ptr, int max_class, const int* rows, const int* sw_rp,
                           const int* sw_col, const double* sw_val, const double* diag, const double* b, double* x,
                           double omega, int flag, int max_row, double mean_row, const int* row_owner,
                           const uint8_t* owner_active, hipStream_t s) {
  return dispatch_sor_coop<double>(K, class_ptr, max_class, rows, sw_rp, sw_col, sw_val, diag, b, x, omega, flag,
                                   max_row, mean_row, row_owner, owner_active, s);
}
bool launch_sor_sweep_coop_f32(int K, const int* class_ptr, int max_class, const int* rows, const int* sw_rp,
                               const int* sw_col, const float* sw_val, const float* diag, const float* b, float* x,
                               float omega, int flag, int max_row, double mean_row, const int* row_owner,
                               const uint8_t* owner_active, hipStream_t s) {
  return dispatch_sor_coop<float>(K, class_ptr, max_class, rows, sw_rp, sw_col, sw_val, diag, b, x, omega, flag,
                                  max_row, mean_row, row_owner, owner_active, s);
}

void launch_sor_sweep_block_perm(int K, const int* class_ptr, const int* rows, const int* sw_rp,
                                 const int* sw_col, const double* sw_val, const double* diag, const double* b,
                                 double* x, double omega, int flag, int max_row, double mean_row, hipStream_t s,
                                 const int* row_owner, const uint8_t* owner_active) {
  dispatch_sweep_block_perm<double>(K, class_ptr, rows, sw_rp, sw_col, sw_val, diag, b, x, omega, flag, max_row,
                                    mean_row, s, row_owner, owner_active);
}
void launch_sor_sweep_block_perm_f32(int K, const int* class_ptr, const int* rows, const int* sw_rp,
                                     const int* sw_col, const float* sw_val, const float* diag, const float* b,
                                     float* x, float omega, int flag, int max_row, double mean_row, hipStream_t s,
                                     const int* row_owner, const uint8_t* owner_active) {
  dispatch_sweep_block_perm<float>(K, class_ptr, rows, sw_rp, sw_col, sw_val, diag, b, x, omega, flag, max_row,
                                   mean_row, s, row_owner, owner_active);
}

void launch_sor_rows_perm(int count, int i0, const int* rows, const int* sw_rp, const int* sw_col,
                          const double* sw_val, const double* diag, const double* b, double* x, double omega,
                          int max_row, double mean_row, hipStream_t s, const int* row_owner,
                          const uint8_t* owner_active) {
  dispatch_sor_perm<double>(count, i0, rows, sw_rp, sw_col, sw_val, diag, b, x, omega, max_row, mean_row, s,
                            row_owner, owner_active);
}
void launch_sor_rows_perm_f32(int count, int i0, const int* rows, const int* sw_rp, const int* sw_col,
                              const float* sw_val, const float* diag, const float* b, float* x, float omega,
                              int max_row, double mean_row, hipStream_t s, const int* row_owner,
                              const uint8_t* owner_active) {
  dispatch_sor_perm<float>(count, i0, rows, sw_rp, sw_col, sw_val, diag, b, x, omega, max_row, mean_row, s,
                           row_owner, owner_active);
}
void launch_permute_rows(int n, const int* rows, const int* rp, const int* sw_rp, const double* val,
                         double* sw_val, hipStream_t s) {
  if(n <= 0) return;
  k_permute_rows<double><<<(int)(((long)n * 8 + 255) / 256), 256, 0, s>>>(n, rows, rp, sw_rp, val, sw_val);
  PMCD_HIP(hipGetLastError());
}
void launch_permute_rows_f32(int n, const int* rows, const int* rp, const int* sw_rp, const float* val,
                             float* sw_val, hipStream_t s) {
  if(n <= 0) return;
  k_permute_rows<float><<<(int)(((long)n * 8 + 255) / 256), 256, 0, s>>>(n, rows, rp, sw_rp, val, sw_val);
  PMCD_HIP(hipGetLastError());
}

void launch_sor_rows(int count, const int* rows, const int* rp, const int* col, const double* val,
                     const double* diag, const double* b, double* x, double omega, bool serial,
                     int max_row, double mean_row, hipStream_t s, const int* row_owner,
                     const uint8_t* owner_active) {
  dispatch_sor_rows<double>(count, rows, rp, col, val, diag, b, x, omega, serial, max_row, mean_row, s,
                            row_owner, owner_active);
}
int sweep_block_rows() { return g_sweep_block_rows; }

void launch_sor_sweep_block(int K, const int* class_ptr, const int* rows, const int* rp, const int* col,
                            const double* val, const double* diag, const double* b, double* x, double omega,
                            int flag, hipStream_t s, const int* row_owner, const uint8_t* owner_active) {
  if(K <= 0) return;
  k_sor_sweep_block<double><<<1, 1024, 0, s>>>(K, class_ptr, rows, rp, col, val, diag, b, x, omega, flag,
                                               row_owner, owner_active);
  PMCD_HIP(hipGetLastError());
}
void launch_sor_sweep_block_f32(int K, const int* class_ptr, const int* rows, const int* rp, const int* col,
                                const float* val, const float* diag, const float* b, float* x, float omega,
                                int flag, hipStream_t s, const int* row_owner, const uint8_t* owner_active) {
  if(K <= 0) return;
  k_sor_sweep_block<float><<<1, 1024, 0, s>>>(K, class_ptr, rows, rp, col, val, diag, b, x, omega, flag,
                                              row_owner, owner_active);
  PMCD_HIP(hipGetLastError());
}

void launch_sor_rows_f32(int count, const int* rows, const int* rp, const int* col, const float* val,
                         const float* diag, const float* b, float* x, float omega, bool serial,
                         int max_row, double mean_row, hipStream_t s, const int* row_owner,
                         const uint8_t* owner_active) {
  dispatch_sor_rows<float>(count, rows, rp, col, val, diag, b, x, omega, serial, max_row, mean_row, s,
                           row_owner, owner_active);
}

int reduce_blocks(int n) {
  int b = (n + 255) / 256;
  return std::max(1, std::min(b, kRedBlocks));
}

void launch_dot(int n, const double* x, const double* y, double* partial, double* out,
                hipStream_t s) {
  const int b = reduce_blocks(n);
  k_dot_partial<<<b, 256, 0, s>>>(n, x, y, partial);
  k_final_sum<<<1, 256, 0, s>>>(b, partial, out);
  PMCD_HIP(hipGetLastError());
}

void launch_sum_max(int n, int stride, const double* v, double* partial, double* out,
                    hipStream_t s) {
  const int b = reduce_blocks(n);
  k_sum_max_partial<<<b, 256, 0, s>>>(n, stride, v, partial);
  k_sum_max_final<<<1, 64, 0, s>>>(b, partial, out);
  PMCD_HIP(hipGetLastError());
}

void launch_axpy(int n, double a, const double* x, double* y, hipStream_t s) {
  k_axpy<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, x, y);
}
void launch_lincomb(int n, const LinComb& lc, double* y, hipStream_t s) {
  PMCD_CHECK(lc.k >= 1 && lc.k <= LinComb::kMax, "linear combination of 1..", LinComb::kMax, " vectors");
  if(n) k_lincomb<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, lc, y);
  PMCD_HIP(hipGetLastError());
}
void launch_fill(int n, double v, double* y, hipStream_t s) {
  if(n) k_fill<double><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, v, y);
}
void launch_fill_f32(int n, float v, float* y, hipStream_t s) {
  if(n) k_fill<float><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, v, y);
}
void launch_stream_copy(long n, const double* x, double* y, hipStream_t s) {
  PMCD_CHECK(n % 2 == 0 && ((uintptr_t)x % 16) == 0 && ((uintptr_t)y % 16) == 0,
             "stream copy needs an even count of 16-byte aligned doubles");
  if(n) k_stream_copy<<<2048, 256, 0, s>>>(n / 2, (const double2*)x, (double2*)y);
  PMCD_HIP(hipGetLastError());
}

void launch_stream_copy4(long n, const double* x, double* y, hipStream_t s) {
  PMCD_CHECK(n % 2 == 0 && ((uintptr_t)x % 16) == 0 && ((uintptr_t)y % 16) == 0,
             "stream copy needs an even count of 16-byte aligned doubles");
  if(n) k_stream_copy4<<<8192, 256, 0, s>>>(n / 2, (const double2*)x, (double2*)y);
  PMCD_HIP(hipGetLastError());
}

void launch_stream_read(long n, const double* x, double* out, hipStream_t s) {
  PMCD_CHECK(n % 2 == 0 && ((uintptr_t)x % 16) == 0, "stream read needs an even count of aligned doubles");
  if(n) k_stream_read<<<kStreamReadBlocks, 256, 0, s>>>(n / 2, (const double2*)x, out);
  PMCD_HIP(hipGetLastError());
}

void launch_copy(int n, const double* x, double* y, hipStream_t s) {
  if(n) k_copy<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, x, y);
}
void launch_mgs_step(int n, const double* h, double* r, const double* v, const double* vn,
                     double* partial, double* out_next, hipStream_t s) {
  const int b = reduce_blocks(n);
  k_mgs_step<<<b, 256, 0, s>>>(n, h, r, v, vn, partial);
  if(vn) k_final_sum<<<1, 256, 0, s>>>(b, partial, out_next);
}
void launch_dot_partial(int n, const double* x, const double* y, double* partial, hipStream_t s) {
  k_dot_partial<<<reduce_blocks(n), 256, 0, s>>>(n, x, y, partial);
  PMCD_HIP(hipGetLastError());
}
void launch_mgs_fused(int n, const double* prev, double* h_out, double* r, const double* v, const double* wn,
                      double* partial, hipStream_t s) {
  const int b = reduce_blocks(n);
  k_mgs_fused<<<b, 256, 0, s>>>(n, b, prev, h_out, r, v, wn, partial);
  PMCD_HIP(hipGetLastError());
}
void launch_scale_norm(int n, const double* prev, double* h_out, const double* r, double* v, hipStream_t s) {
  const int b = reduce_blocks(n);
  k_scale_norm<<<b, 256, 0, s>>>(n, b, prev, h_out, r, v);
  PMCD_HIP(hipGetLastError());
}
void launch_fgmres_update(int n, int k, const double* y, const double* const* z, double* x, hipStream_t s) {
  if(n && k) k_fgmres_update<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, k, y, z, x);
  PMCD_HIP(hipGetLastError());
}
void launch_add_assembled(int n, int n_active, const int* row_ptr, const int* diag_pos, const double* val,
                          const double* f, double* entries, double* rhs, hipStream_t s) {
  int na_entries = 0;
  PMCD_HIP(hipMemcpyAsync(&na_entries, row_ptr + n_active, sizeof(int), hipMemcpyDeviceToHost, s));
  PMCD_HIP(hipStreamSynchronize(s));
  const long long total = (long long)na_entries + n;
  if(total == 0) return;
  const long long g = std::min<long long>((total + 255) / 256, 65535LL * 16);
  k_add_assembled<<<(int)g, 256, 0, s>>>(n, n_active, na_entries, diag_pos, val, f, entries, rhs);
  PMCD_HIP(hipGetLastError());
}
void launch_d2f(int n, const double* x, float* y, hipStream_t s) {
  if(n) k_d2f<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, x, y);
}
void launch_f2d(int n, const float* x, double* y, hipStream_t s) {
  if(n) k_f2d<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, x, y);
}
void launch_gemv(int n, const double* M, const double* x, double* y, hipStream_t s) {
  if(n) k_gemv<double><<<(n + 3) / 4, 256, 0, s>>>(n, M, x, y);
}
}  // namespace pmcd
