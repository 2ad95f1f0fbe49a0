// Hanging-DOF coupling of an assembled matrix and right-hand side
// (Assemble2D.C:3350-3427 with FEMatrix::ModifyMatrixAccordingToCoupling and
// correct_hanging_rows, src/Matrix/FEMatrix.C:244-296, 384-418).
//
// The reference walks the sorted hanging DOFs one after another and adds each
// hanging row, weighted, to the rows of its coupling DOFs. Here every target
// entry gathers its sources instead: item i adds w[s] * v[src[s]] for s in
// sp[i] .. sp[i+1] to v[dst[i]], in the order of the sorted hanging list, so
// each entry takes the same additions in the same order. A hanging row that
// receives from other hanging rows is finished in an earlier launch (its
// depth in the coupling graph), so the sources of a launch are final and no
// launch writes what it reads.
#include "device.hpp"
#include "kernels.hpp"

namespace pmcd {

namespace {

__global__ void __launch_bounds__(256) k_hanging_gather(int n_mat, const int* __restrict__ mat_dst,
                                                        const int* __restrict__ mat_sp,
                                                        const int* __restrict__ mat_src,
                                                        const double* __restrict__ mat_w, double* val, int n_rhs,
                                                        const int* __restrict__ rhs_dst,
                                                        const int* __restrict__ rhs_sp,
                                                        const int* __restrict__ rhs_src,
                                                        const double* __restrict__ rhs_w, double* rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n_mat) {
    double acc = val[mat_dst[i]];
    for(int s = mat_sp[i]; s < mat_sp[i + 1]; ++s) acc += mat_w[s] * val[mat_src[s]];
    val[mat_dst[i]] = acc;
  } else if(i < n_mat + n_rhs) {
    const int j = i - n_mat;
    double acc = rhs[rhs_dst[j]];
    for(int s = rhs_sp[j]; s < rhs_sp[j + 1]; ++s) acc += rhs_w[s] * rhs[rhs_src[s]];
    rhs[rhs_dst[j]] = acc;
  }
}

// hanging rows become u_h - sum_k w_k u_{c_k} = 0, their rhs entries 0
__global__ void __launch_bounds__(256) k_hanging_fix(int n_fix, const int* __restrict__ pos,
                                                     const double* __restrict__ v, double* val, int rhs_first,
                                                     int n_hang, double* rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n_fix) val[pos[i]] = v[i];
  else if(i < n_fix + n_hang) rhs[rhs_first + i - n_fix] = 0.0;
}

}  // namespace

void launch_hanging_gather(int n_mat, const int* mat_dst, const int* mat_sp, const int* mat_src, const double* mat_w,
                           double* val, int n_rhs, const int* rhs_dst, const int* rhs_sp, const int* rhs_src,
                           const double* rhs_w, double* rhs, hipStream_t s) {
  const int n = n_mat + n_rhs;
  if(n) k_hanging_gather<<<(n + 255) / 256, 256, 0, s>>>(n_mat, mat_dst, mat_sp, mat_src, mat_w, val, n_rhs, rhs_dst,
                                                         rhs_sp, rhs_src, rhs_w, rhs);
}

void launch_hanging_fix(int n_fix, const int* pos, const double* v, double* val, int rhs_first, int n_hang,
                        double* rhs, hipStream_t s) {
  const int n = n_fix + n_hang;
  if(n) k_hanging_fix<<<(n + 255) / 256, 256, 0, s>>>(n_fix, pos, v, val, rhs_first, n_hang, rhs);
}

}  // namespace pmcd
