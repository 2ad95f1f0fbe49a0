// Launch wrappers of the HIP kernels (implemented in k_assemble.hip and
// k_linalg.hip). All take device pointers and a stream.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "examples.hpp"

namespace pmcd {

// ---- assembly ------------------------------------------------------------
struct AsmArgs {
  int shape;        // Shape
  int dim;
  int nv;           // vertices per cell
  int nl;           // local DOFs
  int nq;           // quadrature points
  int ncomp;        // derivative components in the reference table (6 or 10)
  int supg;         // 1: add the SUPG/SD local routine
  int n_cells;
  ExampleParams ex;
  // batched instances: cell c uses cell_ex[cell_inst[c]] instead of ex
  const ExampleParams* cell_ex = nullptr;
  const int* cell_inst = nullptr;
  const double* ref;        // [nq][ncomp][nl] reference basis values/derivatives
  const double* qw;         // [nq] reference weights
  const double* qp;         // [nq][3] reference points
  const double* cell_xyz;   // [n_cells][nv][3]
  double* local_A;          // [n_cells][nl][nl]  (row = test, column = ansatz)
  double* local_rhs;        // [n_cells][nl]
  // time-dependent forms (TCDStiff, TCDMass, TCDRhs and their SUPG parts,
  // src/AssembleRoutines/CD_local_assembling_routines.C:9-387): local_A gets
  // the stiffness matrix, local_M the mass matrix
  int tcd = 0;
  double* local_M = nullptr;  // [n_cells][nl][nl]
  // reference tensors T_p[i][j] of the component pairs, [NP padded to 4]
  // [nl^2 padded to 16] (null: not available, the per-point kernels run)
  const double* tensor = nullptr;
};
__host__ __device__ inline const ExampleParams& cell_params(const AsmArgs& a, int cell) {
  return a.cell_inst ? a.cell_ex[a.cell_inst[cell]] : a.ex;
}
// P1 triangles/tetrahedra take a thread-per-cell kernel (same arithmetic;
// not for the time-dependent forms)
bool p1_simplex(const AsmArgs& a);
void launch_assemble_local(const AsmArgs& a, hipStream_t s);
// FP64-MFMA variant; returns false if the element is not covered
bool launch_assemble_mfma(const AsmArgs& a, hipStream_t s);
// reference-tensor variant (constant matrix coefficients, affine cells):
// one FP64-MFMA GEMM per level; returns false if not covered
bool launch_assemble_tensor(const AsmArgs& a, hipStream_t s);
// component pairs of the reference tensors in dimension dim, and the tensors
// themselves ([pairs padded to 4][nl^2 padded to 16]) from the reference
// table ref[q][component][i] and the weights qw of an nq-point rule
int tensor_pairs(int dim);
std::vector<double> build_assembly_tensor(int dim, int nl, int nq, const double* ref, const double* qw);
// local matrices and rhs of every cell by the variant selected (assembly
// variant 0: reference-order vector kernels; 1: per-point FP64-MFMA kernel for
// 27-DOF hexahedra, vector kernels otherwise; 2: reference tensors where
// available, else as 1)
void assemble_cells(const AsmArgs& a, hipStream_t s);
void launch_mfma_probe(double* out, hipStream_t s);
int assembly_variant();

// Sum local matrices into CSR rows in increasing cell order; Dirichlet rows
// (r >= n_active) get 1 on the diagonal and 0 elsewhere. With sell_val the
// active rows are also written into a sliced ELL copy (sell_slot: row ->
// slot, sell_off: first entry of each slice).
void launch_gather_matrix(int n_rows, int n_active, int nl, const int* row_ptr,
                          const int* col, const int* contrib_ptr, const int* contrib,
                          const uint8_t* slot, const double* local_A, double* val,
                          hipStream_t s, const int* sell_slot = nullptr,
                          const long long* sell_off = nullptr, double* sell_val = nullptr);
// rhs[r] = sum of local rhs (plus Neumann terms) in cell order for active rows,
// Dirichlet value for r >= n_active.
void launch_gather_rhs(int n_rows, int n_active, const int* contrib_ptr, const int* contrib,
                       const double* local_rhs, const int* extra_ptr, const double* extra,
                       const double* dirichlet_values, double* rhs, hipStream_t s);

// ---- errors ---------------------------------------------------------------
// per-cell {L2^2, H1semi^2, SD^2, DG, Linf}
void launch_cell_errors(const AsmArgs& a, const int* dof, const double* u, double* cell_err,
                        hipStream_t s);

// ---- sparse / dense linear algebra ----------------------------------------
// y[r] = beta*z[r] + alpha*(A x)[r] for r < n_active, beta*z[r] + alpha*x[r] otherwise
void launch_spmv(int n, int n_active, const int* rp, const int* col, const double* val,
                 const double* x, const double* z, double* y, double alpha, double beta,
                 hipStream_t s, int max_row = -1);
void launch_spmv_f32(int n, int n_active, const int* rp, const int* col, const float* val,
                     const float* x, const float* z, float* y, float alpha, float beta,
                     hipStream_t s, int max_row = -1);
// same, launch shape from the row lengths (8 or 16 lanes, unrolled loads)
void launch_spmv_rowlen(int n, int n_active, const int* rp, const int* col, const double* val,
                        const double* x, const double* z, double* y, double alpha, double beta,
                        hipStream_t s, int max_row, double mean_row);
// damped Jacobi sweep x_out = x + omega * (b - A x) / diag (non-active rows use A = I)
void launch_jacobi(int n, int n_active, const int* rp, const int* col, const double* val,
                   const double* inv_diag_unused, const double* diag, const double* b,
                   const double* x, double* x_out, double omega, hipStream_t s,
                   int max_row = -1);
void launch_jacobi_f32(int n, int n_active, const int* rp, const int* col, const float* val,
                       const float* diag, const float* b, const float* x, float* x_out,
                       float omega, hipStream_t s, int max_row = -1);
// first sweep from x = 0 when b vanishes on the non-active rows (no SpMV needed)
void launch_jacobi_from_zero(int n, int n_active, const double* diag, const double* b,
                             double* x_out, double omega, hipStream_t s);
void launch_jacobi_from_zero_f32(int n, int n_active, const float* diag, const float* b,
                                 float* x_out, float omega, hipStream_t s);
// lanes per row of the CSR kernels: key 0 SpMV, 1 Jacobi; value 4, 8, 16 or 32
void set_tuning(int key, int value);
int get_tuning(int key);
// plain CSR product y = A x over all rows (transfer operators)
// one class of an SOR sweep: rows[0..count) updated in place; serial: one
// thread per row summing in column order (the reference's rounding), else
// 4/8/16 lanes per row (from the mean row length) with a butterfly sum
// (rows of at most max_row entries)
// (row_owner, owner_active): optional per-row owner and per-owner flags; rows
// of inactive owners are skipped
void launch_sor_rows(int count, const int* rows, const int* rp, const int* col, const double* val,
                     const double* diag, const double* b, double* x, double omega, bool serial,
                     int max_row, double mean_row, hipStream_t s, const int* row_owner = nullptr,
                     const uint8_t* owner_active = nullptr);
void launch_sor_rows_f32(int count, const int* rows, const int* rp, const int* col, const float* val,
                         const float* diag, const float* b, float* x, float omega, bool serial,
                         int max_row, double mean_row, hipStream_t s, const int* row_owner = nullptr,
                         const uint8_t* owner_active = nullptr);
int sor_exact_limit();
// one-workgroup sweeps (all classes in one launch) on levels whose largest
// class has at most this many rows (exact sweeps) or row lanes (multicolour
// sweeps, W lanes per row); tuning key 12, 0 switches them off
int sweep_block_rows();
void launch_sor_sweep_block(int K, const int* class_ptr, const int* rows, const int* rp, const int* col,
                            const double* val, const double* diag, const double* b, double* x, double omega,
                            int flag, hipStream_t s, const int* row_owner, const uint8_t* owner_active);
// lanes per row of the multicolour class kernels, from the mean row length
int sweep_lanes(double mean_row);
void launch_sor_sweep_block_perm(int K, const int* class_ptr, const int* rows, const int* sw_rp,
                                 const int* sw_col, const double* sw_val, const double* diag, const double* b,
                                 double* x, double omega, int flag, int max_row, double mean_row, hipStream_t s,
                                 const int* row_owner, const uint8_t* owner_active);
void launch_sor_sweep_block_perm_f32(int K, const int* class_ptr, const int* rows, const int* sw_rp,
                                     const int* sw_col, const float* sw_val, const float* diag, const float* b,
                                     float* x, float omega, int flag, int max_row, double mean_row, hipStream_t s,
                                     const int* row_owner, const uint8_t* owner_active);
void launch_sor_sweep_block_f32(int K, const int* class_ptr, const int* rows, const int* rp, const int* col,
                                const float* val, const float* diag, const float* b, float* x, float omega,
                                int flag, hipStream_t s, const int* row_owner, const uint8_t* owner_active);
bool cooperative_launch_supported();
// a whole multicolour sweep (class-ordered copy) in one cooperative launch
// with a grid-wide barrier between classes; false when the device does not
// support cooperative launches (nothing launched)
bool launch_sor_sweep_coop(int K, const int* class_ptr, int max_class, const int* rows, const int* sw_rp,
                           const int* sw_col, const double* sw_val, const double* diag, const double* b, double* x,
                           double omega, int flag, int max_row, double mean_row, const int* row_owner,
                           const uint8_t* owner_active, hipStream_t s);
bool launch_sor_sweep_coop_f32(int K, const int* class_ptr, int max_class, const int* rows, const int* sw_rp,
                               const int* sw_col, const float* sw_val, const float* diag, const float* b, float* x,
                               float omega, int flag, int max_row, double mean_row, const int* row_owner,
                               const uint8_t* owner_active, hipStream_t s);
// multicolour class [i0, i0 + count) over the class-ordered matrix copy
// (sw_rp, sw_col, sw_val: permuted rows, original entry order)
void launch_sor_rows_perm(int count, int i0, const int* rows, const int* sw_rp, const int* sw_col,
                          const double* sw_val, const double* diag, const double* b, double* x, double omega,
                          int max_row, double mean_row, hipStream_t s, const int* row_owner = nullptr,
                          const uint8_t* owner_active = nullptr);
void launch_sor_rows_perm_f32(int count, int i0, const int* rows, const int* sw_rp, const int* sw_col,
                              const float* sw_val, const float* diag, const float* b, float* x, float omega,
                              int max_row, double mean_row, hipStream_t s, const int* row_owner = nullptr,
                              const uint8_t* owner_active = nullptr);
// sw_val[sw_rp[i] + t] = val[rp[rows[i]] + t]
void launch_permute_rows(int n, const int* rows, const int* rp, const int* sw_rp, const double* val,
                         double* sw_val, hipStream_t s);
void launch_permute_rows_f32(int n, const int* rows, const int* rp, const int* sw_rp, const float* val,
                             float* sw_val, hipStream_t s);
int vcycle_graphs();
// 1: CG with no or Jacobi preconditioning keeps its scalars on the device
int cg_device();
// largest grid of the device CG kernels (1..kCgBlocks)
int cg_block_cap();
int tuning_epoch();
// FGMRES look-ahead factor (tuning key 15): 0 off
double fgmres_lookahead();
// restriction y = R x on active rows, 0 on rows >= n_active
void launch_restrict(int n, int n_active, long nnz, int max_row, const int* rp, const int* col,
                     const double* val, const double* x, double* y, hipStream_t s);
void launch_restrict_f32(int n, int n_active, long nnz, int max_row, const int* rp, const int* col,
                         const float* val, const float* x, float* y, hipStream_t s);
// prolongation y[r] += a (P x)[r] for r < n_active (n rows, nnz entries in P)
void launch_prolong_add(int n, int n_active, long nnz, int max_row, const int* rp, const int* col,
                        const double* val, const double* x, double* y, double a, hipStream_t s);
void launch_prolong_add_f32(int n, int n_active, long nnz, int max_row, const int* rp,
                            const int* col, const float* val, const float* x, float* y, float a,
                            hipStream_t s);

// deterministic reductions: result written to out[0] (device); the first
// pass writes reduce_blocks(n) partials, at most kMaxReduceBlocks
constexpr int kMaxReduceBlocks = 1024;
int reduce_blocks(int n);
// device scratch of the allocation-free entry point: the reduction partials
// (at most kMaxReduceBlocks) and the result, per device
constexpr int kNoAllocScratchDoubles = 2048;
double* noalloc_scratch();
void launch_dot(int n, const double* x, const double* y, double* partial, double* out,
                hipStream_t s);
// first pass of launch_dot only: reduce_blocks(n) partials
void launch_dot_partial(int n, const double* x, const double* y, double* partial, hipStream_t s);
// modified Gram-Schmidt step on the previous reduction's partials (prev, from
// a kernel over the same n): h = sum(prev) into *h_out, r -= h v, partials of
// (r, wn), or of (r, r) when wn is null
void launch_mgs_fused(int n, const double* prev, double* h_out, double* r, const double* v, const double* wn,
                      double* partial, hipStream_t s);
// v = r / sqrt(sum(prev)), the norm into *h_out
void launch_scale_norm(int n, const double* prev, double* h_out, const double* r, double* v, hipStream_t s);
// x += sum_{j<k} y[j] z[j] (device arrays y and z)
void launch_fgmres_update(int n, int k, const double* y, const double* const* z, double* x, hipStream_t s);
void launch_sum_max(int n, int stride, const double* v, double* partial, double* out,
                    hipStream_t s);

// vector kernels
void launch_axpy(int n, double alpha, const double* x, double* y, hipStream_t s);
// y = sum_{j < k} c[j] v[j] (device vectors of length n)
struct LinComb {
  static constexpr int kMax = 32;
  int k = 0;
  const double* v[kMax];
  double c[kMax];
};
void launch_lincomb(int n, const LinComb& lc, double* y, hipStream_t s);
void launch_fill(int n, double v, double* y, hipStream_t s);
void launch_copy(int n, const double* x, double* y, hipStream_t s);
// y = x for n doubles (n even, 16-byte aligned), 16-byte vector accesses
void launch_stream_copy(long n, const double* x, double* y, hipStream_t s);
void launch_stream_copy4(long n, const double* x, double* y, hipStream_t s);
// out[w] = partial sums of x per wave (8192 x 4 entries)
// out must hold kStreamReadBlocks * 4 doubles (one partial sum per wavefront)
constexpr int kStreamReadBlocks = 32768;
void launch_stream_read(long n, const double* x, double* out, hipStream_t s);
// fused MGS step: r -= h_k v_k; h_{k+1} = <r, v_{k+1}> partials (v_next may be null)
void launch_mgs_step(int n, const double* h_dev, double* r, const double* v,
                     const double* v_next, double* partial, double* out_next, hipStream_t s);
void launch_d2f(int n, const double* x, float* y, hipStream_t s);
void launch_f2d(int n, const float* x, double* y, hipStream_t s);
void launch_fill_f32(int n, float v, float* y, hipStream_t s);
// dense y = M x (M row-major n x n)
void launch_gemv(int n, const double* M, const double* x, double* y, hipStream_t s);
// y = a x ; y *= a ; y = x / d
void launch_scale_copy(int n, double a, const double* x, double* y, hipStream_t s);
void launch_scale_inplace(int n, double a, double* y, hipStream_t s);
void launch_div(int n, const double* x, const double* d, double* y, hipStream_t s);
// dst[i] = src[idx[i]]
void launch_gather_idx(int n, const int* idx, const double* src, double* dst, hipStream_t s);
// hanging-DOF coupling: item i < n_mat adds mat_w[s] * val[mat_src[s]],
// s = mat_sp[i] .. mat_sp[i+1], to val[mat_dst[i]] in list order; items
// n_mat .. n_mat + n_rhs do the same on rhs (one launch per coupling depth)
void launch_hanging_gather(int n_mat, const int* mat_dst, const int* mat_sp, const int* mat_src, const double* mat_w,
                           double* val, int n_rhs, const int* rhs_dst, const int* rhs_sp, const int* rhs_src,
                           const double* rhs_w, double* rhs, hipStream_t s);
// val[pos[i]] = v[i] (the hanging rows' constraint entries), rhs[rhs_first ..
// rhs_first + n_hang) = 0
void launch_hanging_fix(int n_fix, const int* pos, const double* v, double* val, int rhs_first, int n_hang,
                        double* rhs, hipStream_t s);
// caller arrays += an assembly (val, f: the gathered CSR values and rhs):
// active-row entries and active rhs added, Dirichlet diagonals set to 1 and
// Dirichlet rhs set (either output may be null)
void launch_add_assembled(int n, int n_active, const int* row_ptr, const int* diag_pos, const double* val,
                          const double* f, double* entries, double* rhs, hipStream_t s);
// scatter CSR rows into a dense column-major n x n matrix (zero-filled first)
void launch_csr_to_dense(int n, const int* rp, const int* col, const double* val, double* M,
                         hipStream_t s);
// Gauss-Jordan inverse with partial pivoting, one launch per pivot: B
// (column-major, overwritten) -> inv (row-major); colbuf holds 2 n doubles,
// perm and colmap n ints each; rows m..n-1 of B must be identity rows (m = n:
// a general matrix)
void dense_inverse(int n, int m, double* B, double* inv, double* colbuf, int* perm, int* colmap,
                   hipStream_t s);

// time-dependent system matrix: S = factor A + M on active rows, A on the
// Dirichlet rows (scale_blocks_actives + add_matrix_actives)
// algebraic flux correction (k_afc.hip)
// the limiter's per-row sums and bounds, kept for the Newton scheme when set
struct AfcLimiterOut {
  double *Pp = nullptr, *Pm = nullptr, *Qp = nullptr, *Qm = nullptr;
  double *umax = nullptr, *umin = nullptr, *q = nullptr;  // BJK17 only
};
struct AfcNewtonArgs {
  int n, n_active;
  const int *rp, *col, *tpos;
  const double *D, *E, *alpha, *u, *f;
  const double *Pp, *Pm, *Qp, *Qm, *Rp, *Rm, *umax, *umin, *q;
  double tau0, tau1;
  int bjk;
  double *S, *rhs;
};
void launch_afc_newton(const AfcNewtonArgs& a, hipStream_t s);
void launch_afc_transpose(int n, const int* rp, const int* col, int* tpos, hipStream_t s);
void launch_afc_matrices(int n, int n_active, const int* rp, const int* col, const int* tpos, const double* A,
                         double* D, double* E, double* S, hipStream_t s);
void launch_afc_rhs(int n, int n_active, const int* rp, const int* col, const int* tpos, const double* D,
                    const double* E, const double* u, const double* f, double* Rp, double* Rm, double* alpha,
                    double* rhs, hipStream_t s, AfcLimiterOut lo = {});
void launch_afc_damp(int n, int n_active, double omega, const double* old, double* u, hipStream_t s);
void launch_afc_fpm(int n, int n_active, const int* rp, const int* col, const double* D, const double* E,
                    const double* alpha, const double* u, const double* f, double tau0, double* S, double* rhs,
                    hipStream_t s);
void launch_afc_bjk_modify(int n_active, const int* rp, const int* col, const int* tpos, double* A,
                           hipStream_t s);
void launch_afc_bjk_rhs(int n, int n_active, const int* rp, const int* col, const double* D, const double* gamma,
                        const double* u, const double* f, double* Rp, double* Rm, double* alpha, double* rhs,
                        hipStream_t s, AfcLimiterOut lo = {});
void launch_tcd_system(int n, int n_active, const int* rp, const double* A, const double* M, double factor,
                       double* S, hipStream_t s, const int* sell_slot = nullptr,
                       const long long* sell_off = nullptr, double* sell_val = nullptr);

// ---- batched (segmented) vector kernels ------------------------------------
// A chunk is a row range [begin, end) owned by one instance.
struct SegChunk {
  int inst, begin, end;
};
// per-instance state of the device-resident batched FGMRES
struct BatchInstState {
  double initial, previous, resid, beta;
  int iterations;  // FGMRES iterations done
  int converged;   // residual below the tolerance or reduced enough
  int update;      // the next update kernel adds sum_{k < n_coef} y_k z_k
  int n_coef;
};
struct BatchSolveParams {
  double res_tol, res_red, div_factor;
  int min_iter, max_iter, m;
};
void launch_seg_mgs2(int n_chunks, const SegChunk* ch, const uint8_t* active, const int* inst_ptr,
                     const double* prev, double* hcol_k, const double* v, double* w, const double* vnext,
                     double* partial, hipStream_t s);
void launch_seg_normscale(int n_chunks, const SegChunk* ch, const uint8_t* active, const int* inst_ptr,
                          const double* prev, double* hcol, const double* w, double* vout, hipStream_t s);
void launch_seg_dot_partial(int n_chunks, const SegChunk* ch, const uint8_t* active, const double* x,
                            const double* y, double* partial, hipStream_t s);
void launch_batch_givens(int ni, int i, const BatchSolveParams& p, const double* hcol, double* H, double* sv,
                         double* cs, double* sn, BatchInstState* st, double* y, uint8_t* active, int* breakdown,
                         hipStream_t s);
void launch_batch_restart_coef(int ni, const BatchSolveParams& p, const double* H, const double* sv,
                               BatchInstState* st, double* y, const uint8_t* active, hipStream_t s);
void launch_seg_update(int n_chunks, const SegChunk* ch, const BatchInstState* st, int m, const double* y,
                       const double* const* z, double* x, hipStream_t s);
// mode 0: start of a solve (test at iteration 0 on beta), 1: restart
void launch_batch_restart_state(int ni, const BatchSolveParams& p, const int* inst_ptr, const double* partial,
                                BatchInstState* st, double* sv, double* scale, uint8_t* active, int mode,
                                hipStream_t s);
// batched solves: iteration graphs (tuning key 16) and cooperative sweeps
// (tuning key 17) on or off
int batch_graphs();
int batch_coop();
constexpr int kChunkRows = 4096;
// out[i] = (x, y) restricted to instance i's rows (partial: one per chunk)
// `active` (may be null): per-instance flags; chunks of inactive instances
// contribute 0 and are not read
void launch_seg_dot(int n_chunks, const SegChunk* ch, int n_inst, const int* inst_ptr,
                    const double* x, const double* y, double* partial, double* out, hipStream_t s,
                    const uint8_t* active = nullptr);
// w -= hprev[inst] vprev (vprev may be null), then out[i] = (w, vnext) per
// instance (vnext may be null)
void launch_seg_mgs(int n_chunks, const SegChunk* ch, int n_inst, const int* inst_ptr,
                    const double* hprev, const double* vprev, double* w, const double* vnext,
                    double* partial, double* out, hipStream_t s, const uint8_t* active = nullptr);
// y = a[inst] x ; y += a[inst] x
void launch_seg_scale(int n_chunks, const SegChunk* ch, const double* a, const double* x, double* y,
                      hipStream_t s, const uint8_t* active = nullptr);
// SpMV with the non-active rule over a batch: rows whose instance is inactive
// are set to 0 without reading the matrix (rows of at most 64 entries)
void launch_spmv_masked(int n, int n_active, const int* rp, const int* col, const double* val,
                        const double* x, const double* z, double* y, double alpha, double beta,
                        int max_row, const int* row_inst, const uint8_t* active, hipStream_t s);
void launch_seg_axpy(int n_chunks, const SegChunk* ch, const double* a, const double* x, double* y,
                     hipStream_t s);

// ---- sliced ELL (one wavefront per slice of kSellRows active rows) ----------
constexpr int kSellRows = 64;
// rows per length-sorting window of the sliced ELL layout: sell_window()
// slices (default 4), read when a level's layout is built
int sell_window();
// levels with fewer active rows keep the unrolled CSR kernels under variant 4
int sell_min_rows();
int sell_blocks_per_cu();
void set_sell_blocks_per_cu(int b);
void set_sell_min_rows(int n);
template <class T>
struct SellView {
  int n_slots;                 // active rows rounded up to whole slices
  const long long* off;        // first entry of each slice
  const int* perm;             // slot -> row (-1: padding)
  const int* len;              // slot -> row length
  const int* col;
  const T* val;
  int window;                  // rows per length-sorted window of the layout
  // column footprints of the 256-slot blocks (null: none): block-local
  // column of every entry, first footprint entry of each block, the
  // footprints' global columns, the widest footprint
  const unsigned short* lcol;
  const int* fp_off;
  const int* fp;
  int fp_max;
};
// widest block footprint (distinct columns) the LDS-staged kernels take
constexpr int kSellMaxFootprint = 4096;
template <class T>
void launch_spmv_sell(int n, int n_active, const SellView<T>& m, const T* x, const T* z, T* y, T alpha, T beta,
                      hipStream_t s);
template <class T>
void launch_jacobi_sell(int n, int n_active, const SellView<T>& m, const T* diag, const T* b, const T* x, T* x_out,
                        T omega, hipStream_t s);
// copy the active rows' CSR entries (values or columns) into the sliced layout
template <class T>
void launch_sell_fill(int n_slots, const long long* off, const int* perm, const int* len,
                      const int* rp, const T* src, T* dst, hipStream_t s);
// sliced ELL launch mode (tuning key 13): bit 0 XCD-contiguous slices, bit 1
// row-ordered output through LDS (used only with 4-slice windows), bit 2
// non-temporal matrix loads, bit 3 pipelined matrix loads, bit 4 x staged in
// LDS per block through the column footprints (with bit 1 and footprints)
int sell_mode();
void set_sell_mode(int mode);
// the mode a launch uses on a layout of `window` rows per window whose
// widest block footprint is fp_max (0: none)
int sell_mode_for(int window, int fp_max);
// entries of a lane loaded per step (4, 8 or 16)
int sell_unroll();
void set_sell_unroll(int u);
void set_sell_window(int slices);
int get_sell_window_slices();

// ---- conjugate gradients with device-resident scalars ----------------------
// State of one CG solve. Scalars that a kernel both reads and writes sit in
// two slots (iteration k in slot k & 1), so that no block reads a slot that
// another block of the same launch writes; block 0 writes, every block
// computes the same values. The host copies the state back now and then to
// see whether the iteration has stopped.
struct CgState {
  double rho[2];     // rho_k = (r_{k-1}, z_k), z_k = M r_{k-1}
  double res[2];     // relative residual after iteration k
  double normb;      // residuals are |r| / normb
  double initial;    // residual before the first iteration
  double res_tol, res_red, div_factor;
  double final_res;  // residual at the stop
  int iter;          // iterations done at the stop
  int done;          // 0 running, 1 stopped (converged or limit), 2 diverged
  int min_iter, max_iter;
};
// Grids of the two CG kernels (the direction kernel's is capped by
// cg_block_cap(), the update kernel's by kCgUpdateBlocks) and the sizes of
// their partial-sum buffers
constexpr int kCgBlocks = 4096;
constexpr int kCgUpdateBlocks = 1024;
int cg_direction_blocks(int n, int max_row);
int cg_update_blocks(int n);
// Iteration `it` (1-based), first kernel: stop test of iteration it - 1 from
// the update kernel's partials pu ([2][nbu]: |r|^2, (r, z)), beta, then
// p_new = z (it = 1) or beta p_old + z, q = A p_new (non-active rule) and the
// per-block partials pd of (p_new, q). No-op once st->done is set; launched
// with it = max_iter + 1 it only applies the final stop test.
void launch_cg_direction(int it, int n, int n_active, const int* rp, const int* col, const double* val,
                         int max_row, const double* z, const double* p_old, double* p_new, double* q,
                         const double* pu, int nbu, double* pd, CgState* st, hipStream_t s);
// Iteration `it`, second kernel: alpha = rho_it / (p, q) from the partials pd
// ([nbd]), x += alpha p, r -= alpha q, z = M r, partials of |r|^2 and (r, z)
// into pu.
void launch_cg_update(int it, int n, const double* p, const double* q, const double* diag, double* x,
                      double* r, double* z, const double* pd, int nbd, double* pu, CgState* st, hipStream_t s);

}  // namespace pmcd
