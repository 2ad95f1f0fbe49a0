// Device-resident convection-diffusion system: mesh hierarchy, FE spaces,
// CSR matrices and right-hand sides per multigrid level, assembly, the
// iterative solvers (CG, flexible GMRES) with Jacobi or geometric multigrid
// preconditioning (FP64 or FP32 V-cycle), and error evaluation.
//
// Reference: ConvectionDiffusion<d> (src/System/ConvectionDiffusion.C:72-459)
// owning one SystemPerGrid per level (finest first there, coarsest first
// here), Solver<> (src/Solver/Solver.C:423-696), Iteration_cg
// (Iteration_cg.C:27-147), Iteration_gmres::flexible_gmres
// (Iteration_gmres.C:430-620), Multigrid (src/Multigrid/Multigrid.C:157-410).
#pragma once
#include <array>
#include <memory>
#include "common.hpp"
#include "mesh.hpp"
#include "fespace.hpp"
#include "structure.hpp"
#include "quadrature.hpp"
#include "elements.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace pmcd {

struct ProblemDesc {
  int geometry;       // DefaultGeometry, 10 = structured hexa cube, 11 = structured tetra cube, 12 = mesh file
  int n_refine;       // regular refinement steps of the default mesh
  int structured_n;   // cells per edge for the structured cubes
  int order;          // ansatz order 1 or 2
  int supg;           // 0 Galerkin, 1 SUPG
  int linear_quad;    // 1: quadrature degree 2p (INTERNAL_PROBLEM_LINEAR), else 3p
  int mg_levels;      // number of levels kept (1 = finest only)
  ExampleParams ex;
  std::string mesh_file;  // geometry 12: path of a 3D .mesh file
  // TDomain::RefineByIndicator steps without conforming closure after the
  // regular refinements (2D, P1/Q1, one level): leaves with hanging vertices
  int adapt_steps = 0;
  double (*indicator)(const double* x, void* ctx) = nullptr;  // evaluated at vertices
  void* indicator_ctx = nullptr;
  int adapt_conforming = 0;  // 1: conforming closure (triangles), no hanging vertices
};

struct SolverDesc {
  int method;          // 0 direct, 1 cg, 2 fgmres, 3 richardson, 4 sor, 5 ssor
  int precond;         // 0 none, 1 jacobi, 2 multigrid, 3 sor, 4 ssor
  int max_iter;
  int min_iter;
  double res_tol;      // residual_tolerance
  double res_red;      // residual_reduction
  double div_factor;   // divergence_factor
  int restart;         // gmres_restart
  int n_pre, n_post;   // multigrid smoothing steps
  double jacobi_damp;  // jacobi_damp
  double mg_damp;      // multigrid_correction_damp_factor
  int mixed;           // 1: V-cycle in FP32
  int smoother;        // multigrid smoother: 0 jacobi, 1 sor (forward), 2 ssor
  double sor_omega;    // sor_omega
  int cycle;           // multigrid_cycle_type: 0 V, 1 W, 2 F
  int coarse_smoother; // multigrid_smoother_coarse: 0 direct solve, 1 jacobi, 2 sor, 3 ssor
  double coarse_residual;  // multigrid_coarse_residual
  int coarse_max_iter;     // multigrid_coarse_max_n_iterations
};

// Time discretization (TimeDiscretization, src/TimeDiscRoutines/
// TimeDiscretizations.C:26-493)
struct TcdDesc {
  int scheme;     // time_discretization: 0 backward_euler, 1 crank_nicolson, 2 bdf_two
  int bdf_first;  // bdf_two_first_step: 0 backward_euler, 1 crank_nicolson
  double tau;     // time_step_length
  double t0;      // time_start
};

struct SolveStats {
  int iterations = 0;
  double initial_residual = 0;
  double final_residual = 0;
  int converged = 0;
  double seconds = 0;
};

// A transfer matrix (the prolongation) in the sliced ELL layout of its active rows,
// with the column footprints of its 256-row blocks (x staged in LDS when the
// launch mode asks for it); slots -1: not built
struct SellMatrix {
  int slots = -1;
  int fp_max = 0;
  DevBuf<long long> off;
  DevBuf<int> perm, len, col, fp_off, fp;
  DevBuf<unsigned short> lcol;
  DevBuf<double> val;
  DevBuf<float> val32;
};

struct Level {
  Mesh mesh;
  FESpace space;
  Structure st;
  std::vector<double> dirichlet_host;
  // device data
  DevBuf<double> cell_xyz;
  DevBuf<int> dof, row_ptr, col, diag_pos, contrib_ptr, contrib;
  DevBuf<uint8_t> slot;
  DevBuf<double> dir_vals, extra;
  DevBuf<int> extra_ptr;
  DevBuf<double> val, rhs, sol, diag;
  DevBuf<float> val32, diag32;
  // multigrid work vectors
  DevBuf<double> mg_sol, mg_sol2, mg_rhs, mg_def, mg_tmp;
  DevBuf<float> f_sol, f_sol2, f_rhs, f_def, f_tmp;
  // transfer from the next coarser level: P (n x n_coarse), R = P^T
  DevBuf<int> p_rp, p_col, r_rp, r_col;
  DevBuf<double> p_val, r_val;
  DevBuf<float> p_val32, r_val32;
  int n_coarse = 0;
  int p_max_row = -1, r_max_row = -1;  // longest rows of P and R
  // P's active fine rows in the sliced ELL layout
  SellMatrix p_sell;
  // SOR sweep classes (built on first use): class k = sweep_rows[sweep_ptr[k] ..
  // sweep_ptr[k+1]); exact = level schedule (bit-identical to the serial sweep)
  std::vector<int> sweep_ptr;
  DevBuf<int> sweep_ptr_dev;  // the same class bounds on the device
  int sweep_max_class = 0;    // rows of the largest class
  DevBuf<int> sweep_rows;
  bool sweep_exact = false;
  // multicolour sweeps read a copy of the matrix with the rows in class order
  // (each row's entries in their original order), refreshed after assembly
  DevBuf<int> sw_rp, sw_col;
  DevBuf<double> sw_val;
  DevBuf<float> sw_val32;
  bool sw_fresh = false, sw_fresh32 = false;
  // time-dependent problems: stiffness and mass matrices (the system matrix
  // tau_factor A + M goes to val)
  DevBuf<double> tcd_A, tcd_M;
  // sliced ELL copy of the active rows (built on first use; values refreshed
  // after assembly)
  int sell_slots = -1;  // -1: not built
  int sell_window = 0;  // rows per length-sorted window of the built layout
  DevBuf<long long> sell_off;
  DevBuf<int> sell_perm, sell_len, sell_col, sell_slot;  // sell_slot: row -> slot
  DevBuf<double> sell_val;
  DevBuf<float> sell_val32;
  bool sell_fresh = false, sell_fresh32 = false;
  // column footprints of the layout's 256-slot blocks (sliced ELL mode bit
  // 4): -1 not built, 0 not available (window not 256, or a block too wide)
  int sell_fp_max = -1;
  DevBuf<unsigned short> sell_lcol;
  DevBuf<int> sell_fp_off, sell_fp;
  // hanging-DOF coupling after assembly: launch k takes matrix items
  // hg_mat_launch[k] .. hg_mat_launch[k+1] and rhs items hg_rhs_launch[k] ..;
  // hg_fix_* set the hanging rows' constraint entries
  std::vector<int> hg_mat_launch, hg_rhs_launch;
  DevBuf<int> hg_mat_dst, hg_mat_sp, hg_mat_src, hg_rhs_dst, hg_rhs_sp, hg_rhs_src, hg_fix_pos;
  DevBuf<double> hg_mat_w, hg_rhs_w, hg_fix_val;
  int n() const { return space.n_dof; }
  int n_active() const { return space.n_active; }
};

// Quadrature rule of the cells and the device reference tables of the
// element: basis values and derivatives at the rule's points (ncomp
// components), weights, points, and (if `tensors`) the reference tensors of
// the tensor assembly.
QuadRule upload_reference_tables(Shape shape, int order, bool linear_quad, bool tensors, int& ncomp,
                                 DevBuf<double>& ref, DevBuf<double>& qw, DevBuf<double>& qp,
                                 DevBuf<double>& tensor, hipStream_t stream);

// Host-side setup only (meshes, FE spaces, sparsity, slot maps): no GPU use.
std::vector<std::unique_ptr<Level>> build_levels(const ProblemDesc& d);

// Hanging-DOF coupling of a level whose space has hanging DOFs
// (Assemble2D.C:3350-3427): hanging_plan uploads the gather lists (once;
// every column of a hanging row must be in the rows of its active coupling
// DOFs), apply_hanging couples the assembled val and rhs and writes the
// constraint rows and zero hanging rhs entries.
void hanging_plan(Level& L, hipStream_t stream);
void apply_hanging(Level& L, hipStream_t stream);

// SOR sweep of a level's matrix in place (TMatrix::sor_sweep, flag 0 forward,
// 1 backward, 2 both), on the level's row classes (built on first use)
void prepare_level_sweep(Level& L, hipStream_t stream);
// the class-ordered (multicolour) copies of val / val32 brought up to date
void refresh_sweep_values(Level& L, bool f32, hipStream_t stream);
// (row_owner, owner_active): rows of inactive owners are skipped (may be null)
void level_sor_sweep(Level& L, const double* b, double* x, double omega, int flag, hipStream_t stream,
                     const int* row_owner = nullptr, const uint8_t* owner_active = nullptr);
void level_sor_sweep32(Level& L, const float* b, float* x, float omega, int flag, hipStream_t stream,
                       const int* row_owner = nullptr, const uint8_t* owner_active = nullptr);
// Matrix kernels of a level: the sliced ELL kernels where the SpMV / Jacobi
// variant is 4 (layout built and values refreshed on first use), the CSR
// kernels otherwise. y = beta z + alpha A x with the non-active rule;
// x_out = x + omega D^{-1} (b - A x).
void level_spmv(Level& L, const double* x, const double* z, double* y, double alpha, double beta,
                hipStream_t stream);
void level_spmv32(Level& L, const float* x, const float* z, float* y, float alpha, float beta, hipStream_t stream);
void level_jacobi(Level& L, const double* b, const double* x, double* x_out, double omega, hipStream_t stream);
void level_jacobi32(Level& L, const float* b, const float* x, float* x_out, float omega, hipStream_t stream);
// builds the sliced ELL layout and refreshes its FP64 (and, with f32, FP32)
// values; called before any graph capture
void prepare_level_sell(Level& L, bool f32, hipStream_t stream);
// builds the sliced ELL layout (positions and column indices) only
void build_level_sell(Level& L, hipStream_t stream);
// whether the SpMV or Jacobi kernels of the level run on the sliced ELL copy
bool level_uses_sell(const Level& L);
// Directions of one multigrid cycle as CycleControl builds them
// (src/Multigrid/CycleControl.C:17-111): a dummy first entry, then for each
// visit of a level with recursion count r, r times "down" followed by the
// coarser level's directions, then "up" ("end" on the finest level); the
// coarsest level contributes one "up". Recursion counts: 1 (V) or 2 (W, F)
// on every level but the finest (1); an F cycle resets a level's count to 1
// after its first visit. type: 0 V, 1 W, 2 F.
enum { kMgDown = 0, kMgUp = 1, kMgEnd = 2, kMgDummy = 3 };
std::vector<int> cycle_directions(int type, int n_levels);
// physical position of the nodal point of local DOF i of cell c
void node_position(const Mesh& m, int c, const ElementInfo& info, int i, double* x);
double now_seconds();

// One level of a caller-owned operator hierarchy (coarsest first): the CSR
// matrix, whose rows n_active.. are the identity rows of the Dirichlet DOFs
// (their entries are not read), and, for every level but the coarsest, the
// prolongation from the next coarser level (n x n_coarse CSR).
struct OperatorLevel {
  int n = 0, n_active = 0;
  std::vector<int> row_ptr, col;
  std::vector<double> val;
  int n_coarse = 0;
  std::vector<int> p_row_ptr, p_col;
  std::vector<double> p_val;
};

class CDSystem {
 public:
  CDSystem(const ProblemDesc& d, hipStream_t stream);
  // operator only (Solver<> on caller-owned matrices): no mesh, no assembly
  CDSystem(const std::vector<OperatorLevel>& levels, hipStream_t stream);
  bool operator_only() const { return operator_only_; }
  // new matrix values of one level (same pattern), from host memory
  void set_level_values(int level, const double* val);
  // right-hand side of the finest level, from host memory
  void set_rhs(const double* b);
  ~CDSystem();
  CDSystem(const CDSystem&) = delete;
  CDSystem& operator=(const CDSystem&) = delete;

  void assemble();                        // every level
  SolveStats solve(const SolverDesc& s);  // finest level
  void errors(double out[5]);             // L2, H1-semi, SD, DG, Linf on the finest level

  int n_levels() const { return (int)levels_.size(); }
  Level& level(int l) { return *levels_[l]; }
  Level& finest() { return *levels_.back(); }
  const ProblemDesc& desc() const { return desc_; }
  hipStream_t stream() const { return stream_; }
  QuadRule cell_rule() const { return rule_; }
  double last_assemble_ms = 0;

  // building blocks (used by the solvers and exposed through the C ABI)
  void apply(Level& L, const double* x, double* y);                 // y = A x
  void apply_scaled_add(Level& L, const double* x, double* y, double a);  // y += a A x
  double dot(int n, const double* x, const double* y);             // synchronous
  void vcycle(const double* rhs, double* out);                     // one V, W or F cycle, FP64 or FP32
  void set_solver(const SolverDesc& s) { sd_ = s; }
  void prepare(const SolverDesc& s);      // coarse factorization, FP32 copies
  void time_kernel(int which, int level, int iters, double* avg_ms);
  // the V-cycle's transfer between `level` and the next coarser one on host
  // arrays: which 0 = prolongation (y += a P x), 1 = restriction (y = R x)
  void apply_transfer(int level, int which, const double* x, double* y, double a);
  // TMatrix::sor_sweep on level L, in place: flag 0 forward, 1 backward, 2 both
  void sor_sweep(Level& L, const double* b, double* x, double omega, int flag);
  void sor_sweep32(Level& L, const float* b, float* x, float omega, int flag);
  void prepare_sweep(Level& L);
  // same mesh, new coefficients / source / boundary values
  void set_example(const ExampleParams& ex);

  // time-dependent problem (TimeConvectionDiffusion<d>,
  // src/System/TimeConvectionDiffusion.C): tcd_init assembles the stiffness
  // and mass matrices and the rhs at t0 and sets the initial condition (the
  // exact solution interpolated at the nodes); tcd_step advances one step of
  // length tau: assemble at the new time, the right-hand side and system
  // matrix of the scheme, solve.
  void tcd_init(const TcdDesc& t);
  SolveStats tcd_step(const SolverDesc& s);
  double tcd_time() const { return tcd_t_; }
  int tcd_n_steps() const { return tcd_n_; }

  // algebraic flux correction (ConvectionDiffusion_AFC<d>,
  // src/System/ConvectionDiffusion_AFC.C): assemble the Galerkin matrix
  // with its Dirichlet rows, add the artificial diffusion D, and iterate the
  // fixed point rhs or fixed point matrix scheme with the Kuzmin or BJK17
  // limiter and dynamic damping from the zero initial iterate. prm: epsilon, maxit, damping factor, its max,
  // min, global max, increase, decrease, max increase, max decrease, min
  // tolerance, limiter (0 Kuzmin, 1 BJK17), scheme (0 fixed point rhs,
  // 1 fixed point matrix), fixed point matrix weight. out: iterations,
  // final residual, rejected steps, initial residual. Single level only.
  SolveStats afc_solve(const SolverDesc& s, const double* prm, double* out);

 private:
  void build_level_device(Level& L);
  void boundary_data(Level& L);
  void build_transfer(Level& fine, const Level& coarse);
  void set_transfer(Level& fine, int n_coarse, const std::vector<int>& rp, const std::vector<int>& col,
                    const std::vector<double>& val);
  // fine += a P coarse on the active fine rows; coarse = R fine on the active
  // coarse rows, 0 on its Dirichlet tail (the V-cycle's grid transfers)
  void prolong_add(Level& fine, const double* coarse, double* y, double a);
  void restrict_to(Level& fine, const Level& coarse, const double* x, double* y);
  void prolong_add32(Level& fine, const float* coarse, float* y, float a);
  void restrict_to32(Level& fine, const Level& coarse, const float* x, float* y);
  void assemble_level(Level& L);
  void coarse_factorize();
  // `kind`: 0 Jacobi, 1 SOR, 2 SSOR; -1 the solver's multigrid smoother
  int smooth(Level& L, double* const x[2], int cur, const double* b, int steps, bool from_zero, int kind = -1);
  // the coarsest level's iterative loop (coarse smoother instead of the direct solve)
  int coarse_iterate(const double* b, double* const x[2], int cur);
  int smooth32(Level& L, float* const x[2], int cur, const float* b, int steps, bool from_zero);
  bool tail_is_zero(const double* v);  // v vanishes on the finest non-active rows
  void vcycle64(const double* rhs, double* out);
  void vcycle32(const double* rhs, double* out);
  void vcycle_eager(const double* rhs, double* out);
  void clear_graphs();
  void tcd_assemble();  // stiffness, mass and rhs at desc_.ex.t on every level
  void precond(const double* v, double* z);
  SolveStats solve_cg();
  SolveStats solve_cg_device(int n, double normb, double resid);
  SolveStats solve_fgmres();
  SolveStats solve_direct();
  SolveStats solve_richardson();
  SolveStats solve_sor(int flag);

  ProblemDesc desc_;
  bool operator_only_ = false;
  SolverDesc sd_{};
  hipStream_t stream_;
  std::vector<std::unique_ptr<Level>> levels_;  // coarsest first
  QuadRule rule_;
  int ncomp_ = 0;
  DevBuf<double> ref_, qw_, qp_, tensor_;
  DevBuf<double> local_A_, local_rhs_;
  DevBuf<double> red_partial_, red_out_, cell_err_;
  // dense inverse of the coarsest matrix (direct coarse solve)
  DevBuf<double> coarse_inv_, coarse_tmp_;
  bool coarse_ready_ = false;
  bool mixed_ready_ = false;
  // set by the Krylov solvers when every preconditioner input vanishes on the
  // non-active rows (true when the initial residual does)
  bool rhs_nonactive_zero_ = false;
  // captured V-cycles, one per (input, output, zero-tail) triple, valid for
  // the solver parameters and launch shapes they were captured with
  struct VGraph {
    const double* rhs;
    double* out;
    bool fine_zero;
    hipGraphExec_t exec;
  };
  std::vector<VGraph> vgraphs_;
  SolverDesc graph_sd_{};
  int graph_epoch_ = -1;
  // time-dependent state
  bool tcd_ = false;
  TcdDesc tcd_desc_{};
  double tcd_t_ = 0;
  int tcd_n_ = 0;
  DevBuf<double> local_M_, tcd_f_old_, tcd_u_old_, tcd_u_old2_;
  // algebraic flux correction state
  DevBuf<int> afc_tpos_;
  DevBuf<double> afc_A_, afc_D_, afc_E_, afc_alpha_, afc_f_, afc_Rp_, afc_Rm_, afc_old_, afc_cur_sol_,
      afc_cur_rhs_, afc_res_, afc_gamma_;
  DevBuf<double> afc_Pp_, afc_Pm_, afc_Qp_, afc_Qm_, afc_umax_, afc_umin_, afc_q_;  // Newton: limiter sums
  DevBuf<double> afc_hX_, afc_hR_, afc_M_, afc_w_;  // Anderson: outputs, updates, QR work space
  bool afc_anderson_alpha(int n, int k, const double* const* R, double* alpha);
  // Krylov work space
  std::vector<DevBuf<double>> kv_, kz_;
  DevBuf<double> kr_, kw_, kh_;
  // FGMRES: device pointers to the flexible basis, the update coefficients,
  // ping-pong reduction partials, pinned host Hessenberg columns (two) and
  // update coefficients, and the events of the two columns' copies
  DevBuf<const double*> kz_ptr_;
  DevBuf<double> fg_y_, fg_partial_;
  double* fg_host_ = nullptr;
  int fg_host_cap_ = 0;
  hipEvent_t fg_ev_[2] = {nullptr, nullptr};
  // CG with a host-side loop (multigrid or SOR preconditioning)
  DevBuf<double> cgh_p_, cgh_z_, cgh_q_;
  // coarse inverse and direct solve work space
  DevBuf<double> coarse_dense_, coarse_colbuf_, direct_b_;
  DevBuf<int> coarse_perm_, coarse_colmap_;
  // device-resident CG: direction ping-pong, state, partial sums and a
  // pinned host copy of the state
  DevBuf<double> cg_p_[2], cg_z_, cg_q_, cg_pd_, cg_pu_;
  DevBuf<CgState> cg_state_;
  CgState* cg_host_ = nullptr;
  AsmArgs asm_args(const Level& L);
};

}  // namespace pmcd
