// Many independent convection-diffusion instances (different meshes and
// coefficients, same element and quadrature) assembled and solved together.
//
// The reference handles one ConvectionDiffusion<d> at a time
// (src/System/ConvectionDiffusion.C:112-452); a batch is a set of them. On
// the device the instances form one block-diagonal system: every instance's
// active rows come first (instance after instance), then every instance's
// Dirichlet rows, so the non-active rows are again a contiguous tail and the
// single-system kernels (assembly, row gather, SpMV, SOR classes) run over
// all instances in one launch each. Inside an instance the row and column
// order is the instance's own, so each row is summed exactly as in a
// single-instance run. Per-instance reductions (FGMRES) use segmented
// kernels over fixed row chunks.
#pragma once
#include <memory>
#include <vector>
#include "kernels.hpp"
#include "system.hpp"

namespace pmcd {

class BatchSystem {
 public:
  BatchSystem(const std::vector<ProblemDesc>& descs, hipStream_t stream);
  ~BatchSystem();

  void assemble();
  // flexible GMRES (restart sd.restart) on every instance, preconditioned by
  // none / Jacobi / SOR / SSOR (FP32 sweeps with sd.mixed); per-instance stats
  std::vector<SolveStats> solve(const SolverDesc& sd);

  int n_inst() const { return (int)inst_.size(); }
  int n_dof() const { return n_; }
  int n_active() const { return na_; }
  long long nnz() const { return (long long)G_.st.nnz; }
  long long n_cells() const { return n_cells_; }
  int inst_n_dof(int i) const { return inst_[i].L->n(); }
  int inst_n_active(int i) const { return inst_[i].L->n_active(); }
  long long inst_nnz(int i) const { return (long long)inst_[i].L->st.nnz; }
  // instance-local copies (host)
  void get_solution(int i, double* out);
  void set_solution(int i, const double* in);
  void get_rhs(int i, double* out);
  void get_matrix(int i, double* out);
  double last_assemble_ms = 0;

 private:
  struct Inst {
    std::unique_ptr<Level> L;  // host-side mesh, space and structure
    int act_off = 0, dir_off = 0, cell_off = 0;
  };
  int grow(int i, int r) const {  // global row of instance i's local row r
    const Inst& I = inst_[i];
    const int na = I.L->n_active();
    return r < na ? I.act_off + r : na_ + I.dir_off + (r - na);
  }
  void to_global(int i, const double* local, double* global_host);
  void precond(const double* v, double* z);
  void spmv(const double* x, const double* z, double* y, double alpha, double beta);

  hipStream_t stream_;
  std::vector<Inst> inst_;
  std::vector<ProblemDesc> descs_;
  int n_ = 0, na_ = 0;
  long long n_cells_ = 0;
  Level G_;  // the global block-diagonal system
  QuadRule rule_;
  int ncomp_ = 0;
  DevBuf<double> ref_, qw_, qp_, tensor_, local_A_, local_rhs_;
  DevBuf<ExampleParams> cell_ex_;
  DevBuf<int> cell_inst_;
  // segmented reductions
  std::vector<SegChunk> chunks_host_;
  DevBuf<SegChunk> chunks_;
  DevBuf<int> inst_ptr_;
  DevBuf<double> partial_;
  // solver work
  SolverDesc sd_{};
  std::vector<DevBuf<double>> kv_, kz_;
  DevBuf<double> kr_, kw_, kh_;
  // converged instances are masked out of the matrix kernels
  DevBuf<int> row_inst_;
  DevBuf<uint8_t> active_;
  std::vector<uint8_t> active_host_;
  DevBuf<float> f_v_, f_z_;
  bool mixed_ready_ = false;
  // device-resident FGMRES state: per-instance Hessenberg matrices, Givens
  // rotations, rhs of the least-squares problem, update coefficients, scale
  // factors, instance states, a breakdown flag, ping-pong reduction partials
  // and the flexible basis' device pointers
  DevBuf<double> bH_, bs_, bcs_, bsn_, by_, bscale_, bpart_;
  DevBuf<BatchInstState> bst_;
  DevBuf<int> bbreak_;
  DevBuf<const double*> kz_ptr_;
  // pinned host copies of the active flags and breakdown flag (two slots,
  // polled one iteration behind), and their events
  uint8_t* poll_host_ = nullptr;
  hipEvent_t poll_ev_[2] = {nullptr, nullptr};
  // one captured graph per iteration index of a restart cycle (the part of
  // an iteration after a cooperative sweep, or the whole iteration)
  std::vector<hipGraphExec_t> graphs_;
  long long graph_key_ = -1;
  void clear_graphs();
  // one preconditioner sweep on the current sweep classes; true if it ran as
  // one cooperative launch
  bool coop_sweep(const void* b, void* x, bool f32, int flag);
};

}  // namespace pmcd
