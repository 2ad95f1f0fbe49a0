// Time-dependent convection-diffusion on a device system:
// TimeConvectionDiffusion<d> (src/System/TimeConvectionDiffusion.C:259-452)
// with the schemes of TimeDiscretization (src/TimeDiscRoutines/
// TimeDiscretizations.C:117-284, 426-483): backward Euler, Crank-Nicolson and
// BDF2 (first step by backward Euler or Crank-Nicolson). Every step assembles
// the stiffness matrix A, the mass matrix M and the rhs f at the new time on
// every level, builds the scheme's right-hand side on the finest level and
// the system matrix factor A + M on every level, and solves with the
// stationary solver stack.
#include "system.hpp"
#include "elements.hpp"

namespace pmcd {

namespace {
// bdf_coefficients of TimeDiscretization (TimeDiscretizations.C:61)
constexpr double kBdf[3] = {4. / 3., -1. / 3., 2. / 3.};
}  // namespace

void CDSystem::tcd_assemble() {
  for(auto& Lp : levels_) {
    Level& L = *Lp;
    if(L.tcd_A.n != L.val.n) {
      L.tcd_A.alloc(L.val.n);
      L.tcd_M.alloc(L.val.n);
    }
    boundary_data(L);  // Dirichlet values at the current time
    AsmArgs a = asm_args(L);
    a.tcd = 1;
    a.local_M = local_M_.p;
    assemble_cells(a, stream_);
    launch_gather_matrix(L.n(), L.n_active(), a.nl, L.row_ptr.p, L.col.p, L.contrib_ptr.p, L.contrib.p, L.slot.p,
                         local_A_.p, L.tcd_A.p, stream_);
    launch_gather_matrix(L.n(), L.n_active(), a.nl, L.row_ptr.p, L.col.p, L.contrib_ptr.p, L.contrib.p, L.slot.p,
                         local_M_.p, L.tcd_M.p, stream_);
    launch_gather_rhs(L.n(), L.n_active(), L.contrib_ptr.p, L.contrib.p, local_rhs_.p, L.extra_ptr.p, L.extra.p,
                      L.dir_vals.p, L.rhs.p, stream_);
  }
}

void CDSystem::tcd_init(const TcdDesc& t) {
  PMCD_CHECK(!operator_only_, "an operator-only system has no mesh to assemble on");
  PMCD_CHECK(t.scheme >= 0 && t.scheme <= 2, "unknown time discretization ", t.scheme);
  PMCD_CHECK(t.bdf_first == 0 || t.bdf_first == 1, "unknown first BDF2 step ", t.bdf_first);
  PMCD_CHECK(t.tau > 0, "time step length must be positive");
  tcd_ = true;
  tcd_desc_ = t;
  tcd_t_ = t.t0;
  tcd_n_ = 0;
  set_example_time(desc_.ex, t.t0);
  if(local_M_.n != local_A_.n) local_M_.alloc(local_A_.n);
  tcd_assemble();
  // initial condition: the exact solution interpolated at the nodal points
  // (FEFunction::Interpolate of the example's initial condition)
  Level& F = finest();
  const int n = F.n();
  const ElementInfo& info = element_info(F.space.element);
  std::vector<double> u0(n, 0.0);
  for(int c = 0; c < F.mesh.n_cells; ++c)
    for(int i = 0; i < F.space.n_local; ++i) {
      double x[3], e[5];
      node_position(F.mesh, c, info, i, x);
      if(example_exact(desc_.ex, x[0], x[1], x[2], e)) u0[F.space.cell_dof(c)[i]] = e[0];
    }
  F.sol.upload(u0, stream_);
  tcd_u_old_.alloc(n);
  tcd_u_old2_.alloc(n);
  tcd_f_old_.alloc(n);
  if(kr_.n != (size_t)n) {
    kr_.alloc(n);
    kw_.alloc(n);
  }
  launch_copy(n, F.sol.p, tcd_u_old_.p, stream_);
  launch_copy(n, F.sol.p, tcd_u_old2_.p, stream_);
  launch_copy(n, F.rhs.p, tcd_f_old_.p, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

SolveStats CDSystem::tcd_step(const SolverDesc& s) {
  PMCD_CHECK(tcd_, "call tcd_init before tcd_step");
  ++tcd_n_;
  // scheme of this step: BDF2 starts with its first-step scheme
  const int sch = tcd_desc_.scheme;
  const int eff = sch == 2 ? (tcd_n_ <= 1 ? tcd_desc_.bdf_first : 2) : sch;
  const double tau = tcd_desc_.tau;
  tcd_t_ += tau;
  set_example_time(desc_.ex, tcd_t_);
  tcd_assemble();
  Level& F = finest();
  const int n = F.n(), na = F.n_active();
  double* f_new = kw_.p;  // the assembled rhs at the new time
  launch_copy(n, F.rhs.p, f_new, stream_);
  // right-hand side of the scheme (prepare_rhs_from_time_disc), built in the
  // solver's rhs vector
  double* r = F.rhs.p;
  auto mass_apply = [&](const double* x, double a) {
    launch_spmv(n, na, F.row_ptr.p, F.col.p, F.tcd_M.p, x, r, r, a, 1.0, stream_, F.st.max_row);
  };
  if(eff == 2) {
    launch_scale_inplace(na, kBdf[2] * tau, r, stream_);
    mass_apply(tcd_u_old_.p, kBdf[0]);
    mass_apply(tcd_u_old2_.p, kBdf[1]);
  } else if(eff == 0) {
    launch_scale_inplace(na, tau, r, stream_);
    mass_apply(tcd_u_old_.p, 1.0);
  } else {
    launch_scale_inplace(na, 0.5 * tau, r, stream_);
    launch_axpy(na, 0.5 * tau, tcd_f_old_.p, r, stream_);
    mass_apply(tcd_u_old_.p, 1.0);
    launch_spmv(n, na, F.row_ptr.p, F.col.p, F.tcd_A.p, tcd_u_old_.p, r, r, -0.5 * tau, 1.0, stream_,
                F.st.max_row);
  }
  // Dirichlet rows: the boundary values at the new time (copy_nonactive)
  launch_copy(n - na, f_new + na, r + na, stream_);
  // system matrix factor A + M on every level (prepare_system_matrix)
  const double factor = eff == 0 ? tau : (eff == 1 ? 0.5 * tau : kBdf[2] * tau);
  for(auto& Lp : levels_) {
    Level& L = *Lp;
    // a level whose sliced ELL copy exists gets it written here too
    const bool sell = L.sell_slots >= 0 && L.sell_val.n > 0 &&
                      L.sell_slot.n == (size_t)std::max(1, L.n_active()) && L.n_active() > 0;
    launch_tcd_system(L.n(), L.n_active(), L.row_ptr.p, L.tcd_A.p, L.tcd_M.p, factor, L.val.p, stream_,
                      sell ? L.sell_slot.p : nullptr, sell ? L.sell_off.p : nullptr, sell ? L.sell_val.p : nullptr);
    launch_gather_idx(L.n(), L.diag_pos.p, L.val.p, L.diag.p, stream_);
    L.sw_fresh = L.sw_fresh32 = L.sell_fresh32 = false;
    L.sell_fresh = sell;
  }
  coarse_ready_ = false;
  mixed_ready_ = false;
  // the iterate keeps the old solution, with the new Dirichlet values
  launch_copy(n - na, r + na, F.sol.p + na, stream_);
  // old rhs for Crank-Nicolson: the assembled rhs, saved before the solve
  // overwrites the work vector
  launch_copy(n, f_new, tcd_f_old_.p, stream_);
  SolveStats st = solve(s);
  launch_copy(n, tcd_u_old_.p, tcd_u_old2_.p, stream_);
  launch_copy(n, F.sol.p, tcd_u_old_.p, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  return st;
}

}  // namespace pmcd
