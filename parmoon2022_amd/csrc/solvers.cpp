// Iterative and direct solvers on the device-resident finest level.
//
// Reference: Solver<>::solve (src/Solver/Solver.C:601-696),
// IterativeMethod::converged (include/Solver/IterativeMethod.h:83-162),
// Iteration_cg::iterate (src/Solver/Iteration_cg.C:27-147),
// Iteration_gmres::flexible_gmres (src/Solver/Iteration_gmres.C:430-620) with
// the Givens/Hessenberg utilities (src/Solver/Utilities_gmres.C:6-108).
// Vectors stay on the device; only the scalars of the Krylov recurrences
// (dot products, Hessenberg column) come to the host, once per iteration.
#include "system.hpp"
#include <chrono>
#include <cmath>

namespace pmcd {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// IterativeMethod::converged
struct Convergence {
  const SolverDesc& sd;
  double initial = 0, previous = 0;
  explicit Convergence(const SolverDesc& s) : sd(s) {}
  bool operator()(double res, int it) {
    if(it == 0) { initial = res; previous = res; }
    const bool small = res < sd.res_tol;
    const bool reduced = res < sd.res_red * initial;
    const bool enough = it >= sd.min_iter;
    const bool stagnation = res > 1e10 * previous;
    const bool diverged = res > sd.div_factor * initial;
    bool conv = (small || reduced || stagnation || diverged) && enough;
    if(it == sd.max_iter) conv = true;
    if(diverged)
      PMCD_FAIL("Execution of the code stopped because of solver divergence: ", initial, " -> ", res,
                " at iteration ", it);
    previous = res;
    return conv;
  }
};

void givens_generate(double dx, double dy, double& cs, double& sn) {
  if(dy == 0.0) { cs = 1.0; sn = 0.0; }
  else if(std::abs(dy) > std::abs(dx)) {
    double t = dx / dy;
    sn = 1.0 / std::sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    double t = dy / dx;
    cs = 1.0 / std::sqrt(1.0 + t * t);
    sn = t * cs;
  }
}

void givens_apply(double& dx, double& dy, double cs, double sn) {
  double t = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = t;
}

}  // namespace

// Solver<>::update_matrix (Solver.C:423-501): factorize the coarse (or, for
// the direct solver, the only) level, and make FP32 copies of the level
// matrices for the mixed-precision V-cycle.
void CDSystem::prepare(const SolverDesc& s) {
  sd_ = s;
  rhs_nonactive_zero_ = false;
  if((s.precond == 2 && n_levels() > 1 && s.coarse_smoother == 0) || s.method == 0) {
    if(s.method == 0 && n_levels() > 1) PMCD_FAIL("the direct solver expects a single level");
    if(!coarse_ready_) coarse_factorize();
  }
  if(s.precond == 2 && n_levels() > 1 && s.coarse_smoother >= 2) prepare_sweep(*levels_.front());
  if(s.precond == 2 && s.smoother != 0)
    for(auto& Lp : levels_) prepare_sweep(*Lp);
  if(s.precond == 3 || s.precond == 4 || s.method == 4 || s.method == 5) prepare_sweep(finest());
  if(s.mixed)
    for(auto& Lp : levels_) {
      Level& L = *Lp;
      if(L.f_sol.n != (size_t)L.n()) {
        L.f_sol.alloc(L.n()); L.f_sol2.alloc(L.n()); L.f_rhs.alloc(L.n());
        L.f_def.alloc(L.n()); L.f_tmp.alloc(L.n());
      }
    }
  if(s.mixed && !mixed_ready_) {
    for(auto& Lp : levels_) {
      Level& L = *Lp;
      if(L.val32.n != L.val.n) L.val32.alloc(L.val.n);
      if(L.diag32.n != (size_t)L.n()) L.diag32.alloc(L.n());
      launch_d2f((int)L.val.n, L.val.p, L.val32.p, stream_);
      launch_d2f(L.n(), L.diag.p, L.diag32.p, stream_);
    }
    mixed_ready_ = true;
  }
  // sliced ELL and class-ordered SOR values are refreshed here, outside any
  // captured V-cycle: a replayed graph reads them but never rewrites them
  if(get_tuning(2) == 4 || get_tuning(3) == 4)
    for(auto& Lp : levels_)
      if(Lp->n_active() >= sell_min_rows()) prepare_level_sell(*Lp, s.mixed != 0, stream_);
  for(auto& Lp : levels_)
    if(!Lp->sweep_ptr.empty()) {
      refresh_sweep_values(*Lp, false, stream_);
      if(s.mixed && Lp->val32.n == Lp->val.n) refresh_sweep_values(*Lp, true, stream_);
    }
}

SolveStats CDSystem::solve(const SolverDesc& s) {
  prepare(s);
  const double t0 = now_s();
  SolveStats st;
  switch(s.method) {
    case 0: st = solve_direct(); break;
    case 1: st = solve_cg(); break;
    case 2: st = solve_fgmres(); break;
    case 3: st = solve_richardson(); break;
    case 4: st = solve_sor(0); break;
    case 5: st = solve_sor(2); break;
    default: PMCD_FAIL("unknown solver method ", s.method);
  }
  PMCD_HIP(hipStreamSynchronize(stream_));
  st.seconds = now_s() - t0;
  return st;
}

SolveStats CDSystem::solve_direct() {
  Level& F = finest();
  SolveStats st;
  if(direct_b_.n != (size_t)F.n()) direct_b_.alloc(F.n());
  launch_copy(F.n(), F.rhs.p, direct_b_.p, stream_);
  launch_gemv(F.n(), coarse_inv_.p, direct_b_.p, F.sol.p, stream_);
  st.converged = 1;
  return st;
}

SolveStats CDSystem::solve_richardson() {
  // x <- x + M^{-1} (b - A x)
  Level& F = finest();
  const int n = F.n();
  Convergence conv(sd_);
  if(kr_.n != (size_t)n) { kr_.alloc(n); kw_.alloc(n); }
  SolveStats st;
  launch_copy(n, F.rhs.p, kr_.p, stream_);
  apply_scaled_add(F, F.sol.p, kr_.p, -1.0);
  double res = std::sqrt(dot(n, kr_.p, kr_.p));
  rhs_nonactive_zero_ = false;
  st.initial_residual = res;
  int it = 0;
  if(!conv(res, 0)) {
    for(it = 1; it <= sd_.max_iter; ++it) {
      precond(kr_.p, kw_.p);
      launch_axpy(n, 1.0, kw_.p, F.sol.p, stream_);
      launch_copy(n, F.rhs.p, kr_.p, stream_);
      apply_scaled_add(F, F.sol.p, kr_.p, -1.0);
      res = std::sqrt(dot(n, kr_.p, kr_.p));
      if(conv(res, it)) break;
    }
  }
  st.iterations = std::min(it, sd_.max_iter);
  st.final_residual = res;
  st.converged = res < sd_.res_tol || res < sd_.res_red * st.initial_residual;
  return st;
}

// Iteration_sor::iterate (Iteration_sor.C:27-120): relative residual
// |b - A x| / |b|, one (symmetric) sweep per iteration on the finest level.
SolveStats CDSystem::solve_sor(int flag) {
  Level& F = finest();
  const int n = F.n();
  Convergence conv(sd_);
  if(kr_.n != (size_t)n) { kr_.alloc(n); kw_.alloc(n); }
  SolveStats st;
  launch_copy(n, F.rhs.p, kr_.p, stream_);
  apply_scaled_add(F, F.sol.p, kr_.p, -1.0);
  double normb = std::sqrt(dot(n, F.rhs.p, F.rhs.p));
  if(normb == 0.0) {
    normb = 1.0;
    launch_fill(n, 0.0, kr_.p, stream_);
    launch_fill(n, 0.0, F.sol.p, stream_);
  }
  double res = std::sqrt(dot(n, kr_.p, kr_.p)) / normb;
  st.initial_residual = res;
  int it = 0;
  if(!conv(res, 0)) {
    for(it = 1; it <= sd_.max_iter; ++it) {
      sor_sweep(F, F.rhs.p, F.sol.p, sd_.sor_omega, flag);
      launch_copy(n, F.rhs.p, kr_.p, stream_);
      apply_scaled_add(F, F.sol.p, kr_.p, -1.0);
      res = std::sqrt(dot(n, kr_.p, kr_.p)) / normb;
      if(conv(res, it)) break;
    }
  }
  st.iterations = std::min(it, sd_.max_iter);
  st.final_residual = res;
  st.converged = res < sd_.res_tol || res < sd_.res_red * st.initial_residual;
  return st;
}

SolveStats CDSystem::solve_cg() {
  Level& F = finest();
  const int n = F.n();
  Convergence conv(sd_);
  if(kr_.n != (size_t)n) { kr_.alloc(n); kw_.alloc(n); }
  DevBuf<double>& r = kr_;
  double* x = F.sol.p;
  SolveStats st;
  launch_copy(n, F.rhs.p, r.p, stream_);
  apply_scaled_add(F, x, r.p, -1.0);
  double normb = std::sqrt(dot(n, F.rhs.p, F.rhs.p));
  if(normb == 0.0) {
    normb = 1.0;
    launch_fill(n, 0.0, r.p, stream_);
    launch_fill(n, 0.0, x, stream_);
  }
  double resid = std::sqrt(dot(n, r.p, r.p)) / normb;
  if((sd_.precond == 0 || sd_.precond == 1) && cg_device()) {
    if(conv(resid, 0)) {
      st.initial_residual = st.final_residual = resid;
      st.converged = 1;
      return st;
    }
    return solve_cg_device(n, normb, resid);
  }
  if(cgh_p_.n != (size_t)n) { cgh_p_.alloc(n); cgh_z_.alloc(n); cgh_q_.alloc(n); }
  DevBuf<double>& p = cgh_p_;
  DevBuf<double>& z = cgh_z_;
  DevBuf<double>& q = cgh_q_;
  rhs_nonactive_zero_ = tail_is_zero(r.p);
  st.initial_residual = resid;
  if(conv(resid, 0)) { st.final_residual = resid; st.converged = 1; return st; }
  double rho_1 = 0.0;
  for(int i = 1; i <= sd_.max_iter; ++i) {
    precond(r.p, z.p);
    const double rho = dot(n, r.p, z.p);
    if(i == 1) launch_copy(n, z.p, p.p, stream_);
    else {
      const double beta = rho / rho_1;
      launch_scale_inplace(n, beta, p.p, stream_);
      launch_axpy(n, 1.0, z.p, p.p, stream_);
    }
    apply(F, p.p, q.p);
    const double alpha = rho / dot(n, p.p, q.p);
    launch_axpy(n, alpha, p.p, x, stream_);
    launch_axpy(n, -alpha, q.p, r.p, stream_);
    resid = std::sqrt(dot(n, r.p, r.p)) / normb;
    st.iterations = i;
    if(conv(resid, i)) break;
    rho_1 = rho;
  }
  st.final_residual = resid;
  st.converged = resid < sd_.res_tol || resid < sd_.res_red * st.initial_residual;
  return st;
}

// CG with no or Jacobi preconditioning, scalars on the device (k_krylov.hip):
// the residual r = b - A x is in kr_, x in the finest solution vector.
// Iterations are launched in chunks; the state after chunk k is read back
// while chunk k + 1 runs, and iterations launched after the stop do nothing.
SolveStats CDSystem::solve_cg_device(int n, double normb, double resid) {
  Level& F = finest();
  if(cg_z_.n != (size_t)n) {
    for(auto& b : cg_p_) b.alloc(n);
    cg_z_.alloc(n);
    cg_q_.alloc(n);
  }
  if(cg_pd_.n == 0) {
    cg_pd_.alloc(kCgBlocks);
    cg_pu_.alloc(2 * kCgUpdateBlocks);
  }
  if(cg_state_.n == 0) cg_state_.alloc(1);
  if(!cg_host_) PMCD_HIP(hipHostMalloc((void**)&cg_host_, 2 * sizeof(CgState), hipHostMallocDefault));
  const double* diag = sd_.precond == 1 ? F.diag.p : nullptr;
  if(diag) launch_div(n, kr_.p, diag, cg_z_.p, stream_);
  else launch_copy(n, kr_.p, cg_z_.p, stream_);
  CgState s0{};
  s0.rho[1] = dot(n, kr_.p, cg_z_.p);
  s0.res[0] = resid;
  s0.normb = normb;
  s0.initial = resid;
  s0.res_tol = sd_.res_tol;
  s0.res_red = sd_.res_red;
  s0.div_factor = sd_.div_factor;
  s0.min_iter = sd_.min_iter;
  s0.max_iter = sd_.max_iter;
  cg_host_[0] = s0;
  PMCD_HIP(hipMemcpyAsync(cg_state_.p, cg_host_, sizeof(CgState), hipMemcpyHostToDevice, stream_));
  PMCD_HIP(hipStreamSynchronize(stream_));
  const int nbd = cg_direction_blocks(n, F.st.max_row), nbu = cg_update_blocks(n);
  auto direction = [&](int it) {
    launch_cg_direction(it, n, F.n_active(), F.row_ptr.p, F.col.p, F.val.p, F.st.max_row, cg_z_.p,
                        cg_p_[(it - 1) & 1].p, cg_p_[it & 1].p, cg_q_.p, cg_pu_.p, nbu, cg_pd_.p, cg_state_.p,
                        stream_);
  };
  hipEvent_t ev[2];
  for(auto& e : ev) PMCD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int launched = 0, chunk = 8, k = 0;
  bool stopped = false;
  while(!stopped) {
    const int m = std::min(chunk, sd_.max_iter - launched);
    for(int j = 1; j <= m; ++j) {
      const int it = launched + j;
      direction(it);
      launch_cg_update(it, n, cg_p_[it & 1].p, cg_q_.p, diag, F.sol.p, kr_.p, cg_z_.p, cg_pd_.p, nbd, cg_pu_.p,
                       cg_state_.p, stream_);
    }
    launched += m;
    const bool last = launched >= sd_.max_iter;
    if(last) direction(sd_.max_iter + 1);  // the stopping test of the final iteration
    PMCD_HIP(hipMemcpyAsync(&cg_host_[k & 1], cg_state_.p, sizeof(CgState), hipMemcpyDeviceToHost, stream_));
    PMCD_HIP(hipEventRecord(ev[k & 1], stream_));
    if(last) {
      PMCD_HIP(hipEventSynchronize(ev[k & 1]));
      stopped = true;
    } else if(k > 0) {
      PMCD_HIP(hipEventSynchronize(ev[(k - 1) & 1]));
      stopped = cg_host_[(k - 1) & 1].done != 0;
    }
    ++k;
    chunk = std::min(2 * chunk, 32);
  }
  PMCD_HIP(hipStreamSynchronize(stream_));
  for(auto& e : ev) PMCD_HIP(hipEventDestroy(e));
  CgState fin;
  PMCD_HIP(hipMemcpy(&fin, cg_state_.p, sizeof(CgState), hipMemcpyDeviceToHost));
  PMCD_CHECK(fin.done != 0, "device CG ended without a stopping decision");
  if(fin.done == 2)
    PMCD_FAIL("Execution of the code stopped because of solver divergence: ", fin.initial, " -> ",
              fin.final_res, " at iteration ", fin.iter);
  SolveStats st;
  st.initial_residual = resid;
  st.iterations = fin.iter;
  st.final_residual = fin.final_res;
  st.converged = fin.final_res < sd_.res_tol || fin.final_res < sd_.res_red * resid;
  return st;
}

// Flexible GMRES(m) (Iteration_gmres.C:430-620) with modified Gram-Schmidt.
// Each iteration is one device chain: preconditioner, A z, the Gram-Schmidt
// steps (each folding the previous dot product's reduction in), the
// normalised next basis vector, then one copy of the Hessenberg column to
// pinned host memory. The host applies the Givens rotations and tests
// convergence. While the residual is far above the stopping threshold
// (fgmres_lookahead() times it, or below min_n_iterations) the next
// iteration's chain is queued before the host waits for this one: it needs
// only device data, and if the test stops the solver first, the queued
// iteration's results (a basis vector, a Hessenberg column beyond the last
// one used) are never read.
SolveStats CDSystem::solve_fgmres() {
  Level& F = finest();
  const int n = F.n();
  const int m = sd_.restart;
  Convergence conv(sd_);
  if(kv_.size() != (size_t)m + 1 || kr_.n != (size_t)n) {
    kv_.clear(); kz_.clear();
    for(int i = 0; i <= m; ++i) kv_.emplace_back(n);
    for(int i = 0; i < m; ++i) kz_.emplace_back(n);
    kr_.alloc(n);
    kw_.alloc(n);
    std::vector<const double*> zp(m);
    for(int i = 0; i < m; ++i) zp[i] = kz_[i].p;
    kz_ptr_.upload(zp, stream_);
  }
  if(kh_.n != (size_t)m + 2) kh_.alloc(m + 2);
  if(fg_y_.n != (size_t)m) fg_y_.alloc(m);
  if(fg_partial_.n != 2 * (size_t)kMaxReduceBlocks) fg_partial_.alloc(2 * kMaxReduceBlocks);
  if(fg_host_cap_ < m + 2) {
    if(fg_host_) PMCD_HIP(hipHostFree(fg_host_));
    fg_host_ = nullptr;
    PMCD_HIP(hipHostMalloc((void**)&fg_host_, sizeof(double) * 3 * (m + 2), hipHostMallocDefault));
    fg_host_cap_ = m + 2;
  }
  if(!fg_ev_[0])
    for(auto& e : fg_ev_) PMCD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  double* hcol[2] = {fg_host_, fg_host_ + (m + 2)};
  double* yhost = fg_host_ + 2 * (m + 2);
  double* P[2] = {fg_partial_.p, fg_partial_.p + kMaxReduceBlocks};
  std::vector<double> s(m + 1), cs(m + 1), sn(m + 1);
  // Hessenberg matrix, column-major (m+1) x m
  std::vector<double> H((size_t)(m + 1) * m, 0.0);
  auto h = [&](int i, int j) -> double& { return H[(size_t)j * (m + 1) + i]; };
  double* x = F.sol.p;
  double* r = kr_.p;
  SolveStats st;
  launch_copy(n, F.rhs.p, r, stream_);
  apply_scaled_add(F, x, r, -1.0);
  double resid = std::sqrt(dot(n, r, r));
  // every Krylov vector inherits the non-active part of r0 (A is the identity
  // there and the V-cycle returns the rhs's non-active part)
  rhs_nonactive_zero_ = tail_is_zero(r);
  double beta = resid;
  st.initial_residual = resid;
  if(conv(resid, 0)) { st.final_residual = resid; st.converged = 1; return st; }
  const double threshold = std::max(sd_.res_tol, sd_.res_red * st.initial_residual);
  const double lookahead = fgmres_lookahead();
  int j = 1;
  auto update = [&](int k) {
    // Hessenberg back-solve (Utilities_gmres.C:90-108), then x += sum y_j z_j
    // in one pass (the host buffer is free again once the copy is done: the
    // next use follows a synchronising reduction or the end of the solve)
    for(int i = 0; i <= k; ++i) yhost[i] = s[i];
    for(int i = k; i >= 0; --i) {
      yhost[i] /= h(i, i);
      for(int jj = i - 1; jj >= 0; --jj) yhost[jj] -= h(jj, i) * yhost[i];
    }
    PMCD_HIP(hipMemcpyAsync(fg_y_.p, yhost, sizeof(double) * (k + 1), hipMemcpyHostToDevice, stream_));
    launch_fgmres_update(n, k + 1, fg_y_.p, kz_ptr_.p, x, stream_);
    PMCD_HIP(hipStreamSynchronize(stream_));
  };
  // the device chain of iteration i of the current cycle (v_i given)
  auto enqueue = [&](int i) {
    precond(kv_[i].p, kz_[i].p);
    apply(F, kz_[i].p, r);
    // modified Gram-Schmidt: H(k,i) = <r, v_k>; r -= H(k,i) v_k
    launch_dot_partial(n, r, kv_[0].p, P[0], stream_);
    for(int k = 0; k <= i; ++k)
      launch_mgs_fused(n, P[k & 1], kh_.p + k, r, kv_[k].p, k < i ? kv_[k + 1].p : nullptr, P[(k + 1) & 1],
                       stream_);
    launch_scale_norm(n, P[(i + 1) & 1], kh_.p + i + 1, r, kv_[i + 1].p, stream_);
    PMCD_HIP(hipMemcpyAsync(hcol[i & 1], kh_.p, sizeof(double) * (i + 2), hipMemcpyDeviceToHost, stream_));
    PMCD_HIP(hipEventRecord(fg_ev_[i & 1], stream_));
  };
  while(j <= sd_.max_iter) {
    launch_scale_copy(n, 1.0 / beta, r, kv_[0].p, stream_);
    std::fill(s.begin(), s.end(), 0.0);
    s[0] = beta;
    int queued = -1;
    double last = beta;
    for(int i = 0; i < m && j <= sd_.max_iter; ++i, ++j) {
      if(queued < i) { enqueue(i); queued = i; }
      if(lookahead > 0 && i + 1 < m && j + 1 <= sd_.max_iter &&
         (j + 1 <= sd_.min_iter || last > lookahead * threshold)) {
        enqueue(i + 1);
        queued = i + 1;
      }
      PMCD_HIP(hipEventSynchronize(fg_ev_[i & 1]));
      const double* hc = hcol[i & 1];
      for(int k = 0; k <= i; ++k) h(k, i) = hc[k];
      h(i + 1, i) = hc[i + 1];
      if(!(h(i + 1, i) != 0.0)) PMCD_FAIL("Unhappy breakdown in flexible gmres at iteration ", j);
      for(int k = 0; k < i; ++k) givens_apply(h(k, i), h(k + 1, i), cs[k], sn[k]);
      givens_generate(h(i, i), h(i + 1, i), cs[i], sn[i]);
      givens_apply(h(i, i), h(i + 1, i), cs[i], sn[i]);
      givens_apply(s[i], s[i + 1], cs[i], sn[i]);
      resid = std::abs(s[i + 1]);
      last = resid;
      st.iterations = j;
      if(conv(resid, j)) {
        update(i);
        st.final_residual = resid;
        st.converged = resid < sd_.res_tol || resid < sd_.res_red * st.initial_residual;
        return st;
      }
    }
    update(m - 1);
    launch_copy(n, F.rhs.p, r, stream_);
    apply_scaled_add(F, x, r, -1.0);
    beta = std::sqrt(dot(n, r, r));
    if(conv(resid, j)) break;
  }
  st.final_residual = resid;
  st.converged = resid < sd_.res_tol || resid < sd_.res_red * st.initial_residual;
  return st;
}

void CDSystem::errors(double out[5]) {
  PMCD_CHECK(!operator_only_, "an operator-only system has no mesh for the error norms");
  Level& F = finest();
  AsmArgs a = asm_args(F);
  launch_cell_errors(a, F.dof.p, F.sol.p, cell_err_.p, stream_);
  launch_sum_max(F.mesh.n_cells, 5, cell_err_.p, red_partial_.p, red_out_.p, stream_);
  double e[5];
  PMCD_HIP(hipMemcpyAsync(e, red_out_.p, sizeof(e), hipMemcpyDeviceToHost, stream_));
  PMCD_HIP(hipStreamSynchronize(stream_));
  for(int k = 0; k < 4; ++k) out[k] = std::sqrt(e[k]);
  out[4] = e[4];
}

}  // namespace pmcd
