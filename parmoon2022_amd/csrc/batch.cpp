#include "batch.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace pmcd {

namespace {

// IterativeMethod::converged (include/Solver/IterativeMethod.h:83-162) for one
// instance; a diverging instance is reported, not thrown, so that the others
// finish
struct InstConv {
  double initial = 0, previous = 0;
  bool diverged = false;
  bool operator()(const SolverDesc& sd, double res, int it) {
    if(it == 0) { initial = res; previous = res; }
    const bool small = res < sd.res_tol;
    const bool reduced = res < sd.res_red * initial;
    const bool enough = it >= sd.min_iter;
    const bool stagnation = res > 1e10 * previous;
    const bool div = res > sd.div_factor * initial;
    bool conv = (small || reduced || stagnation || div) && enough;
    if(it == sd.max_iter) conv = true;
    if(div) diverged = true;
    previous = res;
    return conv;
  }
};

void givens_gen(double dx, double dy, double& cs, double& sn) {
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

void givens_app(double& dx, double& dy, double cs, double sn) {
  double t = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = t;
}

}  // namespace

BatchSystem::BatchSystem(const std::vector<ProblemDesc>& descs, hipStream_t stream)
    : stream_(stream), descs_(descs) {
  PMCD_CHECK(!descs.empty(), "a batch needs at least one instance");
  const ProblemDesc& d0 = descs[0];
  for(const auto& d : descs) {
    PMCD_CHECK(d.order == d0.order && d.supg == d0.supg && d.linear_quad == d0.linear_quad,
               "instances of a batch share ansatz order, discretisation and quadrature");
    PMCD_CHECK(d.mg_levels <= 1, "batched instances are single-level (no multigrid hierarchy)");
  }
  // host set-up per instance
  inst_.resize(descs.size());
  for(size_t i = 0; i < descs.size(); ++i) {
    ProblemDesc d = descs[i];
    d.mg_levels = 1;
    auto lv = build_levels(d);
    inst_[i].L = std::move(lv.back());
  }
  const Shape shape = inst_[0].L->mesh.shape;
  for(auto& I : inst_)
    PMCD_CHECK(I.L->mesh.shape == shape, "instances of a batch share the cell shape");
  // global numbering: all active rows (instance by instance), then all
  // Dirichlet rows
  int act = 0, dir = 0, cells = 0;
  for(auto& I : inst_) {
    I.act_off = act;
    I.dir_off = dir;
    I.cell_off = cells;
    act += I.L->n_active();
    dir += I.L->n() - I.L->n_active();
    cells += I.L->mesh.n_cells;
  }
  na_ = act;
  n_ = act + dir;
  n_cells_ = cells;
  const int nl = inst_[0].L->space.n_local;
  // global structure
  Structure& S = G_.st;
  S.n_rows = n_;
  S.row_ptr.assign(n_ + 1, 0);
  S.contrib_ptr.assign(n_ + 1, 0);
  for(int i = 0; i < n_inst(); ++i) {
    const Structure& s = inst_[i].L->st;
    for(int r = 0; r < s.n_rows; ++r) {
      S.row_ptr[grow(i, r) + 1] = s.row_ptr[r + 1] - s.row_ptr[r];
      S.contrib_ptr[grow(i, r) + 1] = s.contrib_ptr[r + 1] - s.contrib_ptr[r];
    }
    S.max_row = std::max(S.max_row, s.max_row);
  }
  for(int r = 0; r < n_; ++r) {
    S.row_ptr[r + 1] += S.row_ptr[r];
    S.contrib_ptr[r + 1] += S.contrib_ptr[r];
  }
  S.nnz = S.row_ptr[n_];
  PMCD_CHECK(S.nnz < ((int64_t)1 << 31), "batch too large for 32-bit CSR offsets");
  S.col.resize(S.nnz);
  S.diag.resize(n_);
  S.contrib.resize(S.contrib_ptr[n_]);
  S.slot.resize((size_t)S.contrib_ptr[n_] * nl);
  std::vector<int> gdof((size_t)cells * nl);
  std::vector<double> xyz;
  std::vector<double> dirv(n_ - na_, 0.0);
  std::vector<ExampleParams> exs(n_inst());
  std::vector<int> cinst(cells);
  for(int i = 0; i < n_inst(); ++i) {
    const Inst& I = inst_[i];
    const Level& L = *I.L;
    const Structure& s = L.st;
    const Mesh& m = L.mesh;
    for(int r = 0; r < s.n_rows; ++r) {
      const int g = grow(i, r);
      const int64_t o = S.row_ptr[g];
      for(int k = s.row_ptr[r]; k < s.row_ptr[r + 1]; ++k) S.col[o + (k - s.row_ptr[r])] = grow(i, s.col[k]);
      S.diag[g] = (int)(o + (s.diag[r] - s.row_ptr[r]));
      const int co = S.contrib_ptr[g];
      for(int e = s.contrib_ptr[r]; e < s.contrib_ptr[r + 1]; ++e) {
        const int cm = s.contrib[e];
        const int c = cm / nl, lm = cm - c * nl;
        const int ge = co + (e - s.contrib_ptr[r]);
        S.contrib[ge] = (I.cell_off + c) * nl + lm;
        std::memcpy(&S.slot[(size_t)ge * nl], &s.slot[(size_t)e * nl], nl);
      }
    }
    for(int c = 0; c < m.n_cells; ++c) {
      for(int k = 0; k < nl; ++k) gdof[(size_t)(I.cell_off + c) * nl + k] = grow(i, L.space.cell_dof(c)[k]);
      for(int a = 0; a < m.nv; ++a)
        for(int k = 0; k < 3; ++k) xyz.push_back(m.coords[3 * m.cell(c)[a] + k]);
      cinst[I.cell_off + c] = i;
    }
    // Dirichlet values at the nodal points
    const ElementInfo& info = element_info(L.space.element);
    const int na = L.n_active();
    std::vector<char> seen(L.n() - na, 0);
    for(int c = 0; c < m.n_cells; ++c)
      for(int k = 0; k < nl; ++k) {
        const int r = L.space.cell_dof(c)[k];
        if(r < na || seen[r - na]) continue;
        seen[r - na] = 1;
        double x[3];
        node_position(m, c, info, k, x);
        dirv[I.dir_off + (r - na)] = example_bvalue(descs_[i].ex, 0, x[0], x[1], x[2]);
      }
    if(m.dim() == 2)
      for(int c = 0; c < m.n_cells; ++c)
        for(int j = 0; j < m.nj; ++j) {
          const int comp = m.joint_comp[(size_t)c * m.nj + j];
          PMCD_CHECK(comp < 0 || example_bc(descs_[i].ex, comp) != NEUMANN,
                     "batched instances support Dirichlet boundaries only");
        }
    exs[i] = descs_[i].ex;
    exs[i].dim = m.dim();
  }
  G_.space.n_dof = n_;
  G_.space.n_active = na_;
  G_.space.n_local = nl;
  G_.space.element = inst_[0].L->space.element;
  G_.cell_xyz.upload(xyz, stream_);
  G_.dof.upload(gdof, stream_);
  G_.row_ptr.upload(S.row_ptr, stream_);
  G_.col.upload(S.col, stream_);
  G_.diag_pos.upload(S.diag, stream_);
  G_.contrib_ptr.upload(S.contrib_ptr, stream_);
  G_.contrib.upload(S.contrib, stream_);
  G_.slot.upload(S.slot, stream_);
  G_.dir_vals.upload(dirv, stream_);
  G_.val.alloc(S.nnz);
  G_.rhs.alloc(n_);
  G_.sol.alloc(n_);
  G_.diag.alloc(n_);
  cell_ex_.upload(exs, stream_);
  cell_inst_.upload(cinst, stream_);
  // quadrature and reference tables (shared by all instances)
  bool constant = true;
  for(const auto& d : descs) constant = constant && example_constant_matrix_coeffs(d.ex.kind);
  rule_ = upload_reference_tables(shape, d0.order, d0.linear_quad != 0, constant, ncomp_, ref_, qw_, qp_, tensor_,
                                  stream_);
  local_A_.alloc((size_t)cells * nl * nl);
  local_rhs_.alloc((size_t)cells * nl);
  // reduction chunks: each instance's active range, then its Dirichlet range
  std::vector<int> iptr(n_inst() + 1, 0);
  for(int i = 0; i < n_inst(); ++i) {
    const Inst& I = inst_[i];
    const int ranges[2][2] = {{I.act_off, I.act_off + I.L->n_active()},
                              {na_ + I.dir_off, na_ + I.dir_off + (I.L->n() - I.L->n_active())}};
    for(auto& rg : ranges)
      for(int b = rg[0]; b < rg[1]; b += kChunkRows) chunks_host_.push_back({i, b, std::min(rg[1], b + kChunkRows)});
    iptr[i + 1] = (int)chunks_host_.size();
  }
  chunks_.upload(chunks_host_, stream_);
  inst_ptr_.upload(iptr, stream_);
  std::vector<int> rinst(n_);
  for(int i = 0; i < n_inst(); ++i)
    for(int r = 0; r < inst_[i].L->n(); ++r) rinst[grow(i, r)] = i;
  row_inst_.upload(rinst, stream_);
  active_host_.assign(n_inst(), 1);
  active_.upload(active_host_, stream_);
  partial_.alloc(chunks_host_.size() + 1);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

BatchSystem::~BatchSystem() { (void)hipStreamSynchronize(stream_); }

void BatchSystem::assemble() {
  PMCD_HIP(hipStreamSynchronize(stream_));
  const double t0 = now_seconds();
  AsmArgs a;
  const Mesh& m0 = inst_[0].L->mesh;
  a.shape = (int)m0.shape;
  a.dim = m0.dim();
  a.nv = m0.nv;
  a.nl = G_.space.n_local;
  a.nq = quad_formula(rule_).n();
  a.ncomp = ncomp_;
  a.supg = descs_[0].supg;
  a.n_cells = (int)n_cells_;
  a.ex = descs_[0].ex;
  a.ex.dim = a.dim;
  a.cell_ex = cell_ex_.p;
  a.cell_inst = cell_inst_.p;
  a.ref = ref_.p;
  a.qw = qw_.p;
  a.qp = qp_.p;
  a.tensor = tensor_.n ? tensor_.p : nullptr;
  a.cell_xyz = G_.cell_xyz.p;
  a.local_A = local_A_.p;
  a.local_rhs = local_rhs_.p;
  assemble_cells(a, stream_);
  launch_gather_matrix(n_, na_, a.nl, G_.row_ptr.p, G_.col.p, G_.contrib_ptr.p, G_.contrib.p,
                       G_.slot.p, local_A_.p, G_.val.p, stream_);
  launch_gather_rhs(n_, na_, G_.contrib_ptr.p, G_.contrib.p, local_rhs_.p, nullptr, nullptr,
                    G_.dir_vals.p, G_.rhs.p, stream_);
  launch_gather_idx(n_, G_.diag_pos.p, G_.val.p, G_.diag.p, stream_);
  // solution: 0 on active rows, Dirichlet values on the tail
  launch_fill(na_, 0.0, G_.sol.p, stream_);
  launch_copy(n_ - na_, G_.rhs.p + na_, G_.sol.p + na_, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  last_assemble_ms = (now_seconds() - t0) * 1e3;
  mixed_ready_ = false;
  G_.sw_fresh = G_.sw_fresh32 = false;
}

void BatchSystem::seg_dot(const double* x, const double* y, double* out_dev) {
  launch_seg_dot((int)chunks_host_.size(), chunks_.p, n_inst(), inst_ptr_.p, x, y, partial_.p, out_dev,
                 stream_, active_.p);
}

// y = beta z + alpha A x with the non-active rule; rows of converged
// instances become 0 (their vectors are all zero from then on)
void BatchSystem::spmv(const double* x, const double* z, double* y, double alpha, double beta) {
  if(G_.st.max_row <= 64)
    launch_spmv_masked(n_, na_, G_.row_ptr.p, G_.col.p, G_.val.p, x, z, y, alpha, beta, G_.st.max_row,
                       row_inst_.p, active_.p, stream_);
  else
    launch_spmv_rowlen(n_, na_, G_.row_ptr.p, G_.col.p, G_.val.p, x, z, y, alpha, beta, stream_,
                       G_.st.max_row, (double)G_.st.nnz / std::max(1, n_));
}

void BatchSystem::precond(const double* v, double* z) {
  switch(sd_.precond) {
    case 0: launch_copy(n_, v, z, stream_); break;
    case 1: launch_div(n_, v, G_.diag.p, z, stream_); break;
    case 3:
    case 4: {
      const int flag = sd_.precond == 3 ? 0 : 2;
      if(sd_.mixed) {
        launch_d2f(n_, v, f_v_.p, stream_);
        launch_fill_f32(n_, 0.0f, f_z_.p, stream_);
        level_sor_sweep32(G_, f_v_.p, f_z_.p, (float)sd_.sor_omega, flag, stream_, row_inst_.p, active_.p);
        launch_f2d(n_, f_z_.p, z, stream_);
      } else {
        launch_fill(n_, 0.0, z, stream_);
        level_sor_sweep(G_, v, z, sd_.sor_omega, flag, stream_, row_inst_.p, active_.p);
      }
      break;
    }
    default: PMCD_FAIL("batched solves support no, jacobi, sor and ssor preconditioning");
  }
}

// Iteration_gmres::flexible_gmres (src/Solver/Iteration_gmres.C:430-620) on
// every instance at once: the Krylov vectors are global vectors whose
// instance segments are independent; Hessenberg, Givens and convergence are
// per instance on the host (one device->host copy per iteration). A
// converged instance's segments are scaled to zero and stay zero.
std::vector<SolveStats> BatchSystem::solve(const SolverDesc& sd) {
  sd_ = sd;
  PMCD_CHECK(sd.method == 2, "batched solves use flexible GMRES");
  const int n = n_, ni = n_inst(), m = sd.restart;
  const double t0 = now_seconds();
  if(sd.precond == 3 || sd.precond == 4) prepare_level_sweep(G_, stream_);
  if(sd.mixed && !mixed_ready_) {
    G_.val32.alloc(G_.val.n);
    G_.diag32.alloc(n);
    launch_d2f((int)G_.val.n, G_.val.p, G_.val32.p, stream_);
    launch_d2f(n, G_.diag.p, G_.diag32.p, stream_);
    f_v_.alloc(n);
    f_z_.alloc(n);
    mixed_ready_ = true;
  }
  if(kv_.size() != (size_t)m + 1 || kr_.n != (size_t)n) {
    kv_.clear(); kz_.clear();
    for(int i = 0; i <= m; ++i) kv_.emplace_back(n);
    for(int i = 0; i < m; ++i) kz_.emplace_back(n);
    kr_.alloc(n);
    kw_.alloc(n);
  }
  kh_.alloc((size_t)(m + 2) * ni);
  kscale_.alloc(ni);
  kcoef_.alloc((size_t)(m + 2) * ni);
  std::vector<SolveStats> st(ni);
  std::vector<InstConv> conv(ni);
  std::vector<char> active(ni, 1);
  std::vector<double> H((size_t)ni * (m + 1) * m, 0.0), s((size_t)ni * (m + 1)), cs((size_t)ni * (m + 1)),
      sn((size_t)ni * (m + 1)), resid(ni), beta(ni), scale(ni), hcol((size_t)(m + 2) * ni);
  auto h = [&](int inst, int i, int j) -> double& { return H[((size_t)inst * m + j) * (m + 1) + i]; };
  double* x = G_.sol.p;
  double* r = kr_.p;
  auto residual = [&]() {  // r = b - A x, beta = |r| per instance
    launch_copy(n, G_.rhs.p, r, stream_);
    spmv(x, r, r, -1.0, 1.0);
    seg_dot(r, r, kh_.p);
    PMCD_HIP(hipMemcpyAsync(beta.data(), kh_.p, sizeof(double) * ni, hipMemcpyDeviceToHost, stream_));
    PMCD_HIP(hipStreamSynchronize(stream_));
    for(auto& b : beta) b = std::sqrt(b);
  };
  // active flags on the device follow active_host_ (copied after each change;
  // the host array changes again only after a later stream synchronisation)
  auto push_active = [&]() {
    PMCD_HIP(hipMemcpyAsync(active_.p, active_host_.data(), ni, hipMemcpyHostToDevice, stream_));
  };
  std::fill(active_host_.begin(), active_host_.end(), 1);
  push_active();
  auto set_scale = [&](const std::vector<double>& a) {
    PMCD_HIP(hipMemcpyAsync(kscale_.p, a.data(), sizeof(double) * ni, hipMemcpyHostToDevice, stream_));
  };
  // x += sum_k y_k z_k for the instances in `which`, after a back-solve of
  // their Hessenberg systems of size k + 1 (Utilities_gmres.C:90-108); the
  // sum is accumulated in kw_ (r holds the current Arnoldi vector)
  std::vector<double> coef((size_t)(m + 2) * ni);
  auto update = [&](int k, const std::vector<char>& which) {
    std::fill(coef.begin(), coef.end(), 0.0);
    bool any = false;
    std::vector<double> yi(k + 1);
    for(int inst = 0; inst < ni; ++inst) {
      if(!which[inst]) continue;
      any = true;
      for(int i = 0; i <= k; ++i) yi[i] = s[(size_t)inst * (m + 1) + i];
      for(int i = k; i >= 0; --i) {
        yi[i] /= h(inst, i, i);
        for(int jj = i - 1; jj >= 0; --jj) yi[jj] -= h(inst, jj, i) * yi[i];
      }
      for(int i = 0; i <= k; ++i) coef[(size_t)i * ni + inst] = yi[i];
      coef[(size_t)(k + 1) * ni + inst] = 1.0;
    }
    if(!any) return;
    PMCD_HIP(hipMemcpyAsync(kcoef_.p, coef.data(), sizeof(double) * (size_t)(k + 2) * ni,
                            hipMemcpyHostToDevice, stream_));
    launch_fill(n, 0.0, kw_.p, stream_);
    for(int jj = 0; jj <= k; ++jj)
      launch_seg_axpy((int)chunks_host_.size(), chunks_.p, kcoef_.p + (size_t)jj * ni, kz_[jj].p, kw_.p,
                      stream_);
    launch_seg_axpy((int)chunks_host_.size(), chunks_.p, kcoef_.p + (size_t)(k + 1) * ni, kw_.p, x,
                    stream_);
    PMCD_HIP(hipStreamSynchronize(stream_));
  };
  auto sync_active = [&]() {
    for(int inst = 0; inst < ni; ++inst) active_host_[inst] = active[inst] ? 1 : 0;
    push_active();
  };
  residual();
  int n_active_inst = 0;
  for(int inst = 0; inst < ni; ++inst) {
    resid[inst] = beta[inst];
    st[inst].initial_residual = beta[inst];
    if(conv[inst](sd_, beta[inst], 0)) {
      active[inst] = 0;
      st[inst].final_residual = beta[inst];
      st[inst].converged = 1;
    } else {
      ++n_active_inst;
    }
  }
  sync_active();
  int j = 1;
  while(j <= sd_.max_iter && n_active_inst > 0) {
    for(int inst = 0; inst < ni; ++inst) {
      scale[inst] = active[inst] ? 1.0 / beta[inst] : 0.0;
      std::fill(&s[(size_t)inst * (m + 1)], &s[(size_t)inst * (m + 1)] + m + 1, 0.0);
      s[(size_t)inst * (m + 1)] = beta[inst];
    }
    set_scale(scale);
    launch_seg_scale((int)chunks_host_.size(), chunks_.p, kscale_.p, r, kv_[0].p, stream_, active_.p);
    int i = 0;
    for(; i < m && j <= sd_.max_iter && n_active_inst > 0; ++i, ++j) {
      precond(kv_[i].p, kz_[i].p);
      spmv(kz_[i].p, nullptr, r, 1.0, 0.0);
      // modified Gram-Schmidt, one fused step per basis vector
      seg_dot(r, kv_[0].p, kh_.p);
      for(int k = 0; k <= i; ++k)
        launch_seg_mgs((int)chunks_host_.size(), chunks_.p, ni, inst_ptr_.p, kh_.p + (size_t)k * ni, kv_[k].p,
                       r, k < i ? kv_[k + 1].p : nullptr, partial_.p, kh_.p + (size_t)(k + 1) * ni, stream_,
                       active_.p);
      seg_dot(r, r, kh_.p + (size_t)(i + 1) * ni);
      PMCD_HIP(hipMemcpyAsync(hcol.data(), kh_.p, sizeof(double) * (size_t)(i + 2) * ni,
                              hipMemcpyDeviceToHost, stream_));
      PMCD_HIP(hipStreamSynchronize(stream_));
      std::vector<char> done(ni, 0);
      bool any_done = false;
      for(int inst = 0; inst < ni; ++inst) {
        scale[inst] = 0.0;
        if(!active[inst]) continue;
        for(int k = 0; k <= i; ++k) h(inst, k, i) = hcol[(size_t)k * ni + inst];
        h(inst, i + 1, i) = std::sqrt(hcol[(size_t)(i + 1) * ni + inst]);
        PMCD_CHECK(h(inst, i + 1, i) != 0.0, "Unhappy breakdown in flexible gmres at iteration ", j,
                   " of batch instance ", inst);
        scale[inst] = 1.0 / h(inst, i + 1, i);
        double* si = &s[(size_t)inst * (m + 1)];
        double* ci = &cs[(size_t)inst * (m + 1)];
        double* ni_ = &sn[(size_t)inst * (m + 1)];
        for(int k = 0; k < i; ++k) givens_app(h(inst, k, i), h(inst, k + 1, i), ci[k], ni_[k]);
        givens_gen(h(inst, i, i), h(inst, i + 1, i), ci[i], ni_[i]);
        givens_app(h(inst, i, i), h(inst, i + 1, i), ci[i], ni_[i]);
        givens_app(si[i], si[i + 1], ci[i], ni_[i]);
        resid[inst] = std::abs(si[i + 1]);
        st[inst].iterations = j;
        if(conv[inst](sd_, resid[inst], j)) {
          done[inst] = 1;
          any_done = true;
        }
      }
      if(any_done) {
        update(i, done);
        for(int inst = 0; inst < ni; ++inst)
          if(done[inst]) {
            active[inst] = 0;
            scale[inst] = 0.0;
            --n_active_inst;
            st[inst].final_residual = resid[inst];
            st[inst].converged = resid[inst] < sd_.res_tol ||
                                 resid[inst] < sd_.res_red * st[inst].initial_residual;
          }
        sync_active();
      }
      set_scale(scale);
      launch_seg_scale((int)chunks_host_.size(), chunks_.p, kscale_.p, r, kv_[i + 1].p, stream_, active_.p);
    }
    if(n_active_inst == 0) break;
    // restart: update, true residual
    update(i - 1, active);
    residual();
    for(int inst = 0; inst < ni; ++inst) {
      if(!active[inst]) continue;
      if(conv[inst](sd_, resid[inst], j)) {
        active[inst] = 0;
        --n_active_inst;
        st[inst].final_residual = resid[inst];
        st[inst].converged = resid[inst] < sd_.res_tol ||
                             resid[inst] < sd_.res_red * st[inst].initial_residual;
      }
    }
    sync_active();
  }
  for(int inst = 0; inst < ni; ++inst)
    if(active[inst]) {
      st[inst].final_residual = resid[inst];
      st[inst].converged = resid[inst] < sd_.res_tol || resid[inst] < sd_.res_red * st[inst].initial_residual;
    }
  PMCD_HIP(hipStreamSynchronize(stream_));
  const double secs = now_seconds() - t0;
  for(auto& x : st) x.seconds = secs;
  return st;
}

void BatchSystem::get_solution(int i, double* out) {
  PMCD_CHECK(i >= 0 && i < n_inst(), "instance out of range");
  std::vector<double> g(n_);
  G_.sol.download(g.data(), n_, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  for(int r = 0; r < inst_[i].L->n(); ++r) out[r] = g[grow(i, r)];
}

void BatchSystem::set_solution(int i, const double* in) {
  PMCD_CHECK(i >= 0 && i < n_inst(), "instance out of range");
  std::vector<double> g(n_);
  G_.sol.download(g.data(), n_, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  for(int r = 0; r < inst_[i].L->n(); ++r) g[grow(i, r)] = in[r];
  G_.sol.upload(g, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

void BatchSystem::get_rhs(int i, double* out) {
  PMCD_CHECK(i >= 0 && i < n_inst(), "instance out of range");
  std::vector<double> g(n_);
  G_.rhs.download(g.data(), n_, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  for(int r = 0; r < inst_[i].L->n(); ++r) out[r] = g[grow(i, r)];
}

void BatchSystem::get_matrix(int i, double* out) {
  PMCD_CHECK(i >= 0 && i < n_inst(), "instance out of range");
  std::vector<double> g(G_.st.nnz);
  G_.val.download(g.data(), g.size(), stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  const Structure& s = inst_[i].L->st;
  for(int r = 0; r < s.n_rows; ++r) {
    const int64_t o = G_.st.row_ptr[grow(i, r)];
    for(int k = s.row_ptr[r]; k < s.row_ptr[r + 1]; ++k) out[k] = g[o + (k - s.row_ptr[r])];
  }
}

}  // namespace pmcd
