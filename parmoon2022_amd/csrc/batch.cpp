#include "batch.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace pmcd {


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

BatchSystem::~BatchSystem() {
  (void)hipStreamSynchronize(stream_);
  clear_graphs();
  if(poll_host_) (void)hipHostFree(poll_host_);
  for(auto& e : poll_ev_)
    if(e) (void)hipEventDestroy(e);
}

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

bool BatchSystem::coop_sweep(const void* b, void* x, bool f32, int flag) {
  if(!batch_coop() || G_.sweep_exact || G_.sw_rp.n == 0) return false;
  const int K = (int)G_.sweep_ptr.size() - 1;
  const double mean_row = (double)G_.st.nnz / std::max(1, G_.n());
  if(f32)
    return launch_sor_sweep_coop_f32(K, G_.sweep_ptr_dev.p, G_.sweep_max_class, G_.sweep_rows.p, G_.sw_rp.p,
                                     G_.sw_col.p, G_.sw_val32.p, G_.diag32.p, (const float*)b, (float*)x,
                                     (float)sd_.sor_omega, flag, G_.st.max_row, mean_row, row_inst_.p, active_.p,
                                     stream_);
  return launch_sor_sweep_coop(K, G_.sweep_ptr_dev.p, G_.sweep_max_class, G_.sweep_rows.p, G_.sw_rp.p, G_.sw_col.p,
                               G_.sw_val.p, G_.diag.p, (const double*)b, (double*)x, sd_.sor_omega, flag,
                               G_.st.max_row, mean_row, row_inst_.p, active_.p, stream_);
}

void BatchSystem::clear_graphs() {
  for(auto& g : graphs_)
    if(g) (void)hipGraphExecDestroy(g);
  graphs_.clear();
  graph_key_ = -1;
}

// Iteration_gmres::flexible_gmres (src/Solver/Iteration_gmres.C:430-620) on
// every instance at once: the Krylov vectors are global vectors whose
// instance segments are independent. Each instance's Hessenberg matrix,
// Givens rotations, convergence test and update live on the device
// (k_batch_givens, k_seg_update, k_batch_restart_state), so an iteration is a
// fixed kernel sequence: with an SOR/SSOR preconditioner on multicolour
// classes one cooperative launch per sweep, then one replayed graph for the
// rest of the iteration (the product, the Gram-Schmidt steps with their
// reductions folded in, the next basis vector, the rotations and the
// updates of instances that converged). The host polls the active flags one
// iteration behind; a converged instance is masked out of every kernel.
std::vector<SolveStats> BatchSystem::solve(const SolverDesc& sd) {
  sd_ = sd;
  PMCD_CHECK(sd.method == 2, "batched solves use flexible GMRES");
  PMCD_CHECK(sd.precond == 0 || sd.precond == 1 || sd.precond == 3 || sd.precond == 4,
             "batched solves support no, jacobi, sor and ssor preconditioning");
  const int n = n_, ni = n_inst(), m = sd.restart;
  const int nch = (int)chunks_host_.size();
  const double t0 = now_seconds();
  const bool sor = sd.precond == 3 || sd.precond == 4;
  const int flag = sd.precond == 3 ? 0 : 2;
  if(sor) prepare_level_sweep(G_, stream_);
  if(sd.mixed && !mixed_ready_) {
    G_.val32.alloc(G_.val.n);
    G_.diag32.alloc(n);
    launch_d2f((int)G_.val.n, G_.val.p, G_.val32.p, stream_);
    launch_d2f(n, G_.diag.p, G_.diag32.p, stream_);
    f_v_.alloc(n);
    f_z_.alloc(n);
    mixed_ready_ = true;
  }
  // derived matrix copies brought up to date outside any capture
  if(sor) refresh_sweep_values(G_, sd.mixed != 0, stream_);
  if(kv_.size() != (size_t)m + 1 || kr_.n != (size_t)n) {
    clear_graphs();
    kv_.clear(); kz_.clear();
    for(int i = 0; i <= m; ++i) kv_.emplace_back(n);
    for(int i = 0; i < m; ++i) kz_.emplace_back(n);
    kr_.alloc(n);
    kw_.alloc(n);
    std::vector<const double*> zp(m);
    for(int i = 0; i < m; ++i) zp[i] = kz_[i].p;
    kz_ptr_.upload(zp, stream_);
  }
  if(kh_.n != (size_t)(m + 2) * ni) {
    clear_graphs();
    kh_.alloc((size_t)(m + 2) * ni);
    bH_.alloc((size_t)ni * m * (m + 1));
    bs_.alloc((size_t)ni * (m + 1));
    bcs_.alloc((size_t)ni * (m + 1));
    bsn_.alloc((size_t)ni * (m + 1));
    by_.alloc((size_t)ni * m);
    bscale_.alloc(ni);
    bst_.alloc(ni);
    bbreak_.alloc(1);
    bpart_.alloc(2 * (size_t)std::max(1, nch));
  }
  if(!poll_host_) {
    PMCD_HIP(hipHostMalloc((void**)&poll_host_, 2 * ((size_t)ni + 8), hipHostMallocDefault));
    for(auto& e : poll_ev_) PMCD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  const BatchSolveParams prm{sd.res_tol, sd.res_red, sd.div_factor, sd.min_iter, sd.max_iter, m};
  double* P[2] = {bpart_.p, bpart_.p + std::max(1, nch)};
  double* x = G_.sol.p;
  double* r = kr_.p;
  // a cooperative sweep (multicolour classes, device support) leaves the rest
  // of an iteration to the graphs
  const bool coop = sor && batch_coop() && !G_.sweep_exact && G_.sw_rp.n > 0 && cooperative_launch_supported();
  auto sweep_input = [&](int i) {  // before the sweep of iteration i
    if(sd.mixed) {
      launch_d2f(n, kv_[i].p, f_v_.p, stream_);
      launch_fill_f32(n, 0.0f, f_z_.p, stream_);
    } else {
      launch_fill(n, 0.0, kz_[i].p, stream_);
    }
  };
  auto run_sweep = [&](int i) {
    const bool ran = sd.mixed ? coop_sweep(f_v_.p, f_z_.p, true, flag) : coop_sweep(kv_[i].p, kz_[i].p, false, flag);
    PMCD_CHECK(ran, "cooperative sweep not launched");
  };
  // every instance starts active; no breakdown yet (outside any capture)
  std::fill(active_host_.begin(), active_host_.end(), 1);
  PMCD_HIP(hipMemcpyAsync(active_.p, active_host_.data(), ni, hipMemcpyHostToDevice, stream_));
  PMCD_HIP(hipMemsetAsync(bbreak_.p, 0, sizeof(int), stream_));
  auto residual = [&](int mode) {  // r = b - A x, beta, tests
    launch_copy(n, G_.rhs.p, r, stream_);
    spmv(x, r, r, -1.0, 1.0);
    launch_seg_dot_partial(nch, chunks_.p, mode == 0 ? nullptr : active_.p, r, r, P[0], stream_);
    launch_batch_restart_state(ni, prm, inst_ptr_.p, P[0], bst_.p, bs_.p, bscale_.p, active_.p, mode, stream_);
    launch_seg_scale(nch, chunks_.p, bscale_.p, r, kv_[0].p, stream_, active_.p);
  };
  // the part of iteration i after the preconditioner's sweep (or all of it)
  auto tail = [&](int i, bool after_coop) {
    if(after_coop) {
      if(sd.mixed) launch_f2d(n, f_z_.p, kz_[i].p, stream_);
    } else {
      precond(kv_[i].p, kz_[i].p);
    }
    spmv(kz_[i].p, nullptr, r, 1.0, 0.0);
    // modified Gram-Schmidt: H(k,i) = <r, v_k>; r -= H(k,i) v_k; |r|
    launch_seg_dot_partial(nch, chunks_.p, active_.p, r, kv_[0].p, P[0], stream_);
    for(int k = 0; k <= i; ++k)
      launch_seg_mgs2(nch, chunks_.p, active_.p, inst_ptr_.p, P[k & 1], kh_.p + (size_t)k * ni, kv_[k].p, r,
                      k < i ? kv_[k + 1].p : nullptr, P[(k + 1) & 1], stream_);
    launch_seg_normscale(nch, chunks_.p, active_.p, inst_ptr_.p, P[(i + 1) & 1], kh_.p + (size_t)(i + 1) * ni, r,
                         kv_[i + 1].p, stream_);
    launch_batch_givens(ni, i, prm, kh_.p, bH_.p, bs_.p, bcs_.p, bsn_.p, bst_.p, by_.p, active_.p, bbreak_.p,
                        stream_);
    launch_seg_update(nch, chunks_.p, bst_.p, m, by_.p, kz_ptr_.p, x, stream_);
    if(after_coop && i + 1 < m) sweep_input(i + 1);
  };
  // graphs: one per iteration index; valid for the buffers, the solver
  // parameters and the launch shapes they were captured with
  const long long key = ((((long long)m * 8 + sd.precond) * 2 + sd.mixed) * 2 + (coop ? 1 : 0)) * 100003LL +
                        tuning_epoch();
  if(graph_key_ != key) {
    clear_graphs();
    graph_key_ = key;
    graphs_.assign(m, nullptr);
  }
  auto run_tail = [&](int i, bool after_coop) {
    // with cooperative sweeps the rest of the iteration is launched kernel
    // by kernel: graph replay and cooperative sweeps together have ended a
    // replayed solve with a breakdown now and then (FP32 sweeps), which no
    // ordering in this code explains; every launch here is on one stream
    if(!batch_graphs() || coop) {
      tail(i, after_coop);
      return;
    }
    if(!graphs_[i]) {
      PMCD_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      try {
        tail(i, after_coop);
      } catch(...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if(g) (void)hipGraphDestroy(g);
        throw;
      }
      hipGraph_t graph = nullptr;
      PMCD_HIP(hipStreamEndCapture(stream_, &graph));
      hipGraphExec_t exec = nullptr;
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      PMCD_HIP(e);
      graphs_[i] = exec;
    }
    PMCD_HIP(hipGraphLaunch(graphs_[i], stream_));
  };
  // host view of the active flags, one iteration behind the device
  int slot = 0;
  bool pending = false, all_done = false;
  auto poll = [&]() {
    uint8_t* h = poll_host_ + (size_t)slot * (ni + 8);
    PMCD_HIP(hipMemcpyAsync(h, active_.p, ni, hipMemcpyDeviceToHost, stream_));
    PMCD_HIP(hipMemcpyAsync(h + ni, bbreak_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
    PMCD_HIP(hipEventRecord(poll_ev_[slot], stream_));
    if(pending) {
      const int prev = slot ^ 1;
      PMCD_HIP(hipEventSynchronize(poll_ev_[prev]));
      const uint8_t* hp = poll_host_ + (size_t)prev * (ni + 8);
      int brk = 0;
      std::memcpy(&brk, hp + ni, sizeof(int));
      PMCD_CHECK(brk == 0, "Unhappy breakdown in flexible gmres of a batch instance");
      all_done = std::none_of(hp, hp + ni, [](uint8_t a) { return a != 0; });
    }
    pending = true;
    slot ^= 1;
  };
  residual(0);
  poll();
  int j = 1;
  while(j <= sd.max_iter && !all_done) {
    int i = 0;
    if(coop) sweep_input(0);
    for(; i < m && j <= sd.max_iter && !all_done; ++i, ++j) {
      if(coop) run_sweep(i);
      run_tail(i, coop);
      poll();
    }
    if(all_done || j > sd.max_iter) break;
    // restart: the update with the cycle's m coefficients, the true residual
    launch_batch_restart_coef(ni, prm, bH_.p, bs_.p, bst_.p, by_.p, active_.p, stream_);
    launch_seg_update(nch, chunks_.p, bst_.p, m, by_.p, kz_ptr_.p, x, stream_);
    residual(1);
    poll();
  }
  PMCD_HIP(hipStreamSynchronize(stream_));
  {
    // the last flags (breakdown of the final iterations)
    int brk = 0;
    PMCD_HIP(hipMemcpy(&brk, bbreak_.p, sizeof(int), hipMemcpyDeviceToHost));
    PMCD_CHECK(brk == 0, "Unhappy breakdown in flexible gmres of a batch instance");
  }
  std::vector<BatchInstState> fin(ni);
  bst_.download(fin.data(), ni, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  // every instance leaves the active set converged or at max_n_iterations
  std::fill(active_host_.begin(), active_host_.end(), 1);
  PMCD_HIP(hipMemcpyAsync(active_.p, active_host_.data(), ni, hipMemcpyHostToDevice, stream_));
  PMCD_HIP(hipStreamSynchronize(stream_));
  std::vector<SolveStats> st(ni);
  const double secs = now_seconds() - t0;
  for(int inst = 0; inst < ni; ++inst) {
    st[inst].iterations = fin[inst].iterations;
    st[inst].initial_residual = fin[inst].initial;
    st[inst].final_residual = fin[inst].resid;
    st[inst].converged = fin[inst].converged;
    st[inst].seconds = secs;
  }
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
