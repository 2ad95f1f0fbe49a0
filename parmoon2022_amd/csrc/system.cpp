#include "system.hpp"
#include "elements.hpp"
#include <chrono>
#include <cmath>
#include <algorithm>
#include <climits>
#include <functional>

namespace pmcd {

// physical position of the nodal point of local DOF i of cell c
void node_position(const Mesh& m, int c, const ElementInfo& info, int i, double* x) {
  const int* v = m.cell(c);
  const auto& p = info.nodes[i];
  for(int k = 0; k < 3; ++k) x[k] = 0;
  if(shape_is_simplex(m.shape)) {
    double lam[4] = {1 - p[0] - p[1] - (m.dim() == 3 ? p[2] : 0.0), p[0], p[1], p[2]};
    for(int a = 0; a <= m.dim(); ++a)
      for(int k = 0; k < 3; ++k) x[k] += lam[a] * m.coords[3 * v[a] + k];
  } else {
    // multilinear interpolation of the vertices (tensor reference cell [-1,1]^d)
    const int nvx = m.nv;
    for(int a = 0; a < nvx; ++a) {
      // vertex a's reference corner
      static const int q2[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
      int cx = q2[a % 4][0], cy = q2[a % 4][1], cz = a / 4;
      double w = (cx ? (1 + p[0]) : (1 - p[0])) / 2 * ((cy ? (1 + p[1]) : (1 - p[1])) / 2);
      if(m.dim() == 3) w *= (cz ? (1 + p[2]) : (1 - p[2])) / 2;
      for(int k = 0; k < 3; ++k) x[k] += w * m.coords[3 * v[a] + k];
    }
  }
}

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {
double now_s() { return now_seconds(); }
}  // namespace

std::vector<std::unique_ptr<Level>> build_levels(const ProblemDesc& d) {
  PMCD_CHECK(d.order == 1 || d.order == 2, "ansatz order must be 1 or 2");
  PMCD_CHECK(d.n_refine >= 0 && d.n_refine <= 12, "n_refine out of range");
  Mesh base;
  if(d.geometry == 10 || d.geometry == 11) base = structured_cube(d.structured_n, d.geometry == 11);
  else if(d.geometry == 12) {
    const std::string& f = d.mesh_file;
    const bool geo = f.size() > 4 && (f.compare(f.size() - 4, 4, ".GEO") == 0 || f.compare(f.size() - 4, 4, ".geo") == 0);
    base = geo ? read_geo_file(f) : read_mesh_file(f);
  }
  else {
    PMCD_CHECK(d.geometry >= 0 && d.geometry <= 4, "unknown geometry ", d.geometry);
    base = default_mesh((DefaultGeometry)d.geometry);
  }
  std::vector<Mesh> meshes;
  meshes.push_back(std::move(base));
  for(int i = 0; i < d.n_refine; ++i) meshes.push_back(refine_regular(meshes.back()));
  const int keep = std::max(1, std::min(d.mg_levels, (int)meshes.size()));
  const ExampleParams ex = d.ex;
  BoundCondFn bc = [ex](int comp, const double*) { return (BoundCond)example_bc(ex, comp); };
  std::vector<std::unique_ptr<Level>> levels;
  if(d.adapt_steps > 0) {
    PMCD_CHECK(d.indicator, "refinement by indicator without an indicator function");
    PMCD_CHECK(meshes.back().dim() == 2, "hanging vertices are implemented for 2D meshes");
    PMCD_CHECK(d.mg_levels <= 1, "multigrid on meshes refined by indicator is not implemented");
    auto L = std::make_unique<Level>();
    std::vector<HangingVertex> hanging;
    const IndicatorFn f = [&d](const double* x) { return d.indicator(x, d.indicator_ctx); };
    L->mesh = refine_by_indicator(meshes.back(), d.adapt_steps, f, hanging, d.adapt_conforming != 0);
    L->space = build_fespace(L->mesh, d.order, bc, &hanging);
    L->st = build_structure(L->space, L->mesh.n_cells);
    levels.push_back(std::move(L));
    return levels;
  }
  for(int l = (int)meshes.size() - keep; l < (int)meshes.size(); ++l) {
    auto L = std::make_unique<Level>();
    L->mesh = std::move(meshes[l]);
    L->space = build_fespace(L->mesh, d.order, bc);
    L->st = build_structure(L->space, L->mesh.n_cells);
    levels.push_back(std::move(L));
  }
  return levels;
}

QuadRule upload_reference_tables(Shape shape, int order, bool linear_quad, bool tensors, int& ncomp,
                                 DevBuf<double>& ref_d, DevBuf<double>& qw_d, DevBuf<double>& qp_d,
                                 DevBuf<double>& tensor_d, hipStream_t stream) {
  const QuadRule rule = cell_rule_for(shape, order, linear_quad);
  const QuadFormula& qf = quad_formula(rule);
  const Element el = element_for(shape, order);
  const int nl = element_n_dof(el);
  ncomp = shape_dim(shape) == 2 ? kNDeriv2D : kNDeriv3D;
  std::vector<double> ref((size_t)qf.n() * ncomp * nl), qp((size_t)qf.n() * 3);
  for(int q = 0; q < qf.n(); ++q) {
    double p[3] = {qf.xi[q], qf.dim >= 2 ? qf.eta[q] : 0.0, qf.dim >= 3 ? qf.zeta[q] : 0.0};
    eval_basis(el, p, &ref[(size_t)q * ncomp * nl]);
    for(int k = 0; k < 3; ++k) qp[3 * q + k] = p[k];
  }
  ref_d.upload(ref, stream);
  qw_d.upload(qf.w, stream);
  qp_d.upload(qp, stream);
  if(tensors) tensor_d.upload(build_assembly_tensor(shape_dim(shape), nl, qf.n(), ref.data(), qf.w.data()), stream);
  return rule;
}

CDSystem::CDSystem(const ProblemDesc& d, hipStream_t stream) : desc_(d), stream_(stream) {
  levels_ = build_levels(d);
  // quadrature and reference tables; reference tensors when eps, b and c are
  // constant (every current example)
  const Shape shape = levels_.back()->mesh.shape;
  const int nl = element_n_dof(element_for(shape, d.order));
  rule_ = upload_reference_tables(shape, d.order, d.linear_quad != 0, example_constant_matrix_coeffs(d.ex.kind),
                                  ncomp_, ref_, qw_, qp_, tensor_, stream_);
  size_t max_cells = 0, max_n = 0;
  for(auto& L : levels_) {
    build_level_device(*L);
    max_cells = std::max(max_cells, (size_t)L->mesh.n_cells);
    max_n = std::max(max_n, (size_t)L->n());
  }
  for(size_t l = 1; l < levels_.size(); ++l) build_transfer(*levels_[l], *levels_[l - 1]);
  local_A_.alloc(max_cells * nl * nl);
  local_rhs_.alloc(max_cells * nl);
  red_partial_.alloc(1024 * 5 + 64);
  red_out_.alloc(64);
  cell_err_.alloc(max_cells * 5);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

CDSystem::CDSystem(const std::vector<OperatorLevel>& lv, hipStream_t stream) : stream_(stream) {
  PMCD_CHECK(!lv.empty(), "operator hierarchy without levels");
  desc_ = ProblemDesc{};
  desc_.mg_levels = (int)lv.size();
  operator_only_ = true;
  for(size_t l = 0; l < lv.size(); ++l) {
    const OperatorLevel& o = lv[l];
    PMCD_CHECK(o.n > 0 && o.n_active >= 0 && o.n_active <= o.n, "level ", l, ": invalid sizes");
    PMCD_CHECK((int)o.row_ptr.size() == o.n + 1 && o.row_ptr[0] == 0, "level ", l, ": row_ptr size");
    PMCD_CHECK((size_t)o.row_ptr[o.n] == o.col.size() && o.col.size() == o.val.size(), "level ", l,
               ": column and value counts differ from row_ptr");
    auto L = std::make_unique<Level>();
    L->space.n_dof = o.n;
    L->space.n_active = o.n_active;
    L->space.n_dirichlet = o.n - o.n_active;
    Structure& st = L->st;
    st.n_rows = o.n;
    st.row_ptr = o.row_ptr;
    st.col = o.col;
    st.nnz = (int64_t)o.col.size();
    st.diag.assign(o.n, -1);
    for(int r = 0; r < o.n; ++r) {
      st.max_row = std::max(st.max_row, o.row_ptr[r + 1] - o.row_ptr[r]);
      for(int k = o.row_ptr[r]; k < o.row_ptr[r + 1]; ++k) {
        PMCD_CHECK(o.col[k] >= 0 && o.col[k] < o.n, "level ", l, ": column index out of range");
        if(o.col[k] == r) st.diag[r] = k;
      }
      PMCD_CHECK(st.diag[r] >= 0, "level ", l, ": row ", r, " has no diagonal entry");
    }
    st.symmetric = pattern_symmetric(o.n, st.row_ptr.data(), st.col.data());
    L->row_ptr.upload(st.row_ptr, stream_);
    L->col.upload(st.col, stream_);
    L->diag_pos.upload(st.diag, stream_);
    L->val.upload(o.val, stream_);
    const int n = o.n;
    L->rhs.alloc(n);
    L->sol.alloc(n);
    L->diag.alloc(n);
    L->mg_sol.alloc(n); L->mg_sol2.alloc(n); L->mg_rhs.alloc(n); L->mg_def.alloc(n); L->mg_tmp.alloc(n);
    launch_fill(n, 0.0, L->rhs.p, stream_);
    launch_fill(n, 0.0, L->sol.p, stream_);
    launch_gather_idx(n, L->diag_pos.p, L->val.p, L->diag.p, stream_);
    levels_.push_back(std::move(L));
  }
  for(size_t l = 1; l < lv.size(); ++l) {
    const OperatorLevel& o = lv[l];
    PMCD_CHECK(o.n_coarse == lv[l - 1].n, "level ", l, ": prolongation columns differ from the coarser level");
    PMCD_CHECK((int)o.p_row_ptr.size() == o.n + 1 && (size_t)o.p_row_ptr[o.n] == o.p_col.size() &&
                   o.p_col.size() == o.p_val.size(),
               "level ", l, ": prolongation CSR sizes");
    for(int c : o.p_col) PMCD_CHECK(c >= 0 && c < o.n_coarse, "level ", l, ": prolongation column out of range");
    set_transfer(*levels_[l], o.n_coarse, o.p_row_ptr, o.p_col, o.p_val);
  }
  red_partial_.alloc(1024 * 5 + 64);
  red_out_.alloc(64);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

void CDSystem::set_level_values(int level, const double* val) {
  PMCD_CHECK(level >= 0 && level < n_levels(), "level out of range");
  Level& L = *levels_[level];
  L.val.upload(val, L.val.n, stream_);
  launch_gather_idx(L.n(), L.diag_pos.p, L.val.p, L.diag.p, stream_);
  L.sw_fresh = L.sw_fresh32 = L.sell_fresh = L.sell_fresh32 = false;
  coarse_ready_ = false;
  mixed_ready_ = false;
  PMCD_HIP(hipStreamSynchronize(stream_));
}

void CDSystem::set_rhs(const double* b) {
  Level& F = finest();
  F.rhs.upload(b, F.rhs.n, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

CDSystem::~CDSystem() {
  (void)hipStreamSynchronize(stream_);
  clear_graphs();
  if(cg_host_) (void)hipHostFree(cg_host_);
  if(fg_host_) (void)hipHostFree(fg_host_);
  for(auto& e : fg_ev_)
    if(e) (void)hipEventDestroy(e);
}

void CDSystem::clear_graphs() {
  for(auto& g : vgraphs_) (void)hipGraphExecDestroy(g.exec);
  vgraphs_.clear();
}

// Dirichlet values and Neumann terms of the current example on level L
void CDSystem::boundary_data(Level& L) {
  const Mesh& m = L.mesh;
  const FESpace& sp = L.space;
  const ElementInfo& info = element_info(sp.element);
  // Dirichlet values at the nodal points (Assemble3D.C:472-549, Assemble2D.C:3200-3230)
  const int n = sp.n_dof, na = sp.n_active, nl = sp.n_local;
  L.dirichlet_host.assign(n - na, 0.0);
  std::vector<char> seen(n - na, 0);
  for(int c = 0; c < m.n_cells; ++c)
    for(int i = 0; i < nl; ++i) {
      int r = sp.cell_dof(c)[i];
      if(r < na || seen[r - na]) continue;
      seen[r - na] = 1;
      double x[3];
      node_position(m, c, info, i, x);
      L.dirichlet_host[r - na] = example_bvalue(desc_.ex, 0, x[0], x[1], x[2]);
    }
  L.dir_vals.upload(L.dirichlet_host, stream_);
  // Neumann edge integrals in 2D (Assemble2D.C:3214-3255): per (cell, local
  // DOF) the terms in the order the reference adds them
  if(m.dim() == 2) {
    std::vector<std::vector<double>> terms((size_t)m.n_cells * nl);
    bool any = false;
    const auto& faces = shape_faces(m.shape);
    const QuadFormula& lq = quad_formula(line_rule_for_degree(2 * element_order(sp.element)));
    for(int c = 0; c < m.n_cells; ++c)
      for(int j = 0; j < m.nj; ++j) {
        int comp = m.joint_comp[(size_t)c * m.nj + j];
        if(comp < 0 || example_bc(desc_.ex, comp) != NEUMANN) continue;
        any = true;
        const int* v = m.cell(c);
        const double* x0 = &m.coords[3 * v[faces[j][0]]];
        const double* x1 = &m.coords[3 * v[faces[j][1]]];
        const double hE = std::sqrt((x1[0] - x0[0]) * (x1[0] - x0[0]) + (x1[1] - x0[1]) * (x1[1] - x0[1])) / 2;
        // reference coordinates of the edge end points
        double r0[3], r1[3];
        {
          const auto& A = info.nodes;  // corner nodes are nodes of the element
          auto corner = [&](int lv, double* r) {
            for(int i = 0; i < info.n_dof; ++i)
              if(info.entity_vertices[i].size() == 1 && info.entity_vertices[i][0] == lv) {
                for(int k = 0; k < 3; ++k) r[k] = A[i][k];
                return;
              }
          };
          corner(faces[j][0], r0);
          corner(faces[j][1], r1);
        }
        for(int l = 0; l < lq.n(); ++l) {
          const double s = (1 + lq.xi[l]) / 2;
          double rp[3], xp[3], phi[kMaxLocalDof * kNDeriv3D];
          for(int k = 0; k < 3; ++k) { rp[k] = r0[k] + s * (r1[k] - r0[k]); xp[k] = x0[k] + s * (x1[k] - x0[k]); }
          eval_basis(sp.element, rp, phi);
          double val = example_bvalue(desc_.ex, comp, xp[0], xp[1], 0.0);
          val *= hE * lq.w[l];
          for(int k = 0; k < nl; ++k)
            if(sp.cell_dof(c)[k] < na) terms[(size_t)c * nl + k].push_back(val * phi[k]);
        }
      }
    if(any) {
      std::vector<int> ptr((size_t)m.n_cells * nl + 1, 0);
      std::vector<double> flat;
      for(size_t p = 0; p < terms.size(); ++p) {
        ptr[p + 1] = ptr[p] + (int)terms[p].size();
        flat.insert(flat.end(), terms[p].begin(), terms[p].end());
      }
      L.extra_ptr.upload(ptr, stream_);
      L.extra.upload(flat, stream_);
    } else {
      L.extra_ptr.release();
      L.extra.release();
    }
  }
}

// New PDE data (coefficients, source, boundary values) on the same mesh and
// spaces: the device structures stay, the boundary data is recomputed; the
// next assemble() uses the new data.
void CDSystem::set_example(const ExampleParams& ex) {
  PMCD_CHECK(!operator_only_, "an operator-only system has no PDE data");
  desc_.ex = ex;
  for(auto& L : levels_) boundary_data(*L);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

// Gather lists of the hanging-DOF coupling (see k_hanging.hip). Hanging DOF h
// adds its row, weighted, to the rows of its coupling DOFs below n_active
// (hanging ones included); a target entry (c, col) therefore receives the
// entries (h, col) of the hanging rows coupled to c, in the sorted order. A
// hanging row is final once all rows coupled to it have been added: depth(h)
// = 0 without incoming hanging rows, else 1 + the largest depth among them.
// Launch k (0-based) finishes the hanging rows of depth k + 1, the last launch
// the other active rows.
void hanging_plan(Level& L, hipStream_t stream) {
  const FESpace& sp = L.space;
  const Structure& st = L.st;
  const int nh = sp.n_hanging, first = sp.n_active_non_hanging();
  constexpr int MC = FESpace::kMaxCoupling;
  std::vector<int> depth(nh, 0);
  // target DOF -> (hanging DOF, weight), sorted order
  std::vector<std::vector<std::pair<int, double>>> incoming(sp.n_dof);
  for(int h : sp.hang_order)
    for(int k = 0; k < sp.hang_n[h]; ++k) {
      const int c = sp.hang_cpl[MC * h + k];
      if(c >= sp.n_active) continue;
      incoming[c].push_back({h, sp.hang_w[MC * h + k]});
      if(c >= first) depth[c - first] = std::max(depth[c - first], depth[h] + 1);
    }
  int max_depth = 0;
  for(int h = 0; h < nh; ++h) max_depth = std::max(max_depth, depth[h]);
  const int n_launch = max_depth + 1;
  auto launch_of = [&](int c) { return c >= first && c < sp.n_active ? depth[c - first] - 1 : max_depth; };
  auto pos_in_row = [&](int r, int col) {  // a caller's rows need not be sorted
    for(int p = st.row_ptr[r]; p < st.row_ptr[r + 1]; ++p)
      if(st.col[p] == col) return p;
    return -1;
  };
  std::vector<std::vector<int>> mdst(n_launch), msp(n_launch), msrc(n_launch), rdst(n_launch), rsp(n_launch),
      rsrc(n_launch);
  std::vector<std::vector<double>> mw(n_launch), rw(n_launch);
  for(int c = 0; c < sp.n_active; ++c) {
    if(incoming[c].empty()) continue;
    // ModifyMatrixAccordingToCoupling needs each column of the hanging rows
    // in the coupling row (FEMatrix.C:275-291 throws otherwise)
    for(const auto& hw : incoming[c])
      for(int q = st.row_ptr[first + hw.first]; q < st.row_ptr[first + hw.first + 1]; ++q)
        PMCD_CHECK(pos_in_row(c, st.col[q]) >= 0, "hanging row ", first + hw.first, " column ", st.col[q],
                   " could not be added to row ", c, " because the structure does not include an entry here");
    const int k = launch_of(c);
    for(int p = st.row_ptr[c]; p < st.row_ptr[c + 1]; ++p) {
      const size_t before = msrc[k].size();
      for(const auto& hw : incoming[c]) {
        const int q = pos_in_row(first + hw.first, st.col[p]);
        if(q >= 0) { msrc[k].push_back(q); mw[k].push_back(hw.second); }
      }
      if(msrc[k].size() == before) continue;
      mdst[k].push_back(p);
      msp[k].push_back((int)before);
    }
    rdst[k].push_back(c);
    rsp[k].push_back((int)rsrc[k].size());
    for(const auto& hw : incoming[c]) { rsrc[k].push_back(first + hw.first); rw[k].push_back(hw.second); }
  }
  // concatenate the launches; source pointers become global
  std::vector<int> mat_dst, mat_sp, mat_src, rhs_dst, rhs_sp, rhs_src;
  std::vector<double> mat_w, rhs_w;
  L.hg_mat_launch.assign(1, 0);
  L.hg_rhs_launch.assign(1, 0);
  for(int k = 0; k < n_launch; ++k) {
    const int mo = (int)mat_src.size(), ro = (int)rhs_src.size();
    for(size_t i = 0; i < mdst[k].size(); ++i) { mat_dst.push_back(mdst[k][i]); mat_sp.push_back(mo + msp[k][i]); }
    mat_src.insert(mat_src.end(), msrc[k].begin(), msrc[k].end());
    mat_w.insert(mat_w.end(), mw[k].begin(), mw[k].end());
    for(size_t i = 0; i < rdst[k].size(); ++i) { rhs_dst.push_back(rdst[k][i]); rhs_sp.push_back(ro + rsp[k][i]); }
    rhs_src.insert(rhs_src.end(), rsrc[k].begin(), rsrc[k].end());
    rhs_w.insert(rhs_w.end(), rw[k].begin(), rw[k].end());
    L.hg_mat_launch.push_back((int)mat_dst.size());
    L.hg_rhs_launch.push_back((int)rhs_dst.size());
  }
  mat_sp.push_back((int)mat_src.size());
  rhs_sp.push_back((int)rhs_src.size());
  // the hanging rows: 1 on the diagonal, minus the weights at the coupling
  // DOFs, 0 elsewhere
  std::vector<int> fix_pos;
  std::vector<double> fix_val;
  for(int h = 0; h < nh; ++h) {
    const int r = first + h;
    for(int p = st.row_ptr[r]; p < st.row_ptr[r + 1]; ++p) {
      const int col = st.col[p];
      double v = col == r ? 1.0 : 0.0;
      for(int k = 0; k < sp.hang_n[h]; ++k)
        if(col == sp.hang_cpl[MC * h + k]) v = -sp.hang_w[MC * h + k];
      fix_pos.push_back(p);
      fix_val.push_back(v);
    }
  }
  L.hg_mat_dst.upload(mat_dst, stream);
  L.hg_mat_sp.upload(mat_sp, stream);
  L.hg_mat_src.upload(mat_src, stream);
  L.hg_mat_w.upload(mat_w, stream);
  L.hg_rhs_dst.upload(rhs_dst, stream);
  L.hg_rhs_sp.upload(rhs_sp, stream);
  L.hg_rhs_src.upload(rhs_src, stream);
  L.hg_rhs_w.upload(rhs_w, stream);
  L.hg_fix_pos.upload(fix_pos, stream);
  L.hg_fix_val.upload(fix_val, stream);
}

// Assemble2D.C:3350-3427: coupling of the assembled level, then the
// constraint rows and zero rhs entries of the hanging DOFs
void apply_hanging(Level& L, hipStream_t stream) {
  const int n_launch = (int)L.hg_mat_launch.size() - 1;
  for(int k = 0; k < n_launch; ++k) {
    const int m0 = L.hg_mat_launch[k], r0 = L.hg_rhs_launch[k];
    launch_hanging_gather(L.hg_mat_launch[k + 1] - m0, L.hg_mat_dst.p + m0, L.hg_mat_sp.p + m0, L.hg_mat_src.p,
                          L.hg_mat_w.p, L.val.p, L.hg_rhs_launch[k + 1] - r0, L.hg_rhs_dst.p + r0,
                          L.hg_rhs_sp.p + r0, L.hg_rhs_src.p, L.hg_rhs_w.p, L.rhs.p, stream);
  }
  launch_hanging_fix((int)L.hg_fix_pos.n, L.hg_fix_pos.p, L.hg_fix_val.p, L.val.p, L.space.n_active_non_hanging(),
                     L.space.n_hanging, L.rhs.p, stream);
}

void CDSystem::build_level_device(Level& L) {
  const Mesh& m = L.mesh;
  const FESpace& sp = L.space;
  const Structure& st = L.st;
  const ElementInfo& info = element_info(sp.element);
  // packed vertex coordinates per cell
  std::vector<double> xyz((size_t)m.n_cells * m.nv * 3);
  for(int c = 0; c < m.n_cells; ++c)
    for(int a = 0; a < m.nv; ++a)
      for(int k = 0; k < 3; ++k) xyz[((size_t)c * m.nv + a) * 3 + k] = m.coords[3 * m.cell(c)[a] + k];
  L.cell_xyz.upload(xyz, stream_);
  L.dof.upload(sp.dof, stream_);
  L.row_ptr.upload(st.row_ptr, stream_);
  L.col.upload(st.col, stream_);
  L.diag_pos.upload(st.diag, stream_);
  L.contrib_ptr.upload(st.contrib_ptr, stream_);
  L.contrib.upload(st.contrib, stream_);
  L.slot.upload(st.slot, stream_);
  boundary_data(L);
  const int n = sp.n_dof;
  L.val.alloc(st.nnz);
  L.rhs.alloc(n);
  L.sol.alloc(n);
  L.diag.alloc(n);
  L.mg_sol.alloc(n); L.mg_sol2.alloc(n); L.mg_rhs.alloc(n); L.mg_def.alloc(n); L.mg_tmp.alloc(n);
  if(sp.n_hanging) hanging_plan(L, stream_);
}

// P[f][c] = value of coarse basis function c at the nodal point of fine DOF f
// (the interpolation the reference's GridTransfer::Prolongate applies cell by
// cell and averages, src/Multigrid/GridTransfer.C:603-865); R = P^T
// (DefectRestriction, :866-1115).
void CDSystem::build_transfer(Level& F, const Level& C) {
  const Mesh& fm = F.mesh;
  const Mesh& cm = C.mesh;
  const ElementInfo& fi = element_info(F.space.element);
  const ElementInfo& ci = element_info(C.space.element);
  const int nf = F.n(), ncd = C.n();
  std::vector<std::vector<std::pair<int, double>>> rows(nf);
  std::vector<char> done(nf, 0);
  std::vector<double> phi(kMaxLocalDof * kNDeriv3D);
  for(int c = 0; c < fm.n_cells; ++c) {
    const int p = fm.parent[c];
    PMCD_CHECK(p >= 0 && p < cm.n_cells, "level hierarchy mismatch");
    const int* cv = cm.cell(p);
    // inverse affine map of the parent cell
    const int dim = cm.dim();
    double o[3], J[3][3];
    if(shape_is_simplex(cm.shape)) {
      for(int k = 0; k < 3; ++k) {
        o[k] = cm.coords[3 * cv[0] + k];
        for(int a = 0; a < dim; ++a) J[k][a] = cm.coords[3 * cv[a + 1] + k] - o[k];
      }
    } else {
      // x = x0 + J (xi + 1) with the affine map of vertices 0, 1, 3 (, 4)
      const int e[3] = {1, 3, 4};
      for(int k = 0; k < 3; ++k) {
        o[k] = cm.coords[3 * cv[0] + k];
        for(int a = 0; a < dim; ++a) J[k][a] = (cm.coords[3 * cv[e[a]] + k] - o[k]) / 2;
      }
    }
    // invert J (dim x dim)
    double Ji[3][3] = {{0}};
    if(dim == 2) {
      double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      Ji[0][0] = J[1][1] / det; Ji[0][1] = -J[0][1] / det;
      Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
    } else {
      double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                   J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
      Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
      Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
      Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
      Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    }
    for(int i = 0; i < fi.n_dof; ++i) {
      const int f = F.space.cell_dof(c)[i];
      if(done[f]) continue;
      done[f] = 1;
      double x[3];
      node_position(fm, c, fi, i, x);
      double d[3], rp[3] = {0, 0, 0};
      for(int k = 0; k < 3; ++k) d[k] = x[k] - o[k];
      for(int a = 0; a < dim; ++a) {
        double s = 0;
        for(int k = 0; k < dim; ++k) s += Ji[a][k] * d[k];
        rp[a] = shape_is_simplex(cm.shape) ? s : s - 1.0;
      }
      eval_basis(C.space.element, rp, phi.data());
      for(int l = 0; l < ci.n_dof; ++l) {
        double v = phi[l];
        if(std::fabs(v) > 1e-13) rows[f].push_back({C.space.cell_dof(p)[l], v});
      }
    }
  }
  std::vector<int> rp(nf + 1, 0), col;
  std::vector<double> val;
  for(int f = 0; f < nf; ++f) {
    auto& r = rows[f];
    std::sort(r.begin(), r.end());
    for(auto& e : r) { col.push_back(e.first); val.push_back(e.second); }
    rp[f + 1] = (int)col.size();
  }
  set_transfer(F, ncd, rp, col, val);
}

// The sliced ELL copy of the first `rows` rows of a transfer matrix (CSR,
// columns < n_cols) and the column footprints of its 256-row blocks
static void build_sell_matrix(SellMatrix& S, int rows, int n_cols, const std::vector<int>& rp,
                              const std::vector<int>& col, const std::vector<double>& val, hipStream_t stream) {
  S.slots = 0;
  S.fp_max = 0;
  if(rows <= 0) return;
  Structure st;
  st.n_rows = (int)rp.size() - 1;
  st.row_ptr = rp;
  st.col = col;
  const int window = 4 * kSellRows;
  SellLayout o = build_sell(st, rows, window);
  const long long entries = std::max<long long>(1, o.slice_off.back());
  std::vector<int> scol(entries, 0);
  std::vector<double> sval(entries, 0.0);
  for(int t = 0; t < o.n_slots(); ++t) {
    const int r = o.perm[t];
    if(r < 0) continue;
    const long long base = o.slice_off[t / kSellRows] + t % kSellRows;
    for(int q = rp[r]; q < rp[r + 1]; ++q) {
      scol[base + (long long)(q - rp[r]) * kSellRows] = col[q];
      sval[base + (long long)(q - rp[r]) * kSellRows] = val[q];
    }
  }
  S.off.upload(o.slice_off, stream);
  S.perm.upload(o.perm, stream);
  S.len.upload(o.len, stream);
  S.col.upload(scol, stream);
  S.val.upload(sval, stream);
  std::vector<float> s32(sval.begin(), sval.end());
  S.val32.upload(s32, stream);
  SellFootprint f = build_sell_footprint(st, o, window, kSellMaxFootprint, n_cols);
  if(!f.fp_off.empty() && f.fp_max > 0) {
    f.lcol.resize(entries, 0);
    S.lcol.upload(f.lcol, stream);
    S.fp_off.upload(f.fp_off, stream);
    S.fp.upload(f.fp, stream);
    S.fp_max = f.fp_max;
  }
  S.slots = o.n_slots();
  // the host vectors are freed on return: wait for their uploads
  PMCD_HIP(hipStreamSynchronize(stream));
}

template <class T>
static SellView<T> sell_matrix_view(const SellMatrix& S, const T* val) {
  const bool fp = S.fp_max > 0;
  return SellView<T>{S.slots, S.off.p, S.perm.p, S.len.p, S.col.p, val, 4 * kSellRows,
                     fp ? S.lcol.p : nullptr, fp ? S.fp_off.p : nullptr, fp ? S.fp.p : nullptr,
                     fp ? S.fp_max : 0};
}

// the sliced ELL prolongation (tuning key 19, default on) goes with the fine
// level's sliced ELL smoother. The restriction keeps the CSR kernel: R's rows
// are long (up to 125 entries on Q2 hexahedra) and few, and their fine-level
// footprints are too wide for LDS, so one lane per row loses to 16 (C2:
// 60 us against 50 us below the finest level, 24.5 against 8.8 two levels down).
static bool transfer_uses_sell(const Level& F) {
  return get_tuning(19) == 1 && get_tuning(3) == 4 && F.n_active() >= sell_min_rows() && F.p_sell.slots > 0;
}

void CDSystem::prolong_add(Level& F, const double* xc, double* y, double a) {
  if(transfer_uses_sell(F))
    launch_spmv_sell<double>(F.n_active(), F.n_active(), sell_matrix_view<double>(F.p_sell, F.p_sell.val.p), xc, y,
                             y, a, 1.0, stream_);
  else
    launch_prolong_add(F.n(), F.n_active(), (long)F.p_val.n, F.p_max_row, F.p_rp.p, F.p_col.p, F.p_val.p, xc, y, a,
                       stream_);
}

void CDSystem::restrict_to(Level& F, const Level& C, const double* x, double* y) {
  launch_restrict(C.n(), C.n_active(), (long)F.r_val.n, F.r_max_row, F.r_rp.p, F.r_col.p, F.r_val.p, x, y, stream_);
}

void CDSystem::prolong_add32(Level& F, const float* xc, float* y, float a) {
  if(transfer_uses_sell(F))
    launch_spmv_sell<float>(F.n_active(), F.n_active(), sell_matrix_view<float>(F.p_sell, F.p_sell.val32.p), xc, y,
                            y, a, 1.0f, stream_);
  else
    launch_prolong_add_f32(F.n(), F.n_active(), (long)F.p_val32.n, F.p_max_row, F.p_rp.p, F.p_col.p,
                           F.p_val32.p, xc, y, a, stream_);
}

void CDSystem::restrict_to32(Level& F, const Level& C, const float* x, float* y) {
  launch_restrict_f32(C.n(), C.n_active(), (long)F.r_val32.n, F.r_max_row, F.r_rp.p, F.r_col.p, F.r_val32.p, x, y,
                      stream_);
}

void CDSystem::apply_transfer(int level, int which, const double* x, double* y, double a) {
  const int nl = n_levels();
  const int l = level < 0 ? nl + level : level;
  PMCD_CHECK(l >= 1 && l < nl, "level ", level, " has no coarser level");
  PMCD_CHECK(which == 0 || which == 1, "transfer must be 0 (prolongation) or 1 (restriction)");
  Level& F = *levels_[l];
  const Level& C = *levels_[l - 1];
  const int nx = which == 0 ? C.n() : F.n(), ny = which == 0 ? F.n() : C.n();
  DevBuf<double> dx(std::max(1, nx)), dy(std::max(1, ny));
  dx.upload(x, nx, stream_);
  dy.upload(y, ny, stream_);
  if(which == 0)
    prolong_add(F, dx.p, dy.p, a);
  else
    restrict_to(F, C, dx.p, dy.p);
  dy.download(y, ny, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

// P (n_fine x ncd, CSR) and its transpose R on the device, FP64 and FP32
void CDSystem::set_transfer(Level& F, int ncd, const std::vector<int>& rp, const std::vector<int>& col,
                            const std::vector<double>& val) {
  const int nf = (int)rp.size() - 1;
  // transpose
  std::vector<int> trp(ncd + 1, 0), tcol(col.size());
  std::vector<double> tval(col.size());
  for(int c : col) trp[c + 1]++;
  for(int c = 0; c < ncd; ++c) trp[c + 1] += trp[c];
  std::vector<int> fill(trp.begin(), trp.end() - 1);
  for(int f = 0; f < nf; ++f)
    for(int k = rp[f]; k < rp[f + 1]; ++k) {
      int pos = fill[col[k]]++;
      tcol[pos] = f;
      tval[pos] = val[k];
    }
  F.n_coarse = ncd;
  F.p_max_row = 0;
  for(int f = 0; f < nf; ++f) F.p_max_row = std::max(F.p_max_row, rp[f + 1] - rp[f]);
  F.r_max_row = 0;
  for(int c = 0; c < ncd; ++c) F.r_max_row = std::max(F.r_max_row, trp[c + 1] - trp[c]);
  F.p_rp.upload(rp, stream_); F.p_col.upload(col, stream_); F.p_val.upload(val, stream_);
  F.r_rp.upload(trp, stream_); F.r_col.upload(tcol, stream_); F.r_val.upload(tval, stream_);
  std::vector<float> v32(val.begin(), val.end()), t32(tval.begin(), tval.end());
  F.p_val32.upload(v32, stream_);
  F.r_val32.upload(t32, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  build_sell_matrix(F.p_sell, F.n_active(), ncd, rp, col, val, stream_);
}

AsmArgs CDSystem::asm_args(const Level& L) {
  AsmArgs a;
  a.shape = (int)L.mesh.shape;
  a.dim = L.mesh.dim();
  a.nv = L.mesh.nv;
  a.nl = L.space.n_local;
  a.nq = quad_formula(rule_).n();
  a.ncomp = ncomp_;
  a.supg = desc_.supg;
  a.n_cells = L.mesh.n_cells;
  a.ex = desc_.ex;
  a.ex.dim = a.dim;
  a.ref = ref_.p;
  a.qw = qw_.p;
  a.qp = qp_.p;
  a.tensor = tensor_.n && example_constant_matrix_coeffs(desc_.ex.kind) ? tensor_.p : nullptr;
  a.cell_xyz = L.cell_xyz.p;
  a.local_A = local_A_.p;
  a.local_rhs = local_rhs_.p;
  return a;
}

void CDSystem::assemble_level(Level& L) {
  AsmArgs a = asm_args(L);
  assemble_cells(a, stream_);
  // a level whose matrix kernels run on the sliced ELL copy gets it from the
  // same row gather (no separate copy pass)
  const bool sell = level_uses_sell(L);
  if(sell) build_level_sell(L, stream_);
  launch_gather_matrix(L.n(), L.n_active(), a.nl, L.row_ptr.p, L.col.p, L.contrib_ptr.p, L.contrib.p, L.slot.p,
                       local_A_.p, L.val.p, stream_, sell ? L.sell_slot.p : nullptr,
                       sell ? L.sell_off.p : nullptr, sell ? L.sell_val.p : nullptr);
  L.sell_fresh = sell;
  launch_gather_rhs(L.n(), L.n_active(), L.contrib_ptr.p, L.contrib.p, local_rhs_.p,
                    L.extra_ptr.p, L.extra.p, L.dir_vals.p, L.rhs.p, stream_);
  if(L.space.n_hanging) {
    apply_hanging(L, stream_);
    L.sell_fresh = false;  // the fused sliced ELL copy predates the coupling
  }
  launch_gather_idx(L.n(), L.diag_pos.p, L.val.p, L.diag.p, stream_);
}

void CDSystem::assemble() {
  PMCD_CHECK(!operator_only_, "an operator-only system has no mesh to assemble on");
  PMCD_HIP(hipStreamSynchronize(stream_));
  const double t0 = now_s();
  for(auto& L : levels_) assemble_level(*L);
  // the solution starts at zero with the Dirichlet values copied from the rhs
  // (ConvectionDiffusion.C:443, BlockVector::copy_nonactive)
  Level& F = finest();
  launch_fill(F.n_active(), 0.0, F.sol.p, stream_);
  launch_copy(F.n() - F.n_active(), F.rhs.p + F.n_active(), F.sol.p + F.n_active(), stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  last_assemble_ms = (now_s() - t0) * 1e3;
  coarse_ready_ = false;
  mixed_ready_ = false;
  for(auto& L : levels_) L->sw_fresh = L->sw_fresh32 = L->sell_fresh32 = false;
}

void CDSystem::coarse_factorize() {
  Level& C = *levels_.front();
  const int n = C.n();
  PMCD_CHECK(n <= 20000, "coarse direct solve limited to 20000 unknowns, got ", n);
  // work space kept for the next factorization (same size)
  if(coarse_dense_.n != (size_t)n * n) coarse_dense_.alloc((size_t)n * n);
  if(coarse_colbuf_.n != 2 * (size_t)n) coarse_colbuf_.alloc(2 * (size_t)n);
  if(coarse_perm_.n != (size_t)n) { coarse_perm_.alloc(n); coarse_colmap_.alloc(n); }
  if(coarse_inv_.n != (size_t)n * n) coarse_inv_.alloc((size_t)n * n);
  launch_csr_to_dense(n, C.row_ptr.p, C.col.p, C.val.p, coarse_dense_.p, stream_);
  // rows n_active.. of a level matrix are identity rows (Dirichlet rows)
  dense_inverse(n, C.n_active(), coarse_dense_.p, coarse_inv_.p, coarse_colbuf_.p, coarse_perm_.p,
                coarse_colmap_.p, stream_);
  if(coarse_tmp_.n != 2 * (size_t)n) coarse_tmp_.alloc(2 * (size_t)n);
  PMCD_HIP(hipStreamSynchronize(stream_));
  coarse_ready_ = true;
}

// HIP-event timing of one kernel class on this system's stream.
void CDSystem::time_kernel(int which, int level, int iters, double* avg_ms) {
  if(level < 0) level += n_levels();
  PMCD_CHECK(level >= 0 && level < n_levels(), "level out of range");
  Level& L = *levels_[level];
  const int n = L.n();
  DevBuf<double> x(n), y(n), b(n);
  launch_fill(n, 1.0, x.p, stream_);
  launch_fill(n, 0.5, b.p, stream_);
  DevBuf<float> xf, yf, bf;
  if(which == 2 || which == 5) {
    PMCD_CHECK(L.val32.n == L.val.n, "FP32 matrix copies not prepared");
    xf.alloc(n); yf.alloc(n); bf.alloc(n);
    launch_fill_f32(n, 1.0f, xf.p, stream_);
    launch_fill_f32(n, 0.5f, bf.p, stream_);
  }
  DevBuf<double> copy_dst;
  if(which == 6 || which == 7) copy_dst.alloc(L.val.n & ~(size_t)1);
  if(which == 8) copy_dst.alloc((size_t)kStreamReadBlocks * 4);
  AsmArgs a = asm_args(L);
  auto launch = [&]() {
    switch(which) {
      case 0: level_spmv(L, x.p, nullptr, y.p, 1.0, 0.0, stream_); break;
      case 1: level_jacobi(L, b.p, x.p, y.p, 0.8, stream_); break;
      case 2: level_spmv32(L, xf.p, nullptr, yf.p, 1.0f, 0.0f, stream_); break;
      case 3:
        assemble_cells(a, stream_);
        break;
      case 4: {
        const bool sell = level_uses_sell(L) && L.sell_slots >= 0;
        launch_gather_matrix(n, L.n_active(), a.nl, L.row_ptr.p, L.col.p, L.contrib_ptr.p, L.contrib.p,
                             L.slot.p, local_A_.p, L.val.p, stream_, sell ? L.sell_slot.p : nullptr,
                             sell ? L.sell_off.p : nullptr, sell ? L.sell_val.p : nullptr);
        break;
      }
      case 5: level_jacobi32(L, bf.p, xf.p, yf.p, 0.8f, stream_); break;
      case 6: launch_stream_copy((long)copy_dst.n, L.val.p, copy_dst.p, stream_); break;
      case 7: launch_stream_copy4((long)copy_dst.n, L.val.p, copy_dst.p, stream_); break;
      case 8: launch_stream_read((long)(L.val.n & ~(size_t)1), L.val.p, copy_dst.p, stream_); break;
      default: PMCD_FAIL("unknown kernel id ", which);
    }
  };
  launch();  // warm-up
  hipEvent_t e0, e1;
  PMCD_HIP(hipEventCreate(&e0));
  PMCD_HIP(hipEventCreate(&e1));
  PMCD_HIP(hipEventRecord(e0, stream_));
  for(int i = 0; i < iters; ++i) launch();
  PMCD_HIP(hipEventRecord(e1, stream_));
  PMCD_HIP(hipEventSynchronize(e1));
  float ms = 0;
  PMCD_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = ms / iters;
}

void CDSystem::apply(Level& L, const double* x, double* y) {
  level_spmv(L, x, nullptr, y, 1.0, 0.0, stream_);
}

void CDSystem::apply_scaled_add(Level& L, const double* x, double* y, double a) {
  level_spmv(L, x, y, y, a, 1.0, stream_);
}

double CDSystem::dot(int n, const double* x, const double* y) {
  launch_dot(n, x, y, red_partial_.p, red_out_.p, stream_);
  double h;
  PMCD_HIP(hipMemcpyAsync(&h, red_out_.p, sizeof(double), hipMemcpyDeviceToHost, stream_));
  PMCD_HIP(hipStreamSynchronize(stream_));
  return h;
}

// `steps` damped Jacobi sweeps ping-ponging between x[0] and x[1]; returns the
// index of the buffer holding the result. With `from_zero` the iterate is 0
// and b vanishes on the non-active rows, so the first sweep needs no SpMV.
int CDSystem::smooth(Level& L, double* const x[2], int cur, const double* b, int steps,
                     bool from_zero, int kind) {
  if(kind < 0) kind = sd_.smoother;
  if(kind != 0) {
    // SORSmoother::smooth (SORSmoother.C:29-33): one sweep (pair) per step, in place
    for(int s = 0; s < steps; ++s) sor_sweep(L, b, x[cur], sd_.sor_omega, kind == 1 ? 0 : 2);
    return cur;
  }
  for(int s = 0; s < steps; ++s) {
    if(s == 0 && from_zero)
      launch_jacobi_from_zero(L.n(), L.n_active(), L.diag.p, b, x[1 - cur], sd_.jacobi_damp, stream_);
    else
      level_jacobi(L, b, x[cur], x[1 - cur], sd_.jacobi_damp, stream_);
    cur = 1 - cur;
  }
  return cur;
}

int CDSystem::smooth32(Level& L, float* const x[2], int cur, const float* b, int steps,
                       bool from_zero) {
  if(sd_.smoother != 0) {
    for(int s = 0; s < steps; ++s)
      sor_sweep32(L, b, x[cur], (float)sd_.sor_omega, sd_.smoother == 1 ? 0 : 2);
    return cur;
  }
  for(int s = 0; s < steps; ++s) {
    if(s == 0 && from_zero)
      launch_jacobi_from_zero_f32(L.n(), L.n_active(), L.diag32.p, b, x[1 - cur],
                                  (float)sd_.jacobi_damp, stream_);
    else
      level_jacobi32(L, b, x[cur], x[1 - cur], (float)sd_.jacobi_damp, stream_);
    cur = 1 - cur;
  }
  return cur;
}

// Multigrid::cycle with a V-cycle control (Multigrid.C:157-308): finest
// iterate 0 with the non-active entries of the rhs, pre-smoothing and defect
// restriction down to the coarsest level, direct coarse solve, damped
// prolongation and post-smoothing on the way up. The finest iterate lives in
// `out` and one work buffer; coarser iterates start at 0 with a rhs that is 0
// on the non-active rows.
std::vector<int> cycle_directions(int type, int n_levels) {
  PMCD_CHECK(type >= 0 && type <= 2, "unknown multigrid cycle ", type);
  PMCD_CHECK(n_levels >= 1 && n_levels <= 16, "multigrid levels out of range");
  std::vector<int> rec(n_levels, type == 0 ? 1 : 2);
  rec[n_levels - 1] = 1;
  std::vector<int> dirs{kMgDummy};
  const int finest = n_levels - 1;
  std::function<void(int)> fill = [&](int level) {
    if(level == 0) {
      dirs.push_back(kMgUp);
      return;
    }
    for(int j = 0; j < rec[level]; ++j) {
      dirs.push_back(kMgDown);
      fill(level - 1);
    }
    if(type == 2) rec[level] = 1;
    dirs.push_back(level == finest ? kMgEnd : kMgUp);
  };
  fill(finest);
  return dirs;
}

// The coarsest level's loop with an iterative coarse smoother (Multigrid.C:
// 200-240): the residual starts at 1e10, so at least one step runs; then
// smooth and recompute |b - A x| until it is at most coarse_residual or
// coarse_max_iter steps are done. x[cur] holds the start iterate (0).
int CDSystem::coarse_iterate(const double* b, double* const x[2], int cur) {
  Level& C = *levels_[0];
  const int kind = sd_.coarse_smoother - 1;
  double res = 1e10;
  for(int i = 0; i < sd_.coarse_max_iter && res > sd_.coarse_residual; ++i) {
    cur = smooth(C, x, cur, b, 1, false, kind);
    level_spmv(C, x[cur], b, C.mg_def.p, -1.0, 1.0, stream_);
    res = std::sqrt(dot(C.n(), C.mg_def.p, C.mg_def.p));
  }
  return cur;
}

// One multigrid cycle z = MG(v) (Multigrid::cycle / cycle_step,
// Multigrid.C:157-308): the directions of cycle_directions() walked from the
// finest level; going down pre-smooths, computes the defect, restarts the
// coarser iterate at 0 and restricts; the coarsest level is the direct solve;
// coming up prolongates with damping and post-smooths. Jacobi sweeps
// ping-pong between two buffers per level; a level whose iterate is known to
// be 0 takes its first Jacobi sweep without the SpMV.
void CDSystem::vcycle64(const double* rhs, double* out) {
  const int Lf = n_levels() - 1;
  std::vector<const double*> b(n_levels());
  std::vector<std::array<double*, 2>> x(n_levels());
  std::vector<int> cur(n_levels(), 0);
  for(int l = 0; l <= Lf; ++l) {
    b[l] = levels_[l]->mg_rhs.p;
    x[l] = {levels_[l]->mg_sol.p, levels_[l]->mg_sol2.p};
  }
  b[Lf] = rhs;
  x[Lf] = {out, levels_[Lf]->mg_sol2.p};
  // the first Jacobi sweep may skip its SpMV only if it starts from the zero
  // iterate; SOR updates in place and needs the iterate initialised
  const bool jacobi = sd_.smoother == 0;
  const bool fine_zero = jacobi && rhs_nonactive_zero_ && sd_.n_pre > 0;
  if(!fine_zero) {
    Level& F = finest();
    launch_fill(F.n_active(), 0.0, x[Lf][0], stream_);
    launch_copy(F.n() - F.n_active(), rhs + F.n_active(), x[Lf][0] + F.n_active(), stream_);
  }
  if(n_levels() == 1) {
    cur[0] = smooth(*levels_[0], x[0].data(), cur[0], b[0], sd_.n_pre + sd_.n_post, false);
    if(cur[0] != 0) launch_copy(finest().n(), x[0][cur[0]], out, stream_);
    return;
  }
  auto prolong = [&](int l) {  // damped addition of P x_{l-1} on level l
    prolong_add(*levels_[l], x[l - 1][cur[l - 1]], x[l][cur[l]], sd_.mg_damp);
  };
  std::vector<char> zero(n_levels(), 0);
  zero[Lf] = fine_zero;
  const std::vector<int> dirs = cycle_directions(sd_.cycle, n_levels());
  int level = Lf;
  for(size_t i = 1; i < dirs.size(); ++i) {
    if(level == 0) {
      Level& C = *levels_[0];
      if(sd_.coarse_smoother == 0) {
        launch_gemv(C.n(), coarse_inv_.p, b[0], x[0][0], stream_);
        cur[0] = 0;
      } else {
        cur[0] = coarse_iterate(b[0], x[0].data(), cur[0]);
      }
      prolong(1);
      zero[1] = 0;
      level = 1;
      continue;
    }
    Level& L = *levels_[level];
    if(dirs[i - 1] == kMgUp) cur[level] = smooth(L, x[level].data(), cur[level], b[level], sd_.n_post, false);
    if(dirs[i] == kMgDown) {
      Level& C = *levels_[level - 1];
      cur[level] = smooth(L, x[level].data(), cur[level], b[level], sd_.n_pre, zero[level] != 0);
      zero[level] = 0;
      level_spmv(L, x[level][cur[level]], b[level], L.mg_def.p, -1.0, 1.0, stream_);
      // the coarser iterate restarts at 0; a Jacobi sweep from zero does not read it
      const bool from_zero = jacobi && sd_.n_pre > 0;
      if(!from_zero || level - 1 == 0) {
        launch_fill(C.n(), 0.0, x[level - 1][0], stream_);
        cur[level - 1] = 0;
      }
      zero[level - 1] = from_zero;
      restrict_to(L, C, L.mg_def.p, C.mg_rhs.p);
      --level;
    } else if(dirs[i] == kMgUp) {
      prolong(level + 1);
      zero[level + 1] = 0;
      ++level;
    } else {
      break;
    }
  }
  if(cur[Lf] != 0) launch_copy(finest().n(), x[Lf][cur[Lf]], out, stream_);
}

// The same cycle on the FP32 copies (C3): the coarse direct solve stays FP64.
void CDSystem::vcycle32(const double* rhs, double* out) {
  const int Lf = n_levels() - 1;
  for(auto& Lp : levels_)
    PMCD_CHECK(Lp->f_sol.n == (size_t)Lp->n(), "FP32 work vectors not prepared");
  std::vector<std::array<float*, 2>> x(n_levels());
  std::vector<int> cur(n_levels(), 0);
  for(int l = 0; l <= Lf; ++l) x[l] = {levels_[l]->f_sol.p, levels_[l]->f_sol2.p};
  Level& F = finest();
  launch_d2f(F.n(), rhs, F.f_rhs.p, stream_);
  const bool jacobi = sd_.smoother == 0;
  const bool fine_zero = jacobi && rhs_nonactive_zero_ && sd_.n_pre > 0;
  if(!fine_zero) {
    launch_fill_f32(F.n_active(), 0.0f, F.f_sol.p, stream_);
    PMCD_HIP(hipMemcpyAsync(F.f_sol.p + F.n_active(), F.f_rhs.p + F.n_active(),
                            sizeof(float) * (F.n() - F.n_active()), hipMemcpyDeviceToDevice, stream_));
  }
  if(n_levels() == 1) {
    Level& C = *levels_[0];
    cur[0] = smooth32(C, x[0].data(), cur[0], C.f_rhs.p, sd_.n_pre + sd_.n_post, false);
    launch_f2d(F.n(), x[Lf][cur[Lf]], out, stream_);
    return;
  }
  auto prolong = [&](int l) {
    prolong_add32(*levels_[l], x[l - 1][cur[l - 1]], x[l][cur[l]], (float)sd_.mg_damp);
  };
  std::vector<char> zero(n_levels(), 0);
  zero[Lf] = fine_zero;
  const std::vector<int> dirs = cycle_directions(sd_.cycle, n_levels());
  int level = Lf;
  for(size_t i = 1; i < dirs.size(); ++i) {
    if(level == 0) {
      Level& C = *levels_[0];
      if(sd_.coarse_smoother == 0) {
        launch_f2d(C.n(), C.f_rhs.p, coarse_tmp_.p, stream_);
        launch_gemv(C.n(), coarse_inv_.p, coarse_tmp_.p, coarse_tmp_.p + C.n(), stream_);
        launch_d2f(C.n(), coarse_tmp_.p + C.n(), x[0][0], stream_);
      } else {
        // the coarse loop runs in FP64 on the level's FP64 work vectors
        launch_f2d(C.n(), C.f_rhs.p, C.mg_rhs.p, stream_);
        launch_fill(C.n(), 0.0, C.mg_sol.p, stream_);
        double* const xc[2] = {C.mg_sol.p, C.mg_sol2.p};
        const int c = coarse_iterate(C.mg_rhs.p, xc, 0);
        launch_d2f(C.n(), xc[c], x[0][0], stream_);
      }
      cur[0] = 0;
      prolong(1);
      zero[1] = 0;
      level = 1;
      continue;
    }
    Level& L = *levels_[level];
    if(dirs[i - 1] == kMgUp) cur[level] = smooth32(L, x[level].data(), cur[level], L.f_rhs.p, sd_.n_post, false);
    if(dirs[i] == kMgDown) {
      Level& C = *levels_[level - 1];
      cur[level] = smooth32(L, x[level].data(), cur[level], L.f_rhs.p, sd_.n_pre, zero[level] != 0);
      zero[level] = 0;
      level_spmv32(L, x[level][cur[level]], L.f_rhs.p, L.f_def.p, -1.0f, 1.0f, stream_);
      const bool from_zero = jacobi && sd_.n_pre > 0;
      if(!from_zero || level - 1 == 0) {
        launch_fill_f32(C.n(), 0.0f, x[level - 1][0], stream_);
        cur[level - 1] = 0;
      }
      zero[level - 1] = from_zero;
      restrict_to32(L, C, L.f_def.p, C.f_rhs.p);
      --level;
    } else if(dirs[i] == kMgUp) {
      prolong(level + 1);
      zero[level + 1] = 0;
      ++level;
    } else {
      break;
    }
  }
  launch_f2d(F.n(), x[Lf][cur[Lf]], out, stream_);
}

void CDSystem::vcycle_eager(const double* rhs, double* out) {
  if(sd_.mixed) vcycle32(rhs, out);
  else vcycle64(rhs, out);
}

// The V-cycle is a fixed kernel sequence for given solver parameters, launch
// shapes and buffers: it is captured once per (rhs, out, zero-tail) and
// replayed as one graph launch. Graphs are dropped when the parameters or the
// launch shapes change; matrix values are read from the same buffers, so
// reassembly keeps them valid. The derived copies (sliced ELL, class-ordered
// SOR values, FP32 values) are refreshed by prepare() before any cycle runs,
// never inside a capture, so a captured graph holds no refresh node and stays
// valid after the matrix changes. The sequence holds no memset nodes.
void CDSystem::vcycle(const double* rhs, double* out) {
  // (a coarse smoother loop reads its residual on the host: not capturable)
  if(!vcycle_graphs() || sd_.coarse_smoother != 0) {
    vcycle_eager(rhs, out);
    return;
  }
  auto same = [](const SolverDesc& a, const SolverDesc& b) {
    return a.smoother == b.smoother && a.n_pre == b.n_pre && a.n_post == b.n_post &&
           a.jacobi_damp == b.jacobi_damp && a.mg_damp == b.mg_damp && a.sor_omega == b.sor_omega &&
           a.mixed == b.mixed && a.cycle == b.cycle;
  };
  if(graph_epoch_ != tuning_epoch() || !same(graph_sd_, sd_) || vgraphs_.size() > 128) {
    clear_graphs();
    graph_epoch_ = tuning_epoch();
    graph_sd_ = sd_;
  }
  const bool fz = rhs_nonactive_zero_;
  for(const auto& g : vgraphs_)
    if(g.rhs == rhs && g.out == out && g.fine_zero == fz) {
      PMCD_HIP(hipGraphLaunch(g.exec, stream_));
      return;
    }
  PMCD_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  try {
    vcycle_eager(rhs, out);
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
  vgraphs_.push_back({rhs, out, fz, exec});
  PMCD_HIP(hipGraphLaunch(exec, stream_));
}

bool CDSystem::tail_is_zero(const double* v) {
  Level& F = finest();
  const int m = F.n() - F.n_active();
  if(m == 0) return true;
  return dot(m, v + F.n_active(), v + F.n_active()) == 0.0;
}

// Preconditioner application: none (copy), Jacobi (r = z / diag,
// Iteration_jacobi.C:168-182) or one multigrid cycle.
void CDSystem::precond(const double* v, double* z) {
  Level& F = finest();
  switch(sd_.precond) {
    case 0: launch_copy(F.n(), v, z, stream_); break;
    case 1: launch_div(F.n(), v, F.diag.p, z, stream_); break;
    case 2: vcycle(v, z); break;
    case 3:
    case 4:
      // Iteration_sor::apply (Iteration_sor.C:125-140): sweep(s) from zero
      launch_fill(F.n(), 0.0, z, stream_);
      sor_sweep(F, v, z, sd_.sor_omega, sd_.precond == 3 ? 0 : 2);
      break;
    default: PMCD_FAIL("unknown preconditioner ", sd_.precond);
  }
}

void CDSystem::prepare_sweep(Level& L) { prepare_level_sweep(L, stream_); }

void CDSystem::sor_sweep(Level& L, const double* b, double* x, double omega, int flag) {
  level_sor_sweep(L, b, x, omega, flag, stream_);
}

void CDSystem::sor_sweep32(Level& L, const float* b, float* x, float omega, int flag) {
  level_sor_sweep32(L, b, x, omega, flag, stream_);
}

void build_level_sell(Level& L, hipStream_t stream) {
  if(L.sell_slots >= 0) return;
  L.sell_window = sell_window();
  SellLayout o = build_sell(L.st, L.n_active(), L.sell_window);
  PMCD_CHECK(o.slice_off.back() < (1LL << 40), "sliced ELL layout too large");
  std::vector<int> slot_of(std::max(1, L.n_active()), 0);
  for(int t = 0; t < o.n_slots(); ++t)
    if(o.perm[t] >= 0) slot_of[o.perm[t]] = t;
  L.sell_off.upload(o.slice_off, stream);
  L.sell_perm.upload(o.perm, stream);
  L.sell_len.upload(o.len, stream);
  L.sell_slot.upload(slot_of, stream);
  L.sell_col.alloc(std::max<long long>(1, o.slice_off.back()));
  L.sell_val.alloc(L.sell_col.n);
  L.sell_slots = o.n_slots();
  PMCD_HIP(hipMemsetAsync(L.sell_col.p, 0, L.sell_col.n * sizeof(int), stream));
  PMCD_HIP(hipMemsetAsync(L.sell_val.p, 0, L.sell_val.n * sizeof(double), stream));
  launch_sell_fill<int>(L.sell_slots, L.sell_off.p, L.sell_perm.p, L.sell_len.p, L.row_ptr.p, L.col.p,
                        L.sell_col.p, stream);
  L.sell_fresh = L.sell_fresh32 = false;
}

// the column footprints of the layout's 256-slot blocks, built when a
// launch mode first asks for them
static void build_level_footprint(Level& L, hipStream_t stream) {
  if(L.sell_fp_max >= 0) return;
  L.sell_fp_max = 0;
  if(L.sell_window != 4 * kSellRows) return;
  SellFootprint f = build_sell_footprint(L.st, build_sell(L.st, L.n_active(), L.sell_window), 4 * kSellRows,
                                         kSellMaxFootprint);
  if(f.fp_off.empty() || f.fp_max == 0) return;
  PMCD_CHECK((long long)f.lcol.size() == (long long)L.sell_col.n || L.sell_col.n == 1, "footprint layout mismatch");
  L.sell_lcol.upload(f.lcol, stream);
  L.sell_fp_off.upload(f.fp_off, stream);
  L.sell_fp.upload(f.fp, stream);
  L.sell_fp_max = f.fp_max;
}

void prepare_level_sell(Level& L, bool f32, hipStream_t stream) {
  build_level_sell(L, stream);
  if(sell_mode() & 16) build_level_footprint(L, stream);
  if(!L.sell_fresh) {
    launch_sell_fill<double>(L.sell_slots, L.sell_off.p, L.sell_perm.p, L.sell_len.p, L.row_ptr.p, L.val.p,
                             L.sell_val.p, stream);
    L.sell_fresh = true;
  }
  if(f32 && !L.sell_fresh32) {
    // the FP32 copy rounds the FP64 one entry by entry (padding stays 0)
    PMCD_CHECK(L.sell_val.n < (size_t)INT_MAX, "sliced ELL copy too large for one conversion launch");
    if(L.sell_val32.n != L.sell_val.n) L.sell_val32.alloc(L.sell_val.n);
    launch_d2f((int)L.sell_val.n, L.sell_val.p, L.sell_val32.p, stream);
    L.sell_fresh32 = true;
  }
}

template <class T>
static SellView<T> sell_view(const Level& L, const T* val) {
  const bool fp = L.sell_fp_max > 0;
  return SellView<T>{L.sell_slots, L.sell_off.p, L.sell_perm.p, L.sell_len.p, L.sell_col.p, val, L.sell_window,
                     fp ? L.sell_lcol.p : nullptr, fp ? L.sell_fp_off.p : nullptr, fp ? L.sell_fp.p : nullptr,
                     fp ? L.sell_fp_max : 0};
}

// variant 4 on a level with at least sell_min_rows() active rows
static bool use_sell(const Level& L, int key) { return get_tuning(key) == 4 && L.n_active() >= sell_min_rows(); }

bool level_uses_sell(const Level& L) { return use_sell(L, 2) || use_sell(L, 3); }

void level_spmv(Level& L, const double* x, const double* z, double* y, double alpha, double beta,
                hipStream_t stream) {
  if(use_sell(L, 2)) {
    prepare_level_sell(L, false, stream);
    launch_spmv_sell<double>(L.n(), L.n_active(), sell_view<double>(L, L.sell_val.p), x, z, y, alpha, beta, stream);
  } else {
    launch_spmv(L.n(), L.n_active(), L.row_ptr.p, L.col.p, L.val.p, x, z, y, alpha, beta, stream, L.st.max_row);
  }
}

void level_spmv32(Level& L, const float* x, const float* z, float* y, float alpha, float beta, hipStream_t stream) {
  if(use_sell(L, 2)) {
    prepare_level_sell(L, true, stream);
    launch_spmv_sell<float>(L.n(), L.n_active(), sell_view<float>(L, L.sell_val32.p), x, z, y, alpha, beta, stream);
  } else {
    launch_spmv_f32(L.n(), L.n_active(), L.row_ptr.p, L.col.p, L.val32.p, x, z, y, alpha, beta, stream,
                    L.st.max_row);
  }
}

void level_jacobi(Level& L, const double* b, const double* x, double* x_out, double omega, hipStream_t stream) {
  if(use_sell(L, 3)) {
    prepare_level_sell(L, false, stream);
    launch_jacobi_sell<double>(L.n(), L.n_active(), sell_view<double>(L, L.sell_val.p), L.diag.p, b, x, x_out,
                               omega, stream);
  } else {
    launch_jacobi(L.n(), L.n_active(), L.row_ptr.p, L.col.p, L.val.p, nullptr, L.diag.p, b, x, x_out, omega, stream,
                  L.st.max_row);
  }
}

void level_jacobi32(Level& L, const float* b, const float* x, float* x_out, float omega, hipStream_t stream) {
  if(use_sell(L, 3)) {
    prepare_level_sell(L, true, stream);
    launch_jacobi_sell<float>(L.n(), L.n_active(), sell_view<float>(L, L.sell_val32.p), L.diag32.p, b, x, x_out,
                              omega, stream);
  } else {
    launch_jacobi_f32(L.n(), L.n_active(), L.row_ptr.p, L.col.p, L.val32.p, L.diag32.p, b, x, x_out, omega, stream,
                      L.st.max_row);
  }
}

void prepare_level_sweep(Level& L, hipStream_t stream) {
  if(!L.sweep_ptr.empty()) return;
  // the level schedule is exact; beyond sor_exact_limit() classes (deep
  // dependency chains, e.g. 3D meshes) launches per class would dominate and
  // the sweep uses greedy multicolour order
  SweepOrder o = build_sweep_order(L.st, sor_exact_limit());
  L.sweep_ptr = std::move(o.class_ptr);
  L.sweep_exact = o.exact;
  L.sweep_rows.upload(o.rows, stream);
  L.sweep_ptr_dev.upload(L.sweep_ptr, stream);
  L.sweep_max_class = 0;
  for(size_t k = 0; k + 1 < L.sweep_ptr.size(); ++k)
    L.sweep_max_class = std::max(L.sweep_max_class, L.sweep_ptr[k + 1] - L.sweep_ptr[k]);
  if(!o.exact) {
    const Structure& st = L.st;
    std::vector<int> rp(st.n_rows + 1, 0), cl(st.nnz);
    for(int i = 0; i < st.n_rows; ++i) {
      const int r = o.rows[i], len = st.row_ptr[r + 1] - st.row_ptr[r];
      rp[i + 1] = rp[i] + len;
      std::copy(st.col.begin() + st.row_ptr[r], st.col.begin() + st.row_ptr[r + 1], cl.begin() + rp[i]);
    }
    L.sw_rp.upload(rp, stream);
    L.sw_col.upload(cl, stream);
    L.sw_val.alloc(st.nnz);
    L.sw_fresh = false;
  }
  PMCD_HIP(hipStreamSynchronize(stream));
}

// class-ordered values up to date with val (and val32)
void refresh_sweep_values(Level& L, bool f32, hipStream_t stream) {
  if(L.sweep_exact || L.sw_rp.n == 0) return;
  if(!f32 && !L.sw_fresh) {
    launch_permute_rows(L.n(), L.sweep_rows.p, L.row_ptr.p, L.sw_rp.p, L.val.p, L.sw_val.p, stream);
    L.sw_fresh = true;
  }
  if(f32 && !L.sw_fresh32) {
    if(L.sw_val32.n != L.val.n) L.sw_val32.alloc(L.val.n);
    launch_permute_rows_f32(L.n(), L.sweep_rows.p, L.row_ptr.p, L.sw_rp.p, L.val32.p, L.sw_val32.p, stream);
    L.sw_fresh32 = true;
  }
}

void level_sor_sweep(Level& L, const double* b, double* x, double omega, int flag, hipStream_t stream,
                     const int* row_owner, const uint8_t* owner_active) {
  PMCD_CHECK(flag >= 0 && flag <= 2, "sor_sweep flag must be 0, 1 or 2");
  prepare_level_sweep(L, stream);
  refresh_sweep_values(L, false, stream);
  const int K = (int)L.sweep_ptr.size() - 1;
  const double mean_row = (double)L.st.nnz / std::max(1, L.n());
  if(L.sweep_exact && L.sweep_max_class <= sweep_block_rows()) {
    launch_sor_sweep_block(K, L.sweep_ptr_dev.p, L.sweep_rows.p, L.row_ptr.p, L.col.p, L.val.p, L.diag.p, b, x,
                           omega, flag, stream, row_owner, owner_active);
    return;
  }
  // multicolour classes: one workgroup per sweep when every class fits one
  // pass of 1024 threads (W lanes per row)
  if(!L.sweep_exact && (long)L.sweep_max_class * sweep_lanes(mean_row) <= sweep_block_rows()) {
    launch_sor_sweep_block_perm(K, L.sweep_ptr_dev.p, L.sweep_rows.p, L.sw_rp.p, L.sw_col.p, L.sw_val.p,
                                L.diag.p, b, x, omega, flag, L.st.max_row, mean_row, stream, row_owner,
                                owner_active);
    return;
  }
  auto cls = [&](int k) {
    const int cnt = L.sweep_ptr[k + 1] - L.sweep_ptr[k];
    if(L.sweep_exact)
      launch_sor_rows(cnt, L.sweep_rows.p + L.sweep_ptr[k], L.row_ptr.p, L.col.p, L.val.p, L.diag.p, b, x,
                      omega, true, L.st.max_row, mean_row, stream, row_owner, owner_active);
    else
      launch_sor_rows_perm(cnt, L.sweep_ptr[k], L.sweep_rows.p, L.sw_rp.p, L.sw_col.p, L.sw_val.p, L.diag.p,
                           b, x, omega, L.st.max_row, mean_row, stream, row_owner, owner_active);
  };
  if(flag == 0 || flag == 2)
    for(int k = 0; k < K; ++k) cls(k);
  if(flag == 1 || flag == 2)
    for(int k = K - 1; k >= 0; --k) cls(k);
}

void level_sor_sweep32(Level& L, const float* b, float* x, float omega, int flag,
                       hipStream_t stream, const int* row_owner, const uint8_t* owner_active) {
  PMCD_CHECK(flag >= 0 && flag <= 2, "sor_sweep flag must be 0, 1 or 2");
  PMCD_CHECK(L.val32.n == L.val.n, "FP32 matrix copies not prepared");
  prepare_level_sweep(L, stream);
  refresh_sweep_values(L, true, stream);
  const int K = (int)L.sweep_ptr.size() - 1;
  const double mean_row = (double)L.st.nnz / std::max(1, L.n());
  if(L.sweep_exact && L.sweep_max_class <= sweep_block_rows()) {
    launch_sor_sweep_block_f32(K, L.sweep_ptr_dev.p, L.sweep_rows.p, L.row_ptr.p, L.col.p, L.val32.p,
                               L.diag32.p, b, x, omega, flag, stream, row_owner, owner_active);
    return;
  }
  if(!L.sweep_exact && (long)L.sweep_max_class * sweep_lanes(mean_row) <= sweep_block_rows()) {
    launch_sor_sweep_block_perm_f32(K, L.sweep_ptr_dev.p, L.sweep_rows.p, L.sw_rp.p, L.sw_col.p, L.sw_val32.p,
                                    L.diag32.p, b, x, omega, flag, L.st.max_row, mean_row, stream, row_owner,
                                    owner_active);
    return;
  }
  auto cls = [&](int k) {
    const int cnt = L.sweep_ptr[k + 1] - L.sweep_ptr[k];
    if(L.sweep_exact)
      launch_sor_rows_f32(cnt, L.sweep_rows.p + L.sweep_ptr[k], L.row_ptr.p, L.col.p, L.val32.p, L.diag32.p, b,
                          x, omega, true, L.st.max_row, mean_row, stream, row_owner, owner_active);
    else
      launch_sor_rows_perm_f32(cnt, L.sweep_ptr[k], L.sweep_rows.p, L.sw_rp.p, L.sw_col.p, L.sw_val32.p,
                               L.diag32.p, b, x, omega, L.st.max_row, mean_row, stream, row_owner,
                               owner_active);
  };
  if(flag == 0 || flag == 2)
    for(int k = 0; k < K; ++k) cls(k);
  if(flag == 1 || flag == 2)
    for(int k = K - 1; k >= 0; --k) cls(k);
}

}  // namespace pmcd
