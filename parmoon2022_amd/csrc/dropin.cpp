#include "dropin.hpp"

namespace pmcd {

CallerAssembler::CallerAssembler(Shape shape, int order, int n_cells, int n_dof, int n_active,
                                 const double* vertices, const int* global_numbers, const int* begin_index,
                                 const int* row_ptr, const int* col, const ProblemDesc& forms, hipStream_t stream,
                                 int n_hanging, const int* hanging, const double* weights)
    : forms_(forms), stream_(stream), shape_(shape), n_cells_(n_cells) {
  PMCD_CHECK(order == 1 || order == 2, "ansatz order must be 1 or 2");
  PMCD_CHECK(n_cells > 0 && n_dof > 0 && n_active >= 0 && n_active <= n_dof, "invalid space sizes");
  PMCD_CHECK(vertices && global_numbers && begin_index, "null space arrays");
  const Element el = element_for(shape, order);
  const int nl = element_n_dof(el);
  const int nv = shape_n_vertices(shape);
  FESpace& sp = L_.space;
  sp.element = el;
  sp.n_local = nl;
  sp.n_dof = n_dof;
  sp.n_active = n_active;
  sp.n_dirichlet = n_dof - n_active;
  sp.dof.resize((size_t)n_cells * nl);
  for(int c = 0; c < n_cells; ++c) {
    PMCD_CHECK(begin_index[c + 1] - begin_index[c] == nl, "cell ", c, " has ", begin_index[c + 1] - begin_index[c],
               " local DOFs, the element ", nl);
    for(int m = 0; m < nl; ++m) {
      const int g = global_numbers[begin_index[c] + m];
      PMCD_CHECK(g >= 0 && g < n_dof, "cell ", c, ": global DOF ", g, " out of range");
      sp.dof[(size_t)c * nl + m] = g;
    }
  }
  if(n_hanging > 0) {
    // THangingNode::GetDOF / GetCoeff of the space's hanging DOFs
    constexpr int MC = FESpace::kMaxCoupling;
    PMCD_CHECK(hanging && weights, "hanging DOFs without their couplings");
    PMCD_CHECK(n_hanging <= n_active, "more hanging DOFs than active ones");
    sp.n_hanging = n_hanging;
    const int first = n_active - n_hanging;
    sp.hang_n.assign(n_hanging, 0);
    sp.hang_cpl.assign((size_t)MC * n_hanging, -1);
    sp.hang_w.assign((size_t)MC * n_hanging, 0.0);
    std::vector<char> seen(n_hanging, 0);
    for(int i = 0; i < n_hanging; ++i) {
      const int h = hanging[4 * i] - first;
      PMCD_CHECK(h >= 0 && h < n_hanging && !seen[h], "hanging DOF ", hanging[4 * i], " not in ", first, "..",
                 n_active - 1, " or listed twice");
      seen[h] = 1;
      int k = 0;
      for(; k < MC && hanging[4 * i + 1 + k] >= 0; ++k) {
        const int c = hanging[4 * i + 1 + k];
        PMCD_CHECK(c < n_dof && c != hanging[4 * i], "hanging DOF ", hanging[4 * i], ": coupling DOF ", c,
                   " out of range");
        sp.hang_cpl[MC * h + k] = c;
        sp.hang_w[MC * h + k] = weights[3 * i + k];
      }
      PMCD_CHECK(k > 0, "hanging DOF ", hanging[4 * i], " without coupling DOFs");
      sp.hang_n[h] = k;
    }
    sort_hanging(sp);
  }
  L_.st = structure_from_pattern(sp, n_cells, row_ptr, col);
  const Structure& st = L_.st;
  L_.cell_xyz.upload(vertices, (size_t)n_cells * nv * 3, stream_);
  L_.row_ptr.upload(st.row_ptr, stream_);
  L_.col.upload(st.col, stream_);
  L_.diag_pos.upload(st.diag, stream_);
  L_.contrib_ptr.upload(st.contrib_ptr, stream_);
  L_.contrib.upload(st.contrib, stream_);
  L_.slot.upload(st.slot, stream_);
  L_.dir_vals.alloc(std::max(1, sp.n_dirichlet));
  L_.val.alloc(st.nnz);
  L_.rhs.alloc(n_dof);
  rule_ = upload_reference_tables(shape, order, forms.linear_quad != 0, example_constant_matrix_coeffs(forms.ex.kind),
                                  ncomp_, ref_, qw_, qp_, tensor_, stream_);
  local_A_.alloc((size_t)n_cells * nl * nl);
  local_rhs_.alloc((size_t)n_cells * nl);
  if(sp.n_hanging) hanging_plan(L_, stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
}

void CallerAssembler::assemble(const double* dirichlet_values, bool values_on_device) {
  const FESpace& sp = L_.space;
  if(sp.n_dirichlet > 0) {
    if(!dirichlet_values)
      launch_fill(sp.n_dirichlet, 0.0, L_.dir_vals.p, stream_);
    else
      PMCD_HIP(hipMemcpyAsync(L_.dir_vals.p, dirichlet_values, sizeof(double) * sp.n_dirichlet,
                              values_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
  }
  AsmArgs a;
  a.shape = (int)shape_;
  a.dim = shape_dim(shape_);
  a.nv = shape_n_vertices(shape_);
  a.nl = sp.n_local;
  a.nq = quad_formula(rule_).n();
  a.ncomp = ncomp_;
  a.supg = forms_.supg;
  a.n_cells = n_cells_;
  a.ex = forms_.ex;
  a.ex.dim = a.dim;
  a.ref = ref_.p;
  a.qw = qw_.p;
  a.qp = qp_.p;
  a.tensor = tensor_.n ? tensor_.p : nullptr;
  a.cell_xyz = L_.cell_xyz.p;
  a.local_A = local_A_.p;
  a.local_rhs = local_rhs_.p;
  assemble_cells(a, stream_);
  launch_gather_matrix(sp.n_dof, sp.n_active, a.nl, L_.row_ptr.p, L_.col.p, L_.contrib_ptr.p, L_.contrib.p,
                       L_.slot.p, local_A_.p, L_.val.p, stream_);
  launch_gather_rhs(sp.n_dof, sp.n_active, L_.contrib_ptr.p, L_.contrib.p, local_rhs_.p, nullptr, nullptr,
                    L_.dir_vals.p, L_.rhs.p, stream_);
  if(sp.n_hanging) apply_hanging(L_, stream_);
}

void CallerAssembler::add_into_host(double* entries, double* rhs) {
  const FESpace& sp = L_.space;
  const Structure& st = L_.st;
  std::vector<double> val(st.nnz), f(sp.n_dof);
  L_.val.download(val.data(), val.size(), stream_);
  L_.rhs.download(f.data(), f.size(), stream_);
  PMCD_HIP(hipStreamSynchronize(stream_));
  const int first = sp.n_active_non_hanging();
  if(entries) {
    const int64_t nh = st.row_ptr[first], na = st.row_ptr[sp.n_active];
    for(int64_t k = 0; k < nh; ++k) entries[k] += val[k];
    for(int64_t k = nh; k < na; ++k) entries[k] = val[k];  // hanging rows: the constraints
    for(int r = sp.n_active; r < sp.n_dof; ++r) entries[st.diag[r]] = 1.0;
  }
  if(rhs) {
    for(int r = 0; r < first; ++r) rhs[r] += f[r];
    for(int r = first; r < sp.n_active; ++r) rhs[r] = 0.0;
    for(int r = sp.n_active; r < sp.n_dof; ++r) rhs[r] = f[r];
  }
}

void CallerAssembler::add_into_device(double* entries, double* rhs, hipStream_t caller_stream) {
  const FESpace& sp = L_.space;
  // the assembly is complete before the caller's stream updates its arrays
  PMCD_HIP(hipStreamSynchronize(stream_));
  launch_add_assembled(sp.n_dof, sp.n_active, L_.row_ptr.p, L_.diag_pos.p, L_.val.p, L_.rhs.p, entries, rhs,
                       caller_stream);
  if(sp.n_hanging) {
    // hanging rows: the constraints replace what was added, the rhs is 0
    const int first = sp.n_active_non_hanging();
    const int64_t b = L_.st.row_ptr[first], e = L_.st.row_ptr[sp.n_active];
    if(entries && e > b)
      PMCD_HIP(hipMemcpyAsync(entries + b, L_.val.p + b, sizeof(double) * (e - b), hipMemcpyDeviceToDevice,
                              caller_stream));
    if(rhs) launch_fill(sp.n_hanging, 0.0, rhs + first, caller_stream);
  }
  PMCD_HIP(hipStreamSynchronize(caller_stream));
}

}  // namespace pmcd
