// Algebraic flux correction on a device system: ConvectionDiffusion_AFC<d>
// with the Kuzmin limiter and the fixed point rhs scheme
// (src/System/ConvectionDiffusion_AFC.C:132-600, 1032-1180;
// src/AssembleRoutines/AlgebraicFluxCorrection.C:82-165, 2058-2745,
// 2992-3080). The matrices and the limiter run on the device; the damping
// control (accept / reject a damped iterate by its residual norm) runs on
// the host between launches, as in the reference's dynamic_damping.
#include <cmath>

#include "system.hpp"

namespace pmcd {

namespace {

// Compute_Parameter_For_Linearity_Preservation (AlgebraicFluxCorrection.C:
// 1295-1470) for P1 triangles and tetrahedra: per vertex DOF the largest
// distance to another vertex of a cell over the smallest height of the
// vertex over the opposite side (in 2D the distance to a neighbour whose
// angle is not acute). Mesh geometry only: computed once on the host.
std::vector<double> bjk_gamma(const Mesh& m, const FESpace& sp) {
  const int n = sp.n_dof, nv = m.nv;
  const bool d2 = m.dim() == 2;
  std::vector<double> hi(n, 0.0), lo(n, 4711.0);
  for(int c = 0; c < m.n_cells; ++c) {
    double x[4], y[4], z[4];
    for(int a = 0; a < nv; ++a) {
      const double* X = &m.coords[3 * (size_t)m.cell(c)[a]];
      x[a] = X[0]; y[a] = X[1]; z[a] = X[2];
    }
    double measure;
    if(d2) {
      const double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
      measure = std::fabs(det) / 2.0;
    } else {
      const double a1 = x[1] - x[0], a2 = x[2] - x[0], a3 = x[3] - x[0];
      const double b1 = y[1] - y[0], b2 = y[2] - y[0], b3 = y[3] - y[0];
      const double c1 = z[1] - z[0], c2 = z[2] - z[0], c3 = z[3] - z[0];
      const double det = a1 * (b2 * c3 - b3 * c2) - a2 * (b1 * c3 - b3 * c1) + a3 * (b1 * c2 - b2 * c1);
      measure = std::fabs(det) / 6.0;
    }
    auto len = [&](int p, int q) {
      return std::sqrt((x[p] - x[q]) * (x[p] - x[q]) + (y[p] - y[q]) * (y[p] - y[q]) +
                       (z[p] - z[q]) * (z[p] - z[q]));
    };
    for(int j = 0; j < nv; ++j) {
      const int dof = sp.cell_dof(c)[j];
      const int j1 = (j + 1) % nv, j2 = (j + 2) % nv, j3 = (j + 3) % nv;
      double dist, dist_max;
      if(d2) {
        if((x[j2] - x[j1]) * (x[j] - x[j1]) + (y[j2] - y[j1]) * (y[j] - y[j1]) <= 0) {
          dist = len(j, j1);
          dist_max = len(j, j2);
        } else if((x[j1] - x[j2]) * (x[j] - x[j2]) + (y[j1] - y[j2]) * (y[j] - y[j2]) <= 0) {
          dist = len(j, j2);
          dist_max = len(j, j1);
        } else {
          dist = 2. * measure / len(j1, j2);
          dist_max = std::max(len(j, j1), len(j, j2));
        }
      } else {
        dist_max = std::max(std::max(len(j, j1), len(j, j2)), len(j, j3));
        const double A = (z[j3] - z[j1]) * (y[j2] - y[j1]) - (z[j2] - z[j1]) * (y[j3] - y[j1]);
        const double B = (z[j2] - z[j1]) * (x[j3] - x[j1]) - (z[j3] - z[j1]) * (x[j2] - x[j1]);
        const double C = (y[j3] - y[j1]) * (x[j2] - x[j1]) - (x[j3] - x[j1]) * (y[j2] - y[j1]);
        dist = (6.0 * measure) / std::sqrt(A * A + B * B + C * C);
      }
      hi[dof] = std::max(hi[dof], dist_max);
      lo[dof] = std::min(lo[dof], dist);
    }
  }
  std::vector<double> g(n);
  for(int i = 0; i < n; ++i) g[i] = hi[i] / lo[i];
  return g;
}

}  // namespace

SolveStats CDSystem::afc_solve(const SolverDesc& s, const double* prm, double* out) {
  PMCD_CHECK(!operator_only_, "an operator-only system has no mesh to assemble on");
  PMCD_CHECK(n_levels() == 1, "algebraic flux correction needs a single level (the reference refuses multigrid)");
  PMCD_CHECK(!desc_.supg, "algebraic flux correction uses the Galerkin discretization");
  PMCD_CHECK(!tcd_, "algebraic flux correction is for the stationary problem");
  Level& L = finest();
  const int n = L.n(), na = L.n_active();
  const size_t nnz = L.val.n;
  PMCD_CHECK(prm[11] == 0.0 || prm[11] == 1.0, "unknown AFC limiter ", prm[11]);
  const bool bjk = prm[11] == 1.0;
  PMCD_CHECK(prm[12] == 0.0 || prm[12] == 1.0, "unknown AFC iteration scheme ", prm[12]);
  const bool fpm = prm[12] == 1.0;  // fixed point matrix: S changes with the iterate
  const double tau0 = prm[13];
  PMCD_CHECK(!bjk || ((L.mesh.shape == Shape::Triangle || L.mesh.shape == Shape::Tetrahedron) &&
                      L.space.n_local == L.mesh.nv),
             "the BJK17 limiter needs P1 triangles or tetrahedra");
  if(afc_A_.n != nnz) {
    afc_tpos_.alloc(nnz);
    afc_A_.alloc(nnz); afc_D_.alloc(nnz); afc_E_.alloc(nnz); afc_alpha_.alloc(nnz);
    afc_f_.alloc(n); afc_Rp_.alloc(n); afc_Rm_.alloc(n); afc_old_.alloc(n);
    afc_cur_sol_.alloc(n); afc_cur_rhs_.alloc(n); afc_res_.alloc(n);
    launch_afc_transpose(n, L.row_ptr.p, L.col.p, afc_tpos_.p, stream_);
  }
  if(kr_.n != (size_t)n) {
    kr_.alloc(n);
    kw_.alloc(n);
  }
  // Galerkin matrix with its Dirichlet rows assembled (assemble_dirichlet_rows
  // of Assemble2D/3D), rhs with the Dirichlet values
  boundary_data(L);
  AsmArgs a = asm_args(L);
  assemble_cells(a, stream_);
  launch_gather_matrix(n, n, a.nl, L.row_ptr.p, L.col.p, L.contrib_ptr.p, L.contrib.p, L.slot.p, local_A_.p,
                       afc_A_.p, stream_);
  launch_gather_rhs(n, na, L.contrib_ptr.p, L.contrib.p, local_rhs_.p, L.extra_ptr.p, L.extra.p, L.dir_vals.p,
                    afc_f_.p, stream_);
  if(bjk) {
    launch_afc_bjk_modify(na, L.row_ptr.p, L.col.p, afc_tpos_.p, afc_A_.p, stream_);
    if(afc_gamma_.n != (size_t)n) afc_gamma_.alloc(n);
    const std::vector<double> gamma = bjk_gamma(L.mesh, L.space);
    afc_gamma_.upload(gamma, stream_);
    PMCD_HIP(hipStreamSynchronize(stream_));
  }
  // D, E = A + D, and the solver matrix S = E with identity Dirichlet rows
  launch_afc_matrices(n, na, L.row_ptr.p, L.col.p, afc_tpos_.p, afc_A_.p, afc_D_.p, afc_E_.p, L.val.p, stream_);
  launch_gather_idx(n, L.diag_pos.p, L.val.p, L.diag.p, stream_);
  L.sw_fresh = L.sw_fresh32 = L.sell_fresh = L.sell_fresh32 = false;
  coarse_ready_ = false;
  mixed_ready_ = false;
  launch_fill((int)nnz, 1.0, afc_alpha_.p, stream_);
  // initial iterate: zero, Dirichlet values from the rhs
  launch_fill(na, 0.0, L.sol.p, stream_);
  launch_copy(n - na, afc_f_.p + na, L.sol.p + na, stream_);

  auto limited_rhs = [&]() {
    if(bjk)
      launch_afc_bjk_rhs(n, na, L.row_ptr.p, L.col.p, afc_D_.p, afc_gamma_.p, L.sol.p, afc_f_.p, afc_Rp_.p,
                         afc_Rm_.p, afc_alpha_.p, L.rhs.p, stream_);
    else
      launch_afc_rhs(n, na, L.row_ptr.p, L.col.p, afc_tpos_.p, afc_D_.p, afc_E_.p, L.sol.p, afc_f_.p,
                     afc_Rp_.p, afc_Rm_.p, afc_alpha_.p, L.rhs.p, stream_);
    if(fpm) {
      // the system matrix follows the iterate: rebuilt here, refactorized by
      // the next solve
      launch_afc_fpm(n, na, L.row_ptr.p, L.col.p, afc_D_.p, afc_E_.p, afc_alpha_.p, L.sol.p, afc_f_.p, tau0,
                     L.val.p, L.rhs.p, stream_);
      launch_gather_idx(n, L.diag_pos.p, L.val.p, L.diag.p, stream_);
      L.sw_fresh = L.sw_fresh32 = L.sell_fresh = L.sell_fresh32 = false;
      coarse_ready_ = false;
      mixed_ready_ = false;
    }
  };
  auto residual_norm = [&]() {
    launch_spmv(n, na, L.row_ptr.p, L.col.p, L.val.p, L.sol.p, L.rhs.p, afc_res_.p, -1.0, 1.0, stream_,
                L.st.max_row);
    return std::sqrt(dot(n, afc_res_.p, afc_res_.p));
  };

  const double eps = prm[0] * std::sqrt((double)n);
  const int maxit = (int)prm[1];
  double damp = prm[2], damp_max = prm[3];
  const double damp_min = prm[4], max_global = prm[5], inc = prm[6], dec = prm[7], max_inc = prm[8],
               max_dec = prm[9], min_tol = prm[10];
  SolveStats total;
  // iteration 0: fluxes of the initial iterate, residual, first solve
  limited_rhs();
  double residual_old = residual_norm(), residual = residual_old;
  out[3] = residual_old;
  launch_copy(n, L.sol.p, afc_old_.p, stream_);
  SolveStats st = solve(s);
  total.iterations += st.iterations;
  int k = 1, rejected = 0;
  for(;; ++k) {
    double omega = damp;
    bool first_damp = true;
    launch_copy(n, L.rhs.p, afc_cur_rhs_.p, stream_);
    launch_copy(n, L.sol.p, afc_cur_sol_.p, stream_);
    for(;;) {
      launch_afc_damp(n, na, omega, afc_old_.p, L.sol.p, stream_);
      limited_rhs();
      residual = residual_norm();
      if(k == 1) break;
      if(residual < residual_old || omega <= min_tol * damp_min) {
        if(residual < residual_old && first_damp) {
          damp_max = std::min(max_global, max_inc * damp_max);
          omega = std::min(damp_max, inc * omega);
        }
        damp = omega;
        break;
      }
      ++rejected;
      launch_copy(n, afc_cur_sol_.p, L.sol.p, stream_);
      launch_copy(n, afc_cur_rhs_.p, L.rhs.p, stream_);
      omega = std::max(damp_min, omega * dec);
      if(first_damp) {
        damp_max = std::max(damp_min, max_dec * damp_max);
        first_damp = false;
      }
      damp = omega;
    }
    if(residual < eps && k > 1) break;
    if(k == maxit) break;
    launch_copy(n, L.sol.p, afc_old_.p, stream_);
    st = solve(s);
    total.iterations += st.iterations;
    residual_old = residual;
  }
  PMCD_HIP(hipStreamSynchronize(stream_));
  out[0] = k;
  out[1] = residual;
  out[2] = rejected;
  total.converged = residual < eps ? 1 : 0;
  total.initial_residual = out[3];
  total.final_residual = residual;
  return total;
}

}  // namespace pmcd
