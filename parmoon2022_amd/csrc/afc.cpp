// Algebraic flux correction on a device system: ConvectionDiffusion_AFC<d>
// with the Kuzmin or BJK17 limiter and the fixed point rhs, fixed point
// matrix or Newton scheme, with or without Anderson acceleration
// (src/System/ConvectionDiffusion_AFC.C:132-686, 1032-1180;
// src/AssembleRoutines/AlgebraicFluxCorrection.C:82-1140, 2058-2981,
// 2992-3080; src/General/anderson.f90). The matrices, the limiter and the
// Newton matrix run on the device; the damping control (accept / reject a
// damped iterate by its residual norm) and the small least-squares problem
// of the Anderson step run on the host between launches.
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

// Householder reflector H of the ones row: (1 ... 1) H = (0 ... 0 r),
// r = -sqrt(k), H = I - (2 / vv) v v^T with v = (1 ... 1) - r e_k
struct OnesReflector {
  std::vector<double> v;
  double r, vv;
  explicit OnesReflector(int k) : v(k, 1.0), r(-std::sqrt((double)k)) {
    v[k - 1] -= r;
    vv = 0;
    for(double x : v) vv += x * x;
  }
};

}  // namespace

// The Anderson step's coefficients (anderson.f90: DGGLSE with A = the last k
// updates R, B = (1 ... 1), c = 0, d = 1): alpha = argmin |R alpha| subject
// to sum alpha = 1. Solved as DGGLSE does: M = R H with the reflector of the
// ones row, a Householder QR of M on the device (norms and dot products as
// deterministic reductions), x_k = 1 / r, the first k - 1 unknowns from the
// triangular factor on the host, alpha = H x. Returns false (iterate left
// unchanged) on a zero pivot.
bool CDSystem::afc_anderson_alpha(int n, int k, const double* const* R, double* alpha) {
  const OnesReflector h(k);
  double* M = afc_M_.p;
  double* w = afc_w_.p;
  // w = R v, then M_j = R_j - (2 / vv) v_j w
  LinComb lc;
  lc.k = k;
  for(int j = 0; j < k; ++j) {
    lc.v[j] = R[j];
    lc.c[j] = h.v[j];
  }
  launch_lincomb(n, lc, w, stream_);
  for(int j = 0; j < k; ++j) {
    launch_copy(n, R[j], M + (size_t)j * n, stream_);
    launch_axpy(n, -(2.0 / h.vv) * h.v[j], w, M + (size_t)j * n, stream_);
  }
  bool ok = true;
  for(int c = 0; c < k && c < n; ++c) {
    double* x = M + (size_t)c * n;
    const int m = n - c;
    const double nrm = std::sqrt(dot(m, x + c, x + c));
    if(nrm == 0) {
      ok = c >= k - 1;
      continue;
    }
    double xc;
    PMCD_HIP(hipMemcpyAsync(&xc, x + c, sizeof(double), hipMemcpyDeviceToHost, stream_));
    PMCD_HIP(hipStreamSynchronize(stream_));
    const double a = xc > 0 ? -nrm : nrm;
    launch_copy(m, x + c, w + c, stream_);
    const double wc = xc - a;
    PMCD_HIP(hipMemcpyAsync(w + c, &wc, sizeof(double), hipMemcpyHostToDevice, stream_));
    const double ww = dot(m, w + c, w + c);
    for(int cc = c; cc < k; ++cc) {
      double* y = M + (size_t)cc * n;
      const double d = dot(m, w + c, y + c);
      launch_axpy(m, -(2.0 * d / ww), w + c, y + c, stream_);
    }
  }
  // upper triangle of the factor: rows 0..k-1 of every column
  std::vector<double> T((size_t)k * k);
  for(int j = 0; j < k; ++j)
    PMCD_HIP(hipMemcpyAsync(&T[(size_t)j * k], M + (size_t)j * n, sizeof(double) * std::min(k, n),
                            hipMemcpyDeviceToHost, stream_));
  PMCD_HIP(hipStreamSynchronize(stream_));
  if(!ok) return false;
  std::vector<double> xv(k);
  xv[k - 1] = 1.0 / h.r;
  for(int i = k - 2; i >= 0; --i) {
    double s = -T[(size_t)(k - 1) * k + i] * xv[k - 1];
    for(int j = i + 1; j < k - 1; ++j) s -= T[(size_t)j * k + i] * xv[j];
    const double piv = T[(size_t)i * k + i];
    if(piv == 0) return false;
    xv[i] = s / piv;
  }
  double vx = 0;
  for(int j = 0; j < k; ++j) vx += h.v[j] * xv[j];
  for(int j = 0; j < k; ++j) alpha[j] = xv[j] - (2.0 / h.vv) * vx * h.v[j];
  return true;
}

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
  PMCD_CHECK(prm[12] == 0.0 || prm[12] == 1.0 || prm[12] == 2.0, "unknown AFC iteration scheme ", prm[12]);
  const bool fpm = prm[12] == 1.0;     // fixed point matrix: S changes with the iterate
  const bool newton = prm[12] == 2.0;  // Newton: S = the damped Jacobian of the residual
  const double tau0 = prm[13];
  // Newton: the derivative's weight is afc_fixed_point_derivative_weight
  // times a factor that starts at 0 and follows the residual
  // (ConvectionDiffusion_AFC.C:427-460)
  const double deriv_weight = prm[14], threshold = prm[15];
  double deriv_factor = 0.0;
  const bool anderson = prm[16] == 1.0;
  const int kA = (int)prm[17], startA = (int)prm[18];
  PMCD_CHECK(!anderson || (kA >= 1 && kA <= LinComb::kMax), "Anderson acceleration needs 1..", LinComb::kMax,
             " vectors, got ", kA);
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
  if(newton && afc_Pp_.n != (size_t)n) {
    afc_Pp_.alloc(n); afc_Pm_.alloc(n); afc_Qp_.alloc(n); afc_Qm_.alloc(n);
    afc_umax_.alloc(n); afc_umin_.alloc(n); afc_q_.alloc(n);
  }
  if(anderson && afc_hX_.n != (size_t)n * kA) {
    afc_hX_.alloc((size_t)n * kA); afc_hR_.alloc((size_t)n * kA); afc_M_.alloc((size_t)n * kA);
    afc_w_.alloc(n);
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

  AfcLimiterOut lo;
  if(newton) {
    lo.Pp = afc_Pp_.p; lo.Pm = afc_Pm_.p; lo.Qp = afc_Qp_.p; lo.Qm = afc_Qm_.p;
    lo.umax = afc_umax_.p; lo.umin = afc_umin_.p; lo.q = afc_q_.p;
  }
  auto limited_rhs = [&]() {
    if(bjk)
      launch_afc_bjk_rhs(n, na, L.row_ptr.p, L.col.p, afc_D_.p, afc_gamma_.p, L.sol.p, afc_f_.p, afc_Rp_.p,
                         afc_Rm_.p, afc_alpha_.p, L.rhs.p, stream_, lo);
    else
      launch_afc_rhs(n, na, L.row_ptr.p, L.col.p, afc_tpos_.p, afc_D_.p, afc_E_.p, L.sol.p, afc_f_.p,
                     afc_Rp_.p, afc_Rm_.p, afc_alpha_.p, L.rhs.p, stream_, lo);
    if(newton) {
      AfcNewtonArgs na_args{n, na, L.row_ptr.p, L.col.p, afc_tpos_.p, afc_D_.p, afc_E_.p, afc_alpha_.p,
                            L.sol.p, afc_f_.p, afc_Pp_.p, afc_Pm_.p, afc_Qp_.p, afc_Qm_.p, afc_Rp_.p,
                            afc_Rm_.p, afc_umax_.p, afc_umin_.p, afc_q_.p, tau0, deriv_weight * deriv_factor,
                            bjk ? 1 : 0, L.val.p, L.rhs.p};
      launch_afc_newton(na_args, stream_);
      launch_gather_idx(n, L.diag_pos.p, L.val.p, L.diag.p, stream_);
      L.sw_fresh = L.sw_fresh32 = L.sell_fresh = L.sell_fresh32 = false;
      coarse_ready_ = false;
      mixed_ready_ = false;
    }
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
  int n_hist = 0, hist_head = 0;  // ring of the last kA outputs and updates, oldest at hist_head
  std::vector<double> alphaA(kA > 0 ? kA : 1);
  for(;; ++k) {
    const bool acc = anderson && k >= startA;
    if(acc) {
      // anderson_acceleration_damping: append the output and the update,
      // drop the oldest beyond kA, mix once kA are there
      int slot;
      if(n_hist == kA) {
        slot = hist_head;
        hist_head = (hist_head + 1) % kA;
      } else {
        slot = (hist_head + n_hist) % kA;
        ++n_hist;
      }
      double* X = afc_hX_.p + (size_t)slot * n;
      double* R = afc_hR_.p + (size_t)slot * n;
      launch_copy(n, L.sol.p, X, stream_);
      launch_copy(n, L.sol.p, R, stream_);
      launch_axpy(n, -1.0, afc_old_.p, R, stream_);
      if(k > kA + startA && n_hist == kA) {
        std::vector<const double*> cols(kA);
        for(int j = 0; j < kA; ++j) cols[j] = afc_hR_.p + (size_t)((hist_head + j) % kA) * n;
        if(afc_anderson_alpha(n, kA, cols.data(), alphaA.data())) {
          LinComb lc;
          lc.k = kA;
          for(int j = 0; j < kA; ++j) {
            lc.v[j] = afc_hX_.p + (size_t)((hist_head + j) % kA) * n;
            lc.c[j] = alphaA[j];
          }
          launch_lincomb(n, lc, L.sol.p, stream_);
        }
      }
    }
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
      if(!acc) {
        ++rejected;
        launch_copy(n, afc_cur_sol_.p, L.sol.p, stream_);
        launch_copy(n, afc_cur_rhs_.p, L.rhs.p, stream_);
      }
      omega = std::max(damp_min, omega * dec);
      if(first_damp) {
        damp_max = std::max(damp_min, max_dec * damp_max);
        first_damp = false;
      }
      damp = omega;
      // with Anderson acceleration a rejected step only lowers the damping
      if(acc) break;
    }
    if(residual < eps && k > 1) break;
    if(k == maxit) break;
    launch_copy(n, L.sol.p, afc_old_.p, stream_);
    st = solve(s);
    total.iterations += st.iterations;
    if(newton) {
      if(residual <= threshold) {
        double w = deriv_factor;
        if(w < 1e-3) w = 0.25;
        if(residual / residual_old > 0.99) {
          w *= 0.999;
          if(w < 0.1) w = 0.1;
        }
        if(residual / residual_old < 0.99) {
          w *= 1.001;
          if(w > 1.0) w = 1.0;
        }
        deriv_factor = w;
      }
      if(residual > 100 * threshold) deriv_factor = 0.0;
    }
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
