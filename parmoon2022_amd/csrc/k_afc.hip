// Algebraic flux correction kernels (Kuzmin limiter, fixed point rhs):
// AlgebraicFluxCorrection::steady_state_algorithm, non-MPI path
// (src/AssembleRoutines/AlgebraicFluxCorrection.C:2058-2745) and
// compute_artificial_diffusion_matrix (:82-165).
//
// The reference walks the rows in order and scatters: while row i handles
// its link to column c (only when E_ci <= E_ic, E = A + D), it adds to
// P_i, Q_i and Q_c, writes alpha for both (i,c) and (c,i), and adds to
// rhs_i and rhs_c. Here one thread per row gathers everything row i
// receives, in the order the sequential loop adds it: first from rows
// c < i, then its own link terms in entry order, then from rows c > i
// (columns are sorted ascending). The flux of (c,i) is -F_ic exactly (D is
// symmetric entry by entry), and an alpha entry written from both sides
// keeps the value of the larger row index, as in the sequential loop.
#include "kernels.hpp"
#include "device.hpp"

namespace pmcd {

namespace {

// position of (c, i) in row c; row_ptr[c+1] if absent (sorted columns)
__device__ inline int find_in_row(const int* __restrict__ rp, const int* __restrict__ col, int c, int i) {
  int lo = rp[c], hi = rp[c + 1];
  const int end = hi;
  while(lo < hi) {
    const int mid = (lo + hi) >> 1;
    if(col[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && col[lo] == i) ? lo : end;
}

__global__ void k_afc_transpose(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                int* __restrict__ tpos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  for(int j = rp[i]; j < rp[i + 1]; ++j) tpos[j] = find_in_row(rp, col, col[j], i);
}

// D_ic = min(-a_ic, 0, -a_ci) off the diagonal, D_ii = -(row sum);
// E = A + D; S = E with identity Dirichlet rows (rows >= n_active)
__global__ void k_afc_matrices(int n, int n_active, const int* __restrict__ rp, const int* __restrict__ col,
                               const int* __restrict__ tpos, const double* __restrict__ A,
                               double* __restrict__ D, double* __restrict__ E, double* __restrict__ S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double sum = 0.0;
  int diag = -1;
  for(int j = rp[i]; j < rp[i + 1]; ++j) {
    const int c = col[j];
    double d = 0.0;
    if(c != i) {
      const int t = tpos[j];
      const double kij = A[j], kji = t < rp[c + 1] ? A[t] : 0.0;
      d = -kij;  // std::min({-kij, 0.0, -kji}): the first smallest
      if(0.0 < d) d = 0.0;
      if(-kji < d) d = -kji;
    } else {
      diag = j;
    }
    D[j] = d;
    sum += d;
  }
  if(diag >= 0) D[diag] = -sum;
  for(int j = rp[i]; j < rp[i + 1]; ++j) {
    const double e = A[j] + D[j];
    E[j] = e;
    S[j] = i < n_active ? e : (col[j] == i ? 1.0 : 0.0);
  }
}

// upwind test of the link (i, c) seen from row i: row i handles it unless
// E_ci > E_ic
__device__ inline bool handles(const double* __restrict__ E, int j, int t) { return !(E[t] > E[j]); }

// limiter factors R+ and R- of row i
__global__ void k_afc_limiter(int n, int n_active, const int* __restrict__ rp, const int* __restrict__ col,
                              const int* __restrict__ tpos, const double* __restrict__ D,
                              const double* __restrict__ E, const double* __restrict__ u,
                              double* __restrict__ Rp, double* __restrict__ Rm, AfcLimiterOut lo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int b = rp[i], e = rp[i + 1];
  const double ui = u[i];
  double Pp = 0, Pm = 0, Qp = 0, Qm = 0;
  // from rows c < i: their link (c, i) with flux -F_ic
  for(int j = b; j < e && col[j] < i; ++j) {
    const int t = tpos[j];
    if(t >= rp[col[j] + 1] || !handles(E, t, j)) continue;
    const double Fc = D[t] * (ui - u[col[j]]);
    if(Fc > 0) Qp += Fc;
    if(Fc < 0) Qm += Fc;
  }
  // own links
  for(int j = b; j < e; ++j) {
    const int c = col[j];
    if(c == i) continue;
    const int t = tpos[j];
    if(t < rp[c + 1] && !handles(E, j, t)) continue;
    const double F = D[j] * (u[c] - ui);
    if(F > 0) { Pp += F; Qm -= F; }
    if(F < 0) { Pm += F; Qp -= F; }
  }
  // from rows c > i
  for(int j = b; j < e; ++j) {
    if(col[j] <= i) continue;
    const int t = tpos[j];
    if(t >= rp[col[j] + 1] || !handles(E, t, j)) continue;
    const double Fc = D[t] * (ui - u[col[j]]);
    if(Fc > 0) Qp += Fc;
    if(Fc < 0) Qm += Fc;
  }
  if(lo.Pp) {
    lo.Pp[i] = Pp; lo.Pm[i] = Pm; lo.Qp[i] = Qp; lo.Qm[i] = Qm;
  }
  double rp_ = 1.0, rm_ = 1.0;
  if(fabs(Pp) > 0) { rp_ = Qp / Pp; if(rp_ > 1) rp_ = 1; }
  if(fabs(Pm) > 0) { rm_ = Qm / Pm; if(rm_ > 1) rm_ = 1; }
  if(i >= n_active) rp_ = rm_ = 1.0;
  Rp[i] = rp_;
  Rm[i] = rm_;
}

// alpha of entry j = (i, c): the value of the last row that handles the link
__device__ inline double link_alpha(int i, int c, bool from_i, bool from_c, double F, const double* Rp,
                                    const double* Rm) {
  const bool use_c = from_c && (!from_i || c > i);
  double a = 1.0;
  if(use_c) {
    const double Fc = -F;
    if(Fc > 0) a = Rp[c];
    if(Fc < 0) a = Rm[c];
  } else {
    if(F > 0) a = Rp[i];
    if(F < 0) a = Rm[i];
  }
  return a;
}

// rhs_i = f_i + limited fluxes (Dirichlet rows: u_i); alpha per entry
__global__ void k_afc_rhs(int n, int n_active, const int* __restrict__ rp, const int* __restrict__ col,
                          const int* __restrict__ tpos, const double* __restrict__ D,
                          const double* __restrict__ E, const double* __restrict__ u,
                          const double* __restrict__ Rp, const double* __restrict__ Rm,
                          const double* __restrict__ f, double* __restrict__ alpha, double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int b = rp[i], e = rp[i + 1];
  const double ui = u[i];
  double r = f[i];
  // pass 1: rows c < i subtract alpha (-F_ic); pass 2: own links add
  // alpha F_ic; pass 3: rows c > i
  for(int pass = 0; pass < 3; ++pass)
    for(int j = b; j < e; ++j) {
      const int c = col[j];
      if(c == i) {
        if(pass == 0) alpha[j] = 1.0;
        continue;
      }
      const int t = tpos[j];
      const bool has_t = t < rp[c + 1];
      const bool from_i = !has_t || handles(E, j, t);
      const bool from_c = has_t && handles(E, t, j);
      const double F = D[j] * (u[c] - ui);
      const double a = link_alpha(i, c, from_i, from_c, F, Rp, Rm);
      if(pass == 0) {
        alpha[j] = (from_i || from_c) ? a : alpha[j];
        if(c < i && from_c) r -= a * (D[t] * (ui - u[c]));
      } else if(pass == 1) {
        if(from_i) r += a * F;
      } else if(c > i && from_c) {
        r -= a * (D[t] * (ui - u[c]));
      }
    }
  rhs[i] = i < n_active ? r : ui;
}

// u = old + omega (u - old) on active rows, u = old on Dirichlet rows
__global__ void k_afc_damp(int n, int n_active, double omega, const double* __restrict__ old,
                           double* __restrict__ u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double o = old[i];
  u[i] = i < n_active ? (u[i] - o) * omega + o : o;
}

inline int blocks_for(int n) { return (n + 255) / 256; }

}  // namespace

void launch_afc_transpose(int n, const int* rp, const int* col, int* tpos, hipStream_t s) {
  if(n) k_afc_transpose<<<blocks_for(n), 256, 0, s>>>(n, rp, col, tpos);
  PMCD_HIP(hipGetLastError());
}

void launch_afc_matrices(int n, int n_active, const int* rp, const int* col, const int* tpos, const double* A,
                         double* D, double* E, double* S, hipStream_t s) {
  if(n) k_afc_matrices<<<blocks_for(n), 256, 0, s>>>(n, n_active, rp, col, tpos, A, D, E, S);
  PMCD_HIP(hipGetLastError());
}

void launch_afc_rhs(int n, int n_active, const int* rp, const int* col, const int* tpos, const double* D,
                    const double* E, const double* u, const double* f, double* Rp, double* Rm, double* alpha,
                    double* rhs, hipStream_t s, AfcLimiterOut lo) {
  if(!n) return;
  k_afc_limiter<<<blocks_for(n), 256, 0, s>>>(n, n_active, rp, col, tpos, D, E, u, Rp, Rm, lo);
  k_afc_rhs<<<blocks_for(n), 256, 0, s>>>(n, n_active, rp, col, tpos, D, E, u, Rp, Rm, f, alpha, rhs);
  PMCD_HIP(hipGetLastError());
}

void launch_afc_damp(int n, int n_active, double omega, const double* old, double* u, hipStream_t s) {
  if(n) k_afc_damp<<<blocks_for(n), 256, 0, s>>>(n, n_active, omega, old, u);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd

namespace pmcd {

namespace {

// modify_system_matrix (AlgebraicFluxCorrection.C:35-60): a_di = 0 for a
// Dirichlet column d of an active row i with a_id < 0 (each (d, i) entry has
// one writer, the row i)
__global__ void k_afc_bjk_modify(int n_active, const int* __restrict__ rp, const int* __restrict__ col,
                                 const int* __restrict__ tpos, double* __restrict__ A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n_active) return;
  for(int j = rp[i]; j < rp[i + 1]; ++j) {
    const int c = col[j];
    if(c >= n_active && A[j] < 0.0 && tpos[j] < rp[c + 1]) A[tpos[j]] = 0.0;
  }
}

// BJK17 factors R+ and R- of row i (steady_state_algorithm :2200-2270)
__global__ void k_afc_bjk_limiter(int n, int n_active, const int* __restrict__ rp, const int* __restrict__ col,
                                  const double* __restrict__ D, const double* __restrict__ gamma,
                                  const double* __restrict__ u, double* __restrict__ Rp, double* __restrict__ Rm,
                                  AfcLimiterOut lo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int b = rp[i], e = rp[i + 1];
  const double ui = u[i];
  double umin = ui, umax = ui, q = 0, Pp = 0, Pm = 0;
  for(int j = b; j < e; ++j) {
    const int c = col[j];
    const double uc = u[c];
    if(uc < umin) umin = uc;
    if(uc > umax) umax = uc;
    if(c != i) q += D[j];
  }
  q *= gamma[i];
  for(int j = b; j < e; ++j) {
    const int c = col[j];
    if(c == i) continue;
    const double F = D[j] * (u[c] - ui);
    if(F > 0) Pp += F;
    if(F < 0) Pm += F;
  }
  const double Qp = q * (ui - umax), Qm = q * (ui - umin);
  if(lo.Pp) {
    lo.Pp[i] = Pp; lo.Pm[i] = Pm; lo.Qp[i] = Qp; lo.Qm[i] = Qm;
    lo.umax[i] = umax; lo.umin[i] = umin; lo.q[i] = q;
  }
  double rp_ = 1.0, rm_ = 1.0;
  if(fabs(Pp) > 0) { rp_ = Qp / Pp; if(rp_ > 1) rp_ = 1; }
  if(fabs(Pm) > 0) { rm_ = Qm / Pm; if(rm_ > 1) rm_ = 1; }
  if(i >= n_active) rp_ = rm_ = 1.0;
  Rp[i] = rp_;
  Rm[i] = rm_;
}

// rhs_i = f_i + sum_j alpha_ij F_ij, alpha_ij = min of the two rows' factors
// (:2527-2560, 2720-2722); Dirichlet rows: u_i
__global__ void k_afc_bjk_rhs(int n, int n_active, const int* __restrict__ rp, const int* __restrict__ col,
                              const double* __restrict__ D, const double* __restrict__ u,
                              const double* __restrict__ Rp, const double* __restrict__ Rm,
                              const double* __restrict__ f, double* __restrict__ alpha, double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double ui = u[i];
  double r = f[i];
  for(int j = rp[i]; j < rp[i + 1]; ++j) {
    const int c = col[j];
    const double F = D[j] * (u[c] - ui);
    double a = 1.0;
    if(F > 0) { a = Rp[i]; if(Rm[c] < a) a = Rm[c]; }
    if(F < 0) { a = Rm[i]; if(Rp[c] < a) a = Rp[c]; }
    if(c == i) a = 1.0;
    alpha[j] = a;
    r += a * F;
  }
  rhs[i] = i < n_active ? r : ui;
}

}  // namespace

void launch_afc_bjk_modify(int n_active, const int* rp, const int* col, const int* tpos, double* A,
                           hipStream_t s) {
  if(n_active) k_afc_bjk_modify<<<(n_active + 255) / 256, 256, 0, s>>>(n_active, rp, col, tpos, A);
  PMCD_HIP(hipGetLastError());
}

void launch_afc_bjk_rhs(int n, int n_active, const int* rp, const int* col, const double* D, const double* gamma,
                        const double* u, const double* f, double* Rp, double* Rm, double* alpha, double* rhs,
                        hipStream_t s, AfcLimiterOut lo) {
  if(!n) return;
  k_afc_bjk_limiter<<<(n + 255) / 256, 256, 0, s>>>(n, n_active, rp, col, D, gamma, u, Rp, Rm, lo);
  k_afc_bjk_rhs<<<(n + 255) / 256, 256, 0, s>>>(n, n_active, rp, col, D, u, Rp, Rm, f, alpha, rhs);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd

namespace pmcd {

namespace {

// fixed point matrix scheme (steady_state_algorithm :2746-2790): with the
// limiter's alpha, S = E - tau0 alpha D off the diagonal, S_ii = E_ii +
// tau0 sum_j alpha_ij d_ij; rhs = f + tau1 (alpha_ij d_ij u_j, in entry
// order, minus (sum_j alpha_ij d_ij) u_i at the diagonal); Dirichlet rows:
// identity and u_i
__global__ void k_afc_fpm(int n, int n_active, const int* __restrict__ rp, const int* __restrict__ col,
                          const double* __restrict__ D, const double* __restrict__ E,
                          const double* __restrict__ alpha, const double* __restrict__ u,
                          const double* __restrict__ f, double tau0, double* __restrict__ S,
                          double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int b = rp[i], e = rp[i + 1];
  if(i >= n_active) {
    for(int j = b; j < e; ++j) S[j] = col[j] == i ? 1.0 : 0.0;
    rhs[i] = u[i];
    return;
  }
  const double tau1 = 1.0 - tau0;
  double sum_flux = 0.0;
  for(int j = b; j < e; ++j)
    if(col[j] != i) sum_flux += alpha[j] * D[j];
  double r = f[i];
  for(int j = b; j < e; ++j) {
    const int c = col[j];
    double v = E[j];
    if(c != i) {
      v -= alpha[j] * D[j] * tau0;
      r += alpha[j] * D[j] * u[c] * tau1;
    } else {
      v += sum_flux * tau0;
      r -= sum_flux * u[i] * tau1;
    }
    S[j] = v;
  }
  rhs[i] = r;
}

}  // namespace

void launch_afc_fpm(int n, int n_active, const int* rp, const int* col, const double* D, const double* E,
                    const double* alpha, const double* u, const double* f, double tau0, double* S, double* rhs,
                    hipStream_t s) {
  if(n) k_afc_fpm<<<(n + 255) / 256, 256, 0, s>>>(n, n_active, rp, col, D, E, alpha, u, f, tau0, S, rhs);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd

namespace pmcd {

namespace {

// flux of entry m of row r: d_rc (u_c - u_r)
__device__ inline double flux(const AfcNewtonArgs& a, int r, int m) { return a.D[m] * (a.u[a.col[m]] - a.u[r]); }

// position of (r, c) or -1
__device__ inline int entry_of(const AfcNewtonArgs& a, int r, int c) {
  const int p = find_in_row(a.rp, a.col, r, c);
  return p < a.rp[r + 1] ? p : -1;
}

// sum of d over the entries of row r with F < 0 (sign -1) or F > 0 (+1);
// `upwind`: only entries whose E is at least the transposed entry's
__device__ double row_d_sum(const AfcNewtonArgs& a, int r, int sign, bool upwind) {
  double s = 0;
  for(int m = a.rp[r]; m < a.rp[r + 1]; ++m) {
    const double F = flux(a, r, m);
    if(!(sign > 0 ? F > 0 : F < 0)) continue;
    if(upwind && !(a.E[m] >= a.E[a.tpos[m]])) continue;
    s += a.D[m];
  }
  return s;
}

// Compute_Jacobian_times_flux_Kuzmin (AlgebraicFluxCorrection.C:224-578);
// the reference keeps E_kj in an int in the downwind branch, and so does
// this code
__device__ double jac_kuzmin(const AfcNewtonArgs& a, int i, int j, int pij) {
  const double eps = 1e-14, epsinv = 1.0 / eps;
  const double *E = a.E, *D = a.D;
  const int a_ji = entry_of(a, j, i);
  const double Fij = flux(a, i, pij);
  double sum = 0;
  for(int k = a.rp[i]; k < a.rp[i + 1]; ++k) {
    const int kk = a.col[k];
    const double Fk = flux(a, i, k);
    const int a_ki = a.tpos[k];
    if(E[k] >= E[a_ki]) {
      if(Fk > 0 && fabs(a.Rp[i] - 1) < eps) continue;
      if(Fk < 0 && fabs(a.Rm[i] - 1) < eps) continue;
      if(fabs(Fk) < eps) continue;
      if(i != j && fabs(D[pij]) < 1e-14) continue;
      double DQ, DP, val;
      if(Fk > 0) {
        DQ = i != j ? (Fij >= 0 ? 0 : -D[pij]) : row_d_sum(a, i, -1, false);
        DP = i != j ? (Fij <= 0 ? 0 : (E[pij] >= E[a_ji] ? D[pij] : 0)) : -row_d_sum(a, i, +1, true);
        val = DQ / a.Pp[i] - a.Qp[i] * DP / (a.Pp[i] * a.Pp[i]);
      } else {
        DQ = i != j ? (Fij <= 0 ? 0 : -D[pij]) : row_d_sum(a, i, +1, false);
        DP = i != j ? (Fij >= 0 ? 0 : (E[pij] >= E[a_ji] ? D[pij] : 0)) : -row_d_sum(a, i, -1, true);
        val = DQ / a.Pm[i] - a.Qm[i] * DP / (a.Pm[i] * a.Pm[i]);
      }
      if(fabs(val) < epsinv) sum += val * Fk;
    } else {
      if(Fk > 0 && fabs(a.Rm[kk] - 1) < eps) continue;
      if(Fk < 0 && fabs(a.Rp[kk] - 1) < eps) continue;
      if(fabs(Fk) < eps) continue;
      const int a_kj = entry_of(a, kk, j);
      if(a_kj < 0) continue;
      const double f_kj = flux(a, kk, a_kj), d_kj = D[a_kj];
      const int e_kj = (int)E[a_kj];
      if(kk != j && fabs(d_kj) < 1e-14) continue;
      const int a_jk = entry_of(a, j, kk);
      double DQ, DP, val;
      if(Fk > 0) {
        DQ = kk != j ? (f_kj <= 0 ? 0 : -d_kj) : row_d_sum(a, kk, +1, false);
        DP = kk != j ? (f_kj >= 0 ? 0 : (e_kj >= E[a_jk] ? d_kj : 0)) : -row_d_sum(a, kk, -1, true);
        val = DQ / a.Pm[kk] - a.Qm[kk] * DP / (a.Pm[kk] * a.Pm[kk]);
      } else {
        DQ = kk != j ? (f_kj >= 0 ? 0 : -d_kj) : row_d_sum(a, kk, -1, false);
        DP = kk != j ? (f_kj <= 0 ? 0 : (e_kj >= E[a_jk] ? d_kj : 0)) : -row_d_sum(a, kk, +1, true);
        val = DQ / a.Pp[kk] - a.Qp[kk] * DP / (a.Pp[kk] * a.Pp[kk]);
      }
      if(fabs(val) < epsinv) sum += val * Fk;
    }
  }
  return sum;
}

// Compute_Jacobian_times_flux_BJK17 (AlgebraicFluxCorrection.C:957-1140)
__device__ double jac_bjk(const AfcNewtonArgs& a, int i, int j, int pij) {
  const double eps = 1e-14, epsinv = 1.0 / eps;
  const double* D = a.D;
  const double u_j = a.u[j], Fij = flux(a, i, pij);
  double sum = 0;
  for(int k = a.rp[i]; k < a.rp[i + 1]; ++k) {
    const double Fk = flux(a, i, k);
    if(Fk > 0 && fabs(a.Rp[i] - 1.0) < eps) continue;
    if(Fk < 0 && fabs(a.Rm[i] - 1.0) < eps) continue;
    if(fabs(Fk) < eps) continue;
    const int ck = a.col[k];
    double DQ = 0.0, DP = 0.0;
    if(Fk > 0) {
      if(a.Rp[i] <= a.Rm[ck]) {
        if(i != j) {
          if(fabs(a.umax[i] - u_j) < eps) DQ = -a.q[i];
          if(Fij > 0) DP = D[pij];
        } else {
          if(fabs(a.umax[i] - u_j) >= eps) DQ = a.q[i];
          DP = -row_d_sum(a, i, +1, false);
        }
        if(a.Pp[i] > 0.0) {
          const double val = DQ / a.Pp[i] - a.Qp[i] * DP / (a.Pp[i] * a.Pp[i]);
          if(fabs(val) < epsinv) sum += val * Fk;
        }
      } else {
        if(ck != j) {
          if(fabs(a.umin[ck] - u_j) < eps) DQ = -a.q[ck];
          if(Fij < 0) DP = D[pij];
        } else {
          if(fabs(a.umin[j] - u_j) >= eps) DQ = a.q[j];
          DP = -row_d_sum(a, j, -1, false);
        }
        if(a.Pm[ck] < 0.0) {
          const double val = DQ / a.Pm[ck] - a.Qm[ck] * DP / (a.Pm[ck] * a.Pm[ck]);
          if(fabs(val) < epsinv) sum += val * Fk;
        }
      }
    } else {
      if(a.Rm[i] <= a.Rp[ck]) {
        if(i != j) {
          if(fabs(a.umin[i] - u_j) < eps) DQ = -a.q[i];
          if(Fij < 0) DP = D[pij];
        } else {
          if(fabs(a.umin[i] - u_j) >= eps) DQ = a.q[i];
          DP = -row_d_sum(a, i, -1, false);
        }
        if(a.Pm[i] < 0.0) {
          const double val = DQ / a.Pm[i] - a.Qm[i] * DP / (a.Pm[i] * a.Pm[i]);
          if(fabs(val) < epsinv) sum += val * Fk;
        }
      } else {
        if(ck != j) {
          if(fabs(a.umax[ck] - u_j) < eps) DQ = -a.q[ck];
          if(Fij > 0) DP = D[pij];
        } else {
          if(fabs(a.umax[j] - u_j) >= eps) DQ = a.q[j];
          DP = -row_d_sum(a, j, +1, false);
        }
        if(a.Pp[ck] > 0.0) {
          const double val = DQ / a.Pp[ck] - a.Qp[ck] * DP / (a.Pp[ck] * a.Pp[ck]);
          if(fabs(val) < epsinv) sum += val * Fk;
        }
      }
    }
  }
  return sum;
}

// Newton scheme (steady_state_algorithm :2794-2981), one thread per row:
// r_i = f_i + sum_j (-E_ij u_j + alpha_ij F_ij) in entry order, then the
// row of DF (E - tau0 alpha D - tau1 J off the diagonal, E + tau0
// sum alpha D - tau1 J on it) and r_i += DF_ij u_j in entry order; the
// system matrix takes DF. Dirichlet rows: identity and u_i.
__global__ void k_afc_newton(AfcNewtonArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= a.n) return;
  const int b = a.rp[i], e = a.rp[i + 1];
  if(i >= a.n_active) {
    for(int j = b; j < e; ++j) a.S[j] = a.col[j] == i ? 1.0 : 0.0;
    a.rhs[i] = a.u[i];
    return;
  }
  double r = a.f[i];
  for(int j = b; j < e; ++j) r += -(a.E[j]) * a.u[a.col[j]] + a.alpha[j] * flux(a, i, j);
  double sum_flux = 0.0;
  for(int j = b; j < e; ++j)
    if(a.col[j] != i) sum_flux += a.alpha[j] * a.D[j];
  for(int j = b; j < e; ++j) {
    const int c = a.col[j];
    double J = 0;
    if(a.tau1 != 0) J = a.bjk ? jac_bjk(a, i, c, j) : jac_kuzmin(a, i, c, j);
    const double df = c != i ? a.E[j] - a.tau0 * a.alpha[j] * a.D[j] - a.tau1 * J
                             : a.E[j] + a.tau0 * sum_flux - a.tau1 * J;
    a.S[j] = df;
    r += df * a.u[c];
  }
  a.rhs[i] = r;
}

}  // namespace

void launch_afc_newton(const AfcNewtonArgs& a, hipStream_t s) {
  if(a.n) k_afc_newton<<<blocks_for(a.n), 256, 0, s>>>(a);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd
