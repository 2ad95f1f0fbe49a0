// FP64-MFMA assembly of the convection-diffusion local matrices and right-hand
// sides (same forms as k_assemble.hip: BilinearAssembleGalerkin<d> +
// BilinearAssemble_SD<d>, src/AssembleRoutines/ConvDiff.C:1776-1999).
//
// For an affine cell the physical derivatives are fixed linear combinations
// of the reference ones, so with R[q][s][i] the cell-independent reference
// table (value, first and second derivatives of basis function i at point q)
// the local matrix factors as
//     A_K[i][j] = sum_{q, r} R[q][r][i] * V_K[q][r][j],
//     V_K[q][r][j] = sum_s W_K[q][r][s] * R[q][s][j],
// with r over the test components (value, d/dxi_k) and s over the ansatz
// components (value, d/dxi_k, second derivatives). W_K (per point, 1+d by
// components) holds weight * |det J| times the coefficient and metric terms:
//     W[0][0] = c,  W[0][1+l] = beta_l,  W[1+k][0] = delta beta_k c,
//     W[1+k][1+l] = eps M_kl + delta beta_k beta_l,
//     W[1+k][H_t] = -delta beta_k eps m_t,
// where G = J^-T (rows: d/dx_a = sum_k G[a][k] d/dxi_k), M = G^T G,
// beta = G^T b, m_t the Laplacian weights of the second-derivative components,
// and delta the SUPG parameter (0 for Galerkin). The right-hand side is the
// extra column j = n_local: V[q][0][n] = w f, V[q][1+k][n] = w delta f beta_k.
//
// Stage 1 (vector ALUs) builds V in LDS for a chunk of points; stage 2 runs
// the (n_local x 4 n_q) x (4 n_q x (n_local+1)) product as 16x16x4 FP64 MFMA
// tiles, the K dimension being the (point, test component) pairs. One wave
// per cell, four cells per block sharing the reference-table chunk.
//
// Time-dependent problems (TCDStiff + TCDStiffSUPG, TCDMass + TCDMassSUPG,
// TCDRhs + TCDRhsSUPG, CD_local_assembling_routines.C:9-387, SDFEM type 2)
// have the same stiffness and rhs forms; the mass matrix u v + delta u (b.grad
// v) is the same product with WM[q][0] = w, WM[q][1+k] = w delta beta_k on the
// value component only. It is appended as columns 32.. of V, so one pass over
// the reference table yields stiffness, rhs and mass.
#include <vector>

#include "kernels.hpp"
#include "device.hpp"

namespace pmcd {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kCells = 4;
constexpr int kQC = 8;     // quadrature points per chunk
constexpr int kPad = 32;   // padded local dimension (two 16-wide tiles)

__device__ inline double sdfem_delta_m(int dim, double hK, double eps, double b1, double b2,
                                       double b3) {
  if(fabs(eps) < 1e-20) eps = 1e-20;
  double nb = b1 * b1 + b2 * b2;
  if(dim == 3) nb += b3 * b3;
  nb = sqrt(nb);
  if(!(nb > 0)) return 0.0;
  const double alpha = nb * hK / (2 * eps);
  return hK * (1 / tanh(alpha) - 1 / alpha) / (2 * nb);
}

// affine map data of a cell (same maps as the vector kernel)
struct GeoM {
  double c0[3], C[3][3], G[3][3], det, hK;
};

__device__ inline void geo_m(int shape, const double* v, GeoM& g) {
  auto X = [&](int i, int k) { return v[3 * i + k]; };
  double K[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for(int a = 0; a < 3; ++a)
    for(int k = 0; k < 3; ++k) g.C[a][k] = 0;
  double d2 = 0;
  if(shape == (int)Shape::Triangle || shape == (int)Shape::Quadrangle) {
    for(int a = 0; a < 2; ++a) {
      if(shape == (int)Shape::Triangle) {
        g.c0[a] = X(0, a); g.C[a][0] = X(1, a) - X(0, a); g.C[a][1] = X(2, a) - X(0, a);
      } else {
        g.c0[a] = (X(1, a) + X(3, a)) * 0.5;
        g.C[a][0] = (X(1, a) - X(0, a)) * 0.5;
        g.C[a][1] = (X(3, a) - X(0, a)) * 0.5;
      }
    }
    g.c0[2] = 0;
    g.det = g.C[0][0] * g.C[1][1] - g.C[0][1] * g.C[1][0];
    K[0][0] = g.C[1][1]; K[0][1] = -g.C[1][0]; K[1][0] = -g.C[0][1]; K[1][1] = g.C[0][0];
    const int nv = shape == (int)Shape::Triangle ? 3 : 4;
    if(nv == 3) {
      for(int m = 0; m < 3; ++m) {
        int p = m, q = (m + 1) % 3;
        double s = 0;
        for(int k = 0; k < 3; ++k) s += (X(p, k) - X(q, k)) * (X(p, k) - X(q, k));
        d2 = fmax(d2, s);
      }
    } else {
      for(int m = 0; m < 2; ++m) {
        double s = 0;
        for(int k = 0; k < 3; ++k) s += (X(m, k) - X(m + 2, k)) * (X(m, k) - X(m + 2, k));
        d2 = fmax(d2, s);
      }
    }
  } else {
    for(int a = 0; a < 3; ++a) {
      if(shape == (int)Shape::Tetrahedron) {
        g.c0[a] = X(0, a);
        for(int k = 0; k < 3; ++k) g.C[a][k] = X(k + 1, a) - X(0, a);
      } else {
        g.c0[a] = (X(1, a) + X(3, a) + X(4, a) - X(0, a)) * 0.5;
        g.C[a][0] = (X(1, a) - X(0, a)) * 0.5;
        g.C[a][1] = (X(3, a) - X(0, a)) * 0.5;
        g.C[a][2] = (X(4, a) - X(0, a)) * 0.5;
      }
    }
    const double xc1 = g.C[0][0], xc2 = g.C[0][1], xc3 = g.C[0][2];
    const double yc1 = g.C[1][0], yc2 = g.C[1][1], yc3 = g.C[1][2];
    const double zc1 = g.C[2][0], zc2 = g.C[2][1], zc3 = g.C[2][2];
    g.det = xc1 * yc2 * zc3 + xc2 * yc3 * zc1 + xc3 * yc1 * zc2 - xc3 * yc2 * zc1 - xc2 * yc1 * zc3 -
            xc1 * yc3 * zc2;
    K[0][0] = yc2 * zc3 - yc3 * zc2; K[0][1] = yc3 * zc1 - yc1 * zc3; K[0][2] = yc1 * zc2 - yc2 * zc1;
    K[1][0] = xc3 * zc2 - xc2 * zc3; K[1][1] = xc1 * zc3 - xc3 * zc1; K[1][2] = xc2 * zc1 - xc1 * zc2;
    K[2][0] = xc2 * yc3 - xc3 * yc2; K[2][1] = xc3 * yc1 - xc1 * yc3; K[2][2] = xc1 * yc2 - xc2 * yc1;
    if(shape == (int)Shape::Tetrahedron) {
      for(int p = 0; p < 4; ++p)
        for(int q = p + 1; q < 4; ++q) {
          double s = 0;
          for(int k = 0; k < 3; ++k) s += (X(p, k) - X(q, k)) * (X(p, k) - X(q, k));
          d2 = fmax(d2, s);
        }
    } else {
      const int dg[4][2] = {{0, 6}, {1, 7}, {2, 4}, {3, 5}};
      for(int m = 0; m < 4; ++m) {
        double s = 0;
        for(int k = 0; k < 3; ++k) s += (X(dg[m][0], k) - X(dg[m][1], k)) * (X(dg[m][0], k) - X(dg[m][1], k));
        d2 = fmax(d2, s);
      }
    }
  }
  for(int a = 0; a < 3; ++a)
    for(int k = 0; k < 3; ++k) g.G[a][k] = K[a][k] / g.det;
  g.hK = sqrt(d2);
}

// DIM: 2 or 3; NC: reference components (6 or 10); NT = 1 + DIM test
// components; TCD: also the mass matrix (columns kPad.. of V)
template <int DIM, int NL, bool TCD>
__global__ __launch_bounds__(256) void k_assemble_mfma(AsmArgs a) {
  constexpr int NC = DIM == 2 ? 6 : 10;
  constexpr int NT = 1 + DIM;
  constexpr int NH = NC - NT;               // second-derivative components
  constexpr int QC = TCD ? kQC / 2 : kQC;   // points per chunk (LDS: V is twice as wide)
  constexpr int PV = TCD ? 2 * kPad : kPad; // V width
  constexpr int NIT = (NL + 15) / 16;       // i tiles
  constexpr int NJT = TCD ? (kPad + NL + 15) / 16 : (NL + 1 + 15) / 16;  // j tiles (rhs, mass)
  constexpr int KC = QC * NT;               // K extent of one chunk
  static_assert(KC % 4 == 0, "chunk K extent must be a multiple of 4");
  static_assert(NL + 1 <= kPad, "local matrix and rhs column must fit the padded width");
  __shared__ double s_R[QC][NC][kPad];       // reference table chunk, zero padded
  __shared__ double s_V[kCells][KC][PV];     // per-cell V chunk
  __shared__ double s_W[kCells][QC][NT][NC + 1];  // last column: rhs coefficient
  __shared__ double s_WM[kCells][QC][TCD ? NT : 1];  // mass weights

  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int cell = blockIdx.x * kCells + wave;
  const bool live = cell < a.n_cells;
  GeoM g;
  if(live) geo_m(a.shape, a.cell_xyz + (size_t)cell * a.nv * 3, g);
  // metric terms: M = G^T G, m (Laplacian weights), beta needs b per point
  double M[3][3];
  for(int k = 0; k < 3; ++k)
    for(int l = 0; l < 3; ++l) {
      double s = 0;
      if(live)
        for(int d = 0; d < DIM; ++d) s += g.G[d][k] * g.G[d][l];
      M[k][l] = s;
    }
  double mlap[6];
  if(DIM == 2) { mlap[0] = M[0][0]; mlap[1] = 2 * M[0][1]; mlap[2] = M[1][1]; }
  else {
    mlap[0] = M[0][0]; mlap[1] = 2 * M[0][1]; mlap[2] = 2 * M[0][2];
    mlap[3] = M[1][1]; mlap[4] = 2 * M[1][2]; mlap[5] = M[2][2];
  }

  v4d acc[NIT][NJT];
  for(int it = 0; it < NIT; ++it)
    for(int jt = 0; jt < NJT; ++jt) acc[it][jt] = v4d{0, 0, 0, 0};

  for(int q0 = 0; q0 < a.nq; q0 += QC) {
    const int nqc = min(QC, a.nq - q0);
    // reference table chunk (points beyond nq and columns beyond NL are zero)
    for(int t = threadIdx.x; t < QC * NC * kPad; t += 256) {
      const int qq = t / (NC * kPad), rem = t % (NC * kPad), s = rem / kPad, i = rem % kPad;
      double v = 0;
      if(qq < nqc && i < NL) v = a.ref[((size_t)(q0 + qq) * NC + s) * NL + i];
      s_R[qq][s][i] = v;
    }
    // W per point (lanes over points)
    if(live && lane < QC) {
      const int qq = lane;
      double (*W)[NC + 1] = s_W[wave][qq];
      if(qq < nqc) {
        const int q = q0 + qq;
        const double xi = a.qp[3 * q], eta = a.qp[3 * q + 1], zeta = a.qp[3 * q + 2];
        double x[3];
        for(int d = 0; d < 3; ++d) {
          x[d] = g.c0[d] + g.C[d][0] * xi + g.C[d][1] * eta;
          if(DIM == 3) x[d] = x[d] + g.C[d][2] * zeta;
        }
        const double w = a.qw[q] * fabs(g.det);
        double cf[6];
        example_coeffs(cell_params(a, cell), x[0], x[1], DIM == 3 ? x[2] : 0.0, cf);
        const double eps = cf[0], c = cf[4], f = cf[5];
        const double b[3] = {cf[1], cf[2], DIM == 3 ? cf[3] : 0.0};
        const double delta = a.supg ? sdfem_delta_m(DIM, g.hK, eps, b[0], b[1], b[2]) : 0.0;
        double beta[3];
        for(int k = 0; k < 3; ++k) {
          double s = 0;
          for(int d = 0; d < DIM; ++d) s += g.G[d][k] * b[d];
          beta[k] = s;
        }
        W[0][0] = w * c;
        for(int l = 0; l < DIM; ++l) W[0][1 + l] = w * beta[l];
        for(int t = 0; t < NH; ++t) W[0][NT + t] = 0.0;
        W[0][NC] = w * f;
        for(int k = 0; k < DIM; ++k) {
          W[1 + k][0] = w * delta * beta[k] * c;
          for(int l = 0; l < DIM; ++l) W[1 + k][1 + l] = w * (eps * M[k][l] + delta * beta[k] * beta[l]);
          for(int t = 0; t < NH; ++t) W[1 + k][NT + t] = -w * delta * beta[k] * eps * mlap[t];
          W[1 + k][NC] = w * delta * f * beta[k];
        }
        if constexpr(TCD) {
          s_WM[wave][qq][0] = w;
          for(int k = 0; k < DIM; ++k) s_WM[wave][qq][1 + k] = w * delta * beta[k];
        }
      } else {
        for(int r = 0; r < NT; ++r)
          for(int s = 0; s <= NC; ++s) W[r][s] = 0.0;
        if constexpr(TCD)
          for(int r = 0; r < NT; ++r) s_WM[wave][qq][r] = 0.0;
      }
    }
    __syncthreads();
    // stage 1: V[q*NT + r][j] = sum_s W[q][r][s] R[q][s][j]; j == NL -> rhs column
    if(live) {
      const int j = lane % 32, half = lane / 32;
      for(int qq = half; qq < QC; qq += 2) {
        double out[NT];
        for(int r = 0; r < NT; ++r) out[r] = 0.0;
        if(j < NL) {
#pragma unroll
          for(int s = 0; s < NC; ++s) {
            const double rv = s_R[qq][s][j];
#pragma unroll
            for(int r = 0; r < NT; ++r) out[r] += s_W[wave][qq][r][s] * rv;
          }
        } else if(j == NL) {
          for(int r = 0; r < NT; ++r) out[r] = s_W[wave][qq][r][NC];
        }
        for(int r = 0; r < NT; ++r) s_V[wave][qq * NT + r][j] = out[r];
        if constexpr(TCD) {
          // mass columns: WM[q][r] times the ansatz value
          const double rv = j < NL ? s_R[qq][0][j] : 0.0;
          for(int r = 0; r < NT; ++r) s_V[wave][qq * NT + r][kPad + j] = s_WM[wave][qq][r] * rv;
        }
      }
    }
    __syncthreads();
    // stage 2: A[i][j] += sum_k R^T[i][k] V[k][j], k = (point, test component)
    if(live) {
      const int l16 = lane % 16, kq = lane / 16;
      for(int k0 = 0; k0 < KC; k0 += 4) {
        const int k = k0 + kq, qq = k / NT, r = k % NT;
        double av[NIT], bv[NJT];
#pragma unroll
        for(int it = 0; it < NIT; ++it) av[it] = s_R[qq][r][it * 16 + l16];
#pragma unroll
        for(int jt = 0; jt < NJT; ++jt) bv[jt] = s_V[wave][k][jt * 16 + l16];
#pragma unroll
        for(int it = 0; it < NIT; ++it)
#pragma unroll
          for(int jt = 0; jt < NJT; ++jt)
            acc[it][jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[it], bv[jt], acc[it][jt], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  if(live) {
    // D tile layout: register v of lane l holds row l/16 + 4 v, column l%16
    double* A = a.local_A + (size_t)cell * NL * NL;
    double* F = a.local_rhs + (size_t)cell * NL;
    double* Mm = TCD ? a.local_M + (size_t)cell * NL * NL : nullptr;
    const int l16 = lane % 16, rg = lane / 16;
#pragma unroll
    for(int it = 0; it < NIT; ++it)
#pragma unroll
      for(int jt = 0; jt < NJT; ++jt)
#pragma unroll
        for(int v = 0; v < 4; ++v) {
          const int i = it * 16 + rg + 4 * v, j = jt * 16 + l16;
          if(i < NL) {
            if(j < NL) A[i * NL + j] = acc[it][jt][v];
            else if(j == NL) F[i] = acc[it][jt][v];
            else if constexpr(TCD) {
              if(j >= kPad && j < kPad + NL) Mm[i * NL + (j - kPad)] = acc[it][jt][v];
            }
          }
        }
  }
}

// ---- reference-tensor assembly ---------------------------------------------
//
// With constant matrix coefficients (eps, b, c) on an affine cell, W_K of the
// per-point form above is |det J| times a cell constant Wc_K[r][s] times the
// quadrature weight, so
//     A_K[i][j] = |det J| sum_p Wc_K[p] T_p[i][j],
//     T_p[i][j] = sum_q w_q R[q][r][i] R[q][s][j],   p = (r, s),
// over the NP component pairs with a nonzero coefficient (every (r, s) but
// value-test x second-derivative-ansatz). The T_p are cell independent
// (computed once on the host), and the whole level's local matrices are one
// GEMM [cells x NP] x [NP x NL^2] on FP64 MFMA. The right-hand side keeps the
// per-point source f: F_K[i] = sum_{q, r} g_q c_r R[q][r][i] with
// g_q = w_q |det J| f(x_q), c_0 = 1, c_{1+k} = delta beta_k, a second GEMM
// [cells x nq NT] x [nq NT x NL].
//
// Each wave owns 16 cells (the MFMA M dimension) on its own: 16 lanes build
// the cells' coefficients in the wave's part of LDS, the wave evaluates the
// source at the cells' points, then streams the NL^2 output columns in groups
// of TG 16-wide tiles with the coefficient operand held in registers. MQ
// bounds the rule's points (the source values of a wave's cells stay in LDS).
//
// TCD: also the mass matrix of the time-dependent forms (TCDMass +
// TCDMassSUPG, CD_local_assembling_routines.C:156-168,183-327), the same
// reference tensors with the coefficients |det J| on the value-value pair
// and |det J| delta beta_k on the pairs (d/dxi_k test, value trial), into
// local_M; the stiffness and rhs of TCDStiff(SUPG) / TCDRhs(SUPG) are the
// steady forms above.
template <int DIM, int NL, int MQ, bool TCD>
__global__ __launch_bounds__(256) void k_assemble_tensor(AsmArgs a) {
  constexpr int NC = DIM == 2 ? 6 : 10;
  constexpr int NT = 1 + DIM;
  constexpr int NH = NC - NT;
  constexpr int NP = NT * NC - NH;            // component pairs
  constexpr int NPP = (NP + 3) / 4 * 4;       // padded K
  constexpr int KS = NPP / 4;                 // K steps
  constexpr int NE = NL * NL;
  constexpr int NTILE = (NE + 15) / 16;
  constexpr int TG = NTILE < 6 ? NTILE : 6;    // tiles per group
  constexpr int RT = (NL + 15) / 16;          // rhs tiles
  __shared__ double s_W[4][16][NPP];
  __shared__ double s_c[4][16][NT];
  __shared__ double s_geo[4][16][13];         // c0[3], C[3][3], |det|
  __shared__ double s_g[4][16][MQ];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int cell0 = (blockIdx.x * 4 + wave) * 16;
  if(cell0 >= a.n_cells) return;  // whole wave: no block-wide barrier follows
  if(lane < 16) {
    const int cell = cell0 + lane;
    double* W = s_W[wave][lane];
    if(cell < a.n_cells) {
      GeoM g;
      geo_m(a.shape, a.cell_xyz + (size_t)cell * a.nv * 3, g);
      const double ad = fabs(g.det);
      for(int d = 0; d < 3; ++d) {
        s_geo[wave][lane][d] = g.c0[d];
        for(int k = 0; k < 3; ++k) s_geo[wave][lane][3 + 3 * d + k] = g.C[d][k];
      }
      s_geo[wave][lane][12] = ad;
      double M[3][3];
      for(int k = 0; k < 3; ++k)
        for(int l = 0; l < 3; ++l) {
          double t = 0;
          for(int d = 0; d < DIM; ++d) t += g.G[d][k] * g.G[d][l];
          M[k][l] = t;
        }
      double mlap[6];
      if(DIM == 2) { mlap[0] = M[0][0]; mlap[1] = 2 * M[0][1]; mlap[2] = M[1][1]; }
      else {
        mlap[0] = M[0][0]; mlap[1] = 2 * M[0][1]; mlap[2] = 2 * M[0][2];
        mlap[3] = M[1][1]; mlap[4] = 2 * M[1][2]; mlap[5] = M[2][2];
      }
      // the matrix coefficients are constant: any point of the cell
      double cf[6];
      example_coeffs(cell_params(a, cell), g.c0[0], g.c0[1], DIM == 3 ? g.c0[2] : 0.0, cf);
      const double eps = cf[0], c = cf[4];
      const double b[3] = {cf[1], cf[2], DIM == 3 ? cf[3] : 0.0};
      const double delta = a.supg ? sdfem_delta_m(DIM, g.hK, eps, b[0], b[1], b[2]) : 0.0;
      double beta[3];
      for(int k = 0; k < 3; ++k) {
        double t = 0;
        for(int d = 0; d < DIM; ++d) t += g.G[d][k] * b[d];
        beta[k] = t;
      }
      int p = 0;
      W[p++] = ad * c;
      for(int l = 0; l < DIM; ++l) W[p++] = ad * beta[l];
      for(int k = 0; k < DIM; ++k) {
        W[p++] = ad * (delta * beta[k] * c);
        for(int l = 0; l < DIM; ++l) W[p++] = ad * (eps * M[k][l] + delta * beta[k] * beta[l]);
        for(int t = 0; t < NH; ++t) W[p++] = ad * (-delta * beta[k] * eps * mlap[t]);
      }
      for(; p < NPP; ++p) W[p] = 0.0;
      s_c[wave][lane][0] = 1.0;
      for(int k = 0; k < DIM; ++k) s_c[wave][lane][1 + k] = delta * beta[k];
    } else {
      for(int p = 0; p < NPP; ++p) W[p] = 0.0;
      for(int r = 0; r < NT; ++r) s_c[wave][lane][r] = 0.0;
      for(int t = 0; t < 13; ++t) s_geo[wave][lane][t] = 0.0;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // g_q = w_q |det J| f(x_q) for the wave's cells and points
  for(int t = lane; t < 16 * a.nq; t += 64) {
    const int cl = t / a.nq, q = t % a.nq, cell = cell0 + cl;
    double gq = 0.0;
    if(cell < a.n_cells) {
      const double* gm = s_geo[wave][cl];
      const double xi = a.qp[3 * q], eta = a.qp[3 * q + 1], zeta = a.qp[3 * q + 2];
      double x[3];
      for(int d = 0; d < 3; ++d) {
        x[d] = gm[d] + gm[3 + 3 * d] * xi + gm[3 + 3 * d + 1] * eta;
        if(DIM == 3) x[d] = x[d] + gm[3 + 3 * d + 2] * zeta;
      }
      double cf[6];
      example_coeffs(cell_params(a, cell), x[0], x[1], DIM == 3 ? x[2] : 0.0, cf);
      gq = a.qw[q] * gm[12] * cf[5];
    }
    s_g[wave][cl][q] = gq;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int l16 = lane % 16, kq = lane / 16;
  // local matrices: D[cell][e] = sum_p W[cell][p] T[p][e] (and the mass
  // matrices with Wm)
#pragma unroll 1
  for(int pass = 0; pass < (TCD ? 2 : 1); ++pass) {
  double av[KS];
#pragma unroll
  for(int ks = 0; ks < KS; ++ks) av[ks] = s_W[wave][l16][4 * ks + kq];
  if(TCD && pass == 0) {
    // the mass coefficients replace the stiffness ones in LDS (every lane
    // holds its operand in registers by now): |det J| on the value pair,
    // |det J| delta beta_k on the pairs (d/dxi_k test, value trial)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if(lane < 16) {
      double* W = s_W[wave][lane];
      const double ad = s_geo[wave][lane][12];
      for(int p = 0; p < NPP; ++p) W[p] = 0.0;
      W[0] = ad;
      for(int k = 0; k < DIM; ++k) W[1 + DIM + k * NC] = ad * s_c[wave][lane][1 + k];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  double* const out = pass == 0 ? a.local_A : a.local_M;
  // the tensor rows of K step ks + 1 are loaded while the MFMAs of step ks run
  // (the tensor is padded to whole tiles: every load is in bounds)
  const double* Tbase = a.tensor + (size_t)kq * (NTILE * 16) + l16;
  auto tile_group = [&](int t0) {
    v4d acc[TG];
    double bcur[TG], bnext[TG];
#pragma unroll
    for(int u = 0; u < TG; ++u) {
      acc[u] = v4d{0, 0, 0, 0};
      bcur[u] = t0 + u < NTILE ? Tbase[(t0 + u) * 16] : 0.0;
    }
#pragma unroll
    for(int ks = 0; ks < KS; ++ks) {
      if(ks + 1 < KS) {
        const double* Tn = Tbase + (size_t)(4 * (ks + 1)) * (NTILE * 16);
#pragma unroll
        for(int u = 0; u < TG; ++u) bnext[u] = t0 + u < NTILE ? Tn[(t0 + u) * 16] : 0.0;
      }
#pragma unroll
      for(int u = 0; u < TG; ++u)
        if(t0 + u < NTILE) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks], bcur[u], acc[u], 0, 0, 0);
#pragma unroll
      for(int u = 0; u < TG; ++u) bcur[u] = bnext[u];
    }
#pragma unroll
    for(int u = 0; u < TG; ++u) {
      const int e = (t0 + u) * 16 + l16;
      if(t0 + u >= NTILE || e >= NE) continue;
#pragma unroll
      for(int v = 0; v < 4; ++v) {
        const int cell = cell0 + kq + 4 * v;
        if(cell < a.n_cells) out[(size_t)cell * NE + e] = acc[u][v];
      }
    }
  };
  if constexpr(TCD) {
    // the time-dependent kernel (two passes) keeps the tile loop rolled:
    // unrolled, it spills registers
#pragma unroll 1
    for(int t0 = 0; t0 < NTILE; t0 += TG) tile_group(t0);
  } else {
    for(int t0 = 0; t0 < NTILE; t0 += TG) tile_group(t0);
  }
  }
  // right-hand sides: F[cell][i] = sum_{q, r} g[cell][q] c[cell][r] R[q][r][i],
  // the reference-table operands of 8 K steps loaded before their MFMAs
  const int K = a.nq * NT;
#pragma unroll
  for(int rt = 0; rt < RT; ++rt) {
    v4d acc = {0, 0, 0, 0};
    const int i = rt * 16 + l16;
    for(int k0 = 0; k0 < K; k0 += 32) {
      double bv[8], avr[8];
#pragma unroll
      for(int u = 0; u < 8; ++u) {
        const int k = k0 + 4 * u + kq, q = k / NT, r = k % NT;
        bv[u] = (k < K && i < NL) ? a.ref[((size_t)q * NC + r) * NL + i] : 0.0;
        avr[u] = k < K ? s_g[wave][l16][q] * s_c[wave][l16][r] : 0.0;
      }
#pragma unroll
      for(int u = 0; u < 8; ++u)
        if(k0 + 4 * u < K) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(avr[u], bv[u], acc, 0, 0, 0);
    }
    if(i < NL)
#pragma unroll
      for(int v = 0; v < 4; ++v) {
        const int cell = cell0 + kq + 4 * v;
        if(cell < a.n_cells) a.local_rhs[(size_t)cell * NL + i] = acc[v];
      }
  }
}

// lane-layout check of the 16x16x4 FP64 MFMA: D = A B with A[i][k] = i + 100 k,
// B[k][j] = (k == j % 4) ? 1 + j : 0; out[lane*4+v] = D register v of lane
__global__ void k_mfma_probe(double* out) {
  const int lane = threadIdx.x;
  const int i = lane % 16, k = lane / 16;
  const double av = i + 100.0 * k;
  const int j = lane % 16;
  const double bv = (k == j % 4) ? 1.0 + j : 0.0;
  v4d acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  for(int v = 0; v < 4; ++v) out[lane * 4 + v] = acc[v];
}

}  // namespace

bool launch_assemble_mfma(const AsmArgs& a, hipStream_t s) {
  if(a.n_cells == 0) return true;
  // Only a local matrix that fills the 16x16 tiles pays for the MFMA path:
  // measured per element (SUPG, 2^18-2^21 cells), the vector kernels are
  // faster for every element with fewer than 16 local DOFs (Q1 hexahedra
  // 12.9 vs 33.2 ms, P2 tetrahedra 1.2 vs 1.7 ms, P2 triangles 2.1 vs 3.6 ms,
  // Q2 quadrilaterals within 5%), the MFMA kernel for Q2 hexahedra (10.0 vs
  // 27.7 ms at 64 points).
  if(a.nl < 16) return false;
  if(a.tcd && !a.local_M) return false;
  const int blocks = (a.n_cells + kCells - 1) / kCells;
  if(a.dim == 3 && a.ncomp == 10 && a.nl == 27) {
    if(a.tcd)
      k_assemble_mfma<3, 27, true><<<blocks, 256, 0, s>>>(a);
    else
      k_assemble_mfma<3, 27, false><<<blocks, 256, 0, s>>>(a);
  } else {
    return false;
  }
  PMCD_HIP(hipGetLastError());
  return true;
}

int tensor_pairs(int dim) { return dim == 2 ? 3 * 6 - 3 : 4 * 10 - 6; }

std::vector<double> build_assembly_tensor(int dim, int nl, int nq, const double* ref, const double* qw) {
  const int NC = dim == 2 ? 6 : 10, NT = 1 + dim, NH = NC - NT;
  const int np = tensor_pairs(dim), npp = (np + 3) / 4 * 4;
  const int ne = nl * nl, nep = (ne + 15) / 16 * 16;
  std::vector<double> T((size_t)npp * nep, 0.0);
  int p = 0;
  for(int r = 0; r < NT; ++r)
    for(int s = 0; s < NC; ++s) {
      if(r == 0 && s >= NT) continue;
      for(int i = 0; i < nl; ++i)
        for(int j = 0; j < nl; ++j) {
          double t = 0.0;
          for(int q = 0; q < nq; ++q)
            t += qw[q] * ref[((size_t)q * NC + r) * nl + i] * ref[((size_t)q * NC + s) * nl + j];
          T[(size_t)p * nep + i * nl + j] = t;
        }
      ++p;
    }
  (void)NH;
  return T;
}

template <int MQ, bool TCD>
static bool launch_tensor_mq(const AsmArgs& a, int blocks, hipStream_t s) {
  if(a.dim == 3) {
    switch(a.nl) {
      case 27: k_assemble_tensor<3, 27, MQ, TCD><<<blocks, 256, 0, s>>>(a); break;
      case 8: k_assemble_tensor<3, 8, MQ, TCD><<<blocks, 256, 0, s>>>(a); break;
      case 10: k_assemble_tensor<3, 10, MQ, TCD><<<blocks, 256, 0, s>>>(a); break;
      case 4: k_assemble_tensor<3, 4, MQ, TCD><<<blocks, 256, 0, s>>>(a); break;
      default: return false;
    }
  } else {
    switch(a.nl) {
      case 9: k_assemble_tensor<2, 9, MQ, TCD><<<blocks, 256, 0, s>>>(a); break;
      case 6: k_assemble_tensor<2, 6, MQ, TCD><<<blocks, 256, 0, s>>>(a); break;
      case 4: k_assemble_tensor<2, 4, MQ, TCD><<<blocks, 256, 0, s>>>(a); break;
      case 3: k_assemble_tensor<2, 3, MQ, TCD><<<blocks, 256, 0, s>>>(a); break;
      default: return false;
    }
  }
  PMCD_HIP(hipGetLastError());
  return true;
}

bool launch_assemble_tensor(const AsmArgs& a, hipStream_t s) {
  if(a.n_cells == 0) return true;
  if(!a.tensor || a.nq > 64 || (a.tcd && !a.local_M)) return false;
  // P1 simplices: the thread-per-cell kernel measured faster (P1 tetrahedra
  // 0.65 vs 0.83 ms for 1.57M cells, P1 triangles 0.45 vs 0.85 ms for 2.1M)
  if(p1_simplex(a)) return false;
  const int blocks = (a.n_cells + 63) / 64;
  if(a.tcd)
    return a.nq <= 32 ? launch_tensor_mq<32, true>(a, blocks, s) : launch_tensor_mq<64, true>(a, blocks, s);
  return a.nq <= 32 ? launch_tensor_mq<32, false>(a, blocks, s) : launch_tensor_mq<64, false>(a, blocks, s);
}

void assemble_cells(const AsmArgs& a, hipStream_t s) {
  const int v = assembly_variant();
  if(v == 2 && launch_assemble_tensor(a, s)) return;
  if(v >= 1 && launch_assemble_mfma(a, s)) return;
  launch_assemble_local(a, s);
}

void launch_mfma_probe(double* out, hipStream_t s) {
  k_mfma_probe<<<1, 64, 0, s>>>(out);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd
