// Cell-loop assembly of the convection-diffusion (Galerkin + SUPG) local
// matrices, the deterministic row gather into CSR, and the error integrals.
//
// Reference hot loop: src/AssembleRoutines/Assemble3D.C:183-769 (2D:
// Assemble2D.C:2798-3430) -> FEDatabase::GetOrig (reference -> physical basis
// values, src/FE/HexaAffin.C:73-219, TetraAffin.C, TriaAffin.C, QuadAffin.C)
// -> LocalAssembling::GetLocalForms (LocalAssembling.C:734-970) calling, per
// quadrature point in rule order, BilinearAssembleGalerkin<d> and
// BilinearAssemble_SD<d> (ConvDiff.C:1776-1999) with the SUPG parameter of
// Compute_SDFEM_delta<d> (ConvDiff.C:269-340, SDFEM_TYPE 2).
//
// Layout: one 64-lane wave per cell, four cells per 256-thread block. Per
// cell the wave evaluates geometry and coefficients for all quadrature points
// (lanes over points), then streams the points in chunks through LDS as
// physical basis values (lanes over point x basis function) and accumulates
// the local matrix entries in registers (lanes over entries). Each entry is
// accumulated over the points in rule order with the reference's operation
// order, so the result does not depend on the launch shape.
#include "kernels.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace pmcd {

namespace {

constexpr int kWave = 64;
constexpr int kCellsPerBlock = 4;
constexpr int kMaxQ = 64;
constexpr int kQChunk = 8;

struct Geo {
  double c0[3];
  double C[3][3];   // C[a][k] = d x_a / d xi_k
  double K[3][3];   // cofactor numerators: d xi_k / d x_a = K[a][k] / det
  double det;
  double hK;
};

// Affine reference map of a cell (TriaAffin.C:197-227, QuadAffin.C:196-204,
// TetraAffin.C:246-262, HexaAffin.C:221-250) and its diameter.
__device__ inline void cell_geometry(int shape, const double* v, Geo& g) {
  auto X = [&](int i, int k) { return v[3 * i + k]; };
  for(int a = 0; a < 3; ++a)
    for(int k = 0; k < 3; ++k) { g.C[a][k] = 0; g.K[a][k] = 0; }
  if(shape == (int)Shape::Triangle || shape == (int)Shape::Quadrangle) {
    for(int a = 0; a < 2; ++a) {
      if(shape == (int)Shape::Triangle) {
        g.c0[a] = X(0, a);
        g.C[a][0] = X(1, a) - X(0, a);
        g.C[a][1] = X(2, a) - X(0, a);
      } else {
        g.c0[a] = (X(1, a) + X(3, a)) * 0.5;
        g.C[a][0] = (X(1, a) - X(0, a)) * 0.5;
        g.C[a][1] = (X(3, a) - X(0, a)) * 0.5;
      }
    }
    g.c0[2] = 0;
    const double xc1 = g.C[0][0], xc2 = g.C[0][1], yc1 = g.C[1][0], yc2 = g.C[1][1];
    g.det = xc1 * yc2 - xc2 * yc1;
    g.K[0][0] = yc2;  g.K[0][1] = -yc1;
    g.K[1][0] = -xc2; g.K[1][1] = xc1;
    double d2;
    if(shape == (int)Shape::Triangle) {
      double e[3];
      const int ed[3][2] = {{0, 1}, {1, 2}, {2, 0}};
      for(int m = 0; m < 3; ++m) {
        double dx = X(ed[m][0], 0) - X(ed[m][1], 0), dy = X(ed[m][0], 1) - X(ed[m][1], 1),
               dz = X(ed[m][0], 2) - X(ed[m][1], 2);
        e[m] = dx * dx + dy * dy + dz * dz;
      }
      d2 = fmax(e[0], fmax(e[1], e[2]));
    } else {
      double dx1 = X(0, 0) - X(2, 0), dy1 = X(0, 1) - X(2, 1), dz1 = X(0, 2) - X(2, 2);
      double dx2 = X(1, 0) - X(3, 0), dy2 = X(1, 1) - X(3, 1), dz2 = X(1, 2) - X(3, 2);
      d2 = fmax(dx1 * dx1 + dy1 * dy1 + dz1 * dz1, dx2 * dx2 + dy2 * dy2 + dz2 * dz2);
    }
    g.hK = sqrt(d2);
    return;
  }
  for(int a = 0; a < 3; ++a) {
    if(shape == (int)Shape::Tetrahedron) {
      g.c0[a] = X(0, a);
      g.C[a][0] = X(1, a) - X(0, a);
      g.C[a][1] = X(2, a) - X(0, a);
      g.C[a][2] = X(3, a) - X(0, a);
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
  g.det = xc1 * yc2 * zc3 + xc2 * yc3 * zc1 + xc3 * yc1 * zc2 - xc3 * yc2 * zc1 -
          xc2 * yc1 * zc3 - xc1 * yc3 * zc2;
  g.K[0][0] = yc2 * zc3 - yc3 * zc2; g.K[0][1] = yc3 * zc1 - yc1 * zc3; g.K[0][2] = yc1 * zc2 - yc2 * zc1;
  g.K[1][0] = xc3 * zc2 - xc2 * zc3; g.K[1][1] = xc1 * zc3 - xc3 * zc1; g.K[1][2] = xc2 * zc1 - xc1 * zc2;
  g.K[2][0] = xc2 * yc3 - xc3 * yc2; g.K[2][1] = xc3 * yc1 - xc1 * yc3; g.K[2][2] = xc1 * yc2 - xc2 * yc1;
  double d2 = 0;
  if(shape == (int)Shape::Tetrahedron) {
    const int ed[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for(int m = 0; m < 6; ++m) {
      double dx = X(ed[m][0], 0) - X(ed[m][1], 0), dy = X(ed[m][0], 1) - X(ed[m][1], 1),
             dz = X(ed[m][0], 2) - X(ed[m][1], 2);
      double l = dx * dx + dy * dy + dz * dz;
      if(l > d2) d2 = l;
    }
  } else {
    const int dg[4][2] = {{0, 6}, {1, 7}, {2, 4}, {3, 5}};
    for(int m = 0; m < 4; ++m) {
      double dx = X(dg[m][0], 0) - X(dg[m][1], 0), dy = X(dg[m][0], 1) - X(dg[m][1], 1),
             dz = X(dg[m][0], 2) - X(dg[m][1], 2);
      double l = dx * dx + dy * dy + dz * dz;
      if(l > d2) d2 = l;
    }
  }
  g.hK = sqrt(d2);
}

// Compute_SDFEM_delta<d>, SDFEM_TYPE 2, DELTA0 = 1, polynomial degree 1.
__device__ inline double sdfem_delta(int dim, double hK, double eps, double b1, double b2,
                                     double b3) {
  if(fabs(eps) < 1e-20) eps = 1e-20;
  double norm_b = b1 * b1 + b2 * b2;
  if(dim == 3) norm_b += b3 * b3;
  norm_b = sqrt(norm_b);
  const double nu = 1.0 / 1;
  double delta;
  if(norm_b > 0) {
    double alpha = nu * norm_b * hK / (2 * eps);
    delta = nu * hK * (1 / tanh(alpha) - 1 / alpha) / (2 * norm_b);
  } else {
    delta = 0;
  }
  delta *= 1.0;
  return delta;
}

// physical value, gradient and Laplacian of reference basis function i at
// point q; second derivatives use J^-T H J^-1 of the affine map.
__device__ inline void phys_values(const AsmArgs& a, const Geo& g, int q, int i, bool need2,
                                   double* out /*u, ux, uy, uz, lap*/) {
  const double* r = a.ref + (size_t)q * a.ncomp * a.nl;
  const int nl = a.nl;
  out[0] = r[i];
  if(a.dim == 2) {
    const double rx = r[nl + i], ry = r[2 * nl + i];
    out[1] = (g.K[0][0] * rx + g.K[0][1] * ry) / g.det;
    out[2] = (g.K[1][0] * rx + g.K[1][1] * ry) / g.det;
    out[3] = 0;
    if(need2) {
      const double h[3] = {r[3 * nl + i], r[4 * nl + i], r[5 * nl + i]};  // 20, 11, 02
      double uaa[2];
      for(int d = 0; d < 2; ++d) {
        const double j0 = g.K[d][0] / g.det, j1 = g.K[d][1] / g.det;
        uaa[d] = j0 * j0 * h[0] + 2 * j0 * j1 * h[1] + j1 * j1 * h[2];
      }
      out[4] = uaa[0] + uaa[1];
    } else out[4] = 0;
  } else {
    const double rx = r[nl + i], ry = r[2 * nl + i], rz = r[3 * nl + i];
    if(a.shape == (int)Shape::Tetrahedron) {
      for(int d = 0; d < 3; ++d)
        out[1 + d] = (g.K[d][0] / g.det) * rx + (g.K[d][1] / g.det) * ry + (g.K[d][2] / g.det) * rz;
    } else {
      for(int d = 0; d < 3; ++d)
        out[1 + d] = (g.K[d][0] * rx + g.K[d][1] * ry + g.K[d][2] * rz) / g.det;
    }
    if(need2) {
      // 200, 110, 101, 020, 011, 002
      const double h[6] = {r[4 * nl + i], r[5 * nl + i], r[6 * nl + i],
                           r[7 * nl + i], r[8 * nl + i], r[9 * nl + i]};
      double uaa[3];
      for(int d = 0; d < 3; ++d) {
        const double j0 = g.K[d][0] / g.det, j1 = g.K[d][1] / g.det, j2 = g.K[d][2] / g.det;
        uaa[d] = j0 * j0 * h[0] + 2 * j0 * j1 * h[1] + 2 * j0 * j2 * h[2] + j1 * j1 * h[3] +
                 2 * j1 * j2 * h[4] + j2 * j2 * h[5];
      }
      out[4] = uaa[0] + uaa[1] + uaa[2];
    } else out[4] = 0;
  }
}

template <int NL>
__global__ __launch_bounds__(256) void k_assemble_local(AsmArgs a) {
  constexpr int EPL = (NL * NL + kWave - 1) / kWave;  // entries per lane
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int cell = blockIdx.x * kCellsPerBlock + wave;
  const bool live = cell < a.n_cells;
  __shared__ double s_coef[kCellsPerBlock][kMaxQ][8];
  __shared__ double s_phys[kCellsPerBlock][kQChunk][5][NL];

  Geo g;
  if(live) cell_geometry(a.shape, a.cell_xyz + (size_t)cell * a.nv * 3, g);
  const bool need2 = a.supg != 0;
  // per-point weight, coefficients and SUPG parameter
  if(live && lane < a.nq) {
    const int q = lane;
    const double xi = a.qp[3 * q], eta = a.qp[3 * q + 1], zeta = a.qp[3 * q + 2];
    double x[3];
    for(int d = 0; d < 3; ++d) {
      x[d] = g.c0[d] + g.C[d][0] * xi + g.C[d][1] * eta;
      if(a.dim == 3) x[d] = x[d] + g.C[d][2] * zeta;
    }
    const double mult = a.qw[q] * fabs(g.det);
    double cf[6];
    example_coeffs(cell_params(a, cell), x[0], x[1], a.dim == 3 ? x[2] : 0.0, cf);
    double* s = s_coef[wave][q];
    s[0] = mult;
    for(int k = 0; k < 6; ++k) s[1 + k] = cf[k];
    s[7] = need2 ? sdfem_delta(a.dim, g.hK, cf[0], cf[1], cf[2], cf[3]) : 0.0;
  }
  double acc[EPL], accm[EPL];
  for(int t = 0; t < EPL; ++t) acc[t] = accm[t] = 0.0;
  double rhs = 0.0;
  const bool tcd = a.tcd != 0;
  __syncthreads();

  for(int q0 = 0; q0 < a.nq; q0 += kQChunk) {
    if(live) {
      for(int p = lane; p < kQChunk * NL; p += kWave) {
        const int qq = p / NL, i = p % NL, q = q0 + qq;
        if(q < a.nq) {
          double o[5];
          phys_values(a, g, q, i, need2, o);
          for(int k = 0; k < 5; ++k) s_phys[wave][qq][k][i] = o[k];
        }
      }
    }
    __syncthreads();
    if(live) {
      const int qe = min(kQChunk, a.nq - q0);
      for(int qq = 0; qq < qe; ++qq) {
        const double* cf = s_coef[wave][q0 + qq];
        const double Mult = cf[0], eps = cf[1], b1 = cf[2], b2 = cf[3], b3 = cf[4], c = cf[5],
                     f = cf[6], delta = cf[7];
        const double(*P)[NL] = s_phys[wave][qq];
#pragma unroll
        for(int t = 0; t < EPL; ++t) {
          const int e = lane + kWave * t;
          if(e < NL * NL) {
            const int i = e / NL, j = e % NL;
            const double t0 = P[0][i], tx = P[1][i], ty = P[2][i], tz = P[3][i];
            const double a0 = P[0][j], ax = P[1][j], ay = P[2][j], az = P[3][j];
            if(tcd) {
              // TCDStiff<d> (CD_local_assembling_routines.C:9-45; in 2D the
              // z components are 0)
              double val = eps * (tx * ax + ty * ay + tz * az);
              val += (b1 * ax + b2 * ay + b3 * az) * t0;
              val += c * t0 * a0;
              acc[t] += Mult * val;
              // TCDMass<d> (:156-168)
              accm[t] += Mult * t0 * a0;
              if(need2) {
                // TCDStiffSUPG<d> / TCDMassSUPG<d> (:183-327) with the
                // unscaled SDFEM type 2 parameter
                const double bgradv = delta * (b1 * tx + b2 * ty + b3 * tz);
                double v2 = (b1 * ax + b2 * ay + b3 * az + c * a0) * bgradv;
                v2 -= eps * P[4][j] * bgradv;
                acc[t] += Mult * v2;
                accm[t] += Mult * a0 * bgradv;
              }
            } else if(a.dim == 3) {
              // BilinearAssembleGalerkin<3> (ConvDiff.C:1842-1883)
              double val = eps * (tx * ax + ty * ay + tz * az);
              val += (b1 * ax + b2 * ay + b3 * az) * t0;
              val += c * a0 * t0;
              acc[t] += val * Mult;
              if(need2) {
                // BilinearAssemble_SD<3> (ConvDiff.C:1951-1999)
                const double bgradv = b1 * tx + b2 * ty + b3 * tz;
                double v2 = -eps * P[4][j];
                v2 += b1 * ax + b2 * ay + b3 * az + c * a0;
                acc[t] += v2 * delta * bgradv * Mult;
              }
            } else {
              // BilinearAssembleGalerkin<2> (ConvDiff.C:1776-1839)
              const double c3 = c;
              double val = eps * (tx * ax + ty * ay);
              val += (b1 * ax + b2 * ay) * t0;
              val += c3 * a0 * t0;
              val *= Mult;
              acc[t] += val;
              if(need2) {
                // BilinearAssemble_SD<2> (ConvDiff.C:1886-1948)
                const double bgradv = b1 * tx + b2 * ty;
                double v2 = delta * (-eps * P[4][j] + b1 * ax + b2 * ay + c3 * a0) * bgradv;
                v2 *= Mult;
                acc[t] += v2;
              }
            }
          }
        }
        if(lane < NL) {
          const int i = lane;
          const double t0 = P[0][i], tx = P[1][i], ty = P[2][i], tz = P[3][i];
          if(tcd) {
            // TCDRhs<d> (:170-181) and TCDRhsSUPG<d> (:328-387)
            rhs += Mult * t0 * f;
            if(need2) rhs += Mult * f * (delta * (b1 * tx + b2 * ty + b3 * tz));
          } else if(a.dim == 3) {
            rhs += Mult * t0 * f;
            if(need2) {
              const double bgradv = b1 * tx + b2 * ty + b3 * tz;
              rhs += Mult * (delta * bgradv) * f;
            }
          } else {
            const double f2 = f;
            rhs += Mult * t0 * f2;
            if(need2) {
              const double bgradv = b1 * tx + b2 * ty;
              rhs += Mult * delta * bgradv * f2;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if(live) {
    double* A = a.local_A + (size_t)cell * NL * NL;
    double* Mm = a.local_M ? a.local_M + (size_t)cell * NL * NL : nullptr;
#pragma unroll
    for(int t = 0; t < EPL; ++t) {
      const int e = lane + kWave * t;
      if(e < NL * NL) {
        A[e] = acc[t];
        if(Mm) Mm[e] = accm[t];
      }
    }
    if(lane < NL) a.local_rhs[(size_t)cell * NL + lane] = rhs;
  }
}

// One wave per row: zero the row in LDS, add the local rows of the cells
// containing the row DOF in increasing cell order, write the row out.
__global__ __launch_bounds__(64) void k_gather_matrix(int n_rows, int n_active, int nl,
                                                       const int* __restrict__ row_ptr,
                                                       const int* __restrict__ col,
                                                       const int* __restrict__ contrib_ptr,
                                                       const int* __restrict__ contrib,
                                                       const uint8_t* __restrict__ slot,
                                                       const double* __restrict__ local_A,
                                                       double* __restrict__ val,
                                                       const int* __restrict__ sell_slot,
                                                       const long long* __restrict__ sell_off,
                                                       double* __restrict__ sell_val) {
  __shared__ double s_row[256];
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const int r0 = row_ptr[r], len = row_ptr[r + 1] - r0;
  if(r >= n_active) {
    for(int k = lane; k < len; k += 64) val[r0 + k] = (col[r0 + k] == r) ? 1.0 : 0.0;
    return;
  }
  for(int k = lane; k < len; k += 64) s_row[k] = 0.0;
  __syncthreads();
  const int e1 = contrib_ptr[r + 1];
  for(int e = contrib_ptr[r]; e < e1; ++e) {
    const int cm = contrib[e];
    const int c = cm / nl, m = cm - c * nl;
    if(lane < nl) {
      const double v = local_A[((size_t)c * nl + m) * nl + lane];
      const int p = slot[(size_t)e * nl + lane];
      s_row[p] += v;
    }
    __syncthreads();
  }
  for(int k = lane; k < len; k += 64) val[r0 + k] = s_row[k];
  if(sell_val) {
    // the same row into its sliced ELL slot (entry k at base + 64 k)
    const int t = sell_slot[r];
    const long long base = sell_off[t / kSellRows] + (t % kSellRows);
    for(int k = lane; k < len; k += 64) sell_val[base + (long long)k * kSellRows] = s_row[k];
  }
}

__global__ void k_gather_rhs(int n_rows, int n_active, const int* __restrict__ contrib_ptr,
                             const int* __restrict__ contrib, const double* __restrict__ local_rhs,
                             const int* __restrict__ extra_ptr, const double* __restrict__ extra,
                             const double* __restrict__ dir_vals, double* __restrict__ rhs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= n_rows) return;
  if(r >= n_active) { rhs[r] = dir_vals[r - n_active]; return; }
  double s = 0.0;
  for(int e = contrib_ptr[r]; e < contrib_ptr[r + 1]; ++e) {
    const int cm = contrib[e];
    s += local_rhs[cm];
    if(extra_ptr)
      for(int t = extra_ptr[cm]; t < extra_ptr[cm + 1]; ++t) s += extra[t];
  }
  rhs[r] = s;
}

// Error integrals of one cell (FEFunction3D.C:60-248 with
// conv_diff_l2_h1_linf_error<d>, ConvDiff.C:2124-2183): lanes over points,
// then lane 0 sums the point contributions in rule order.
// P1 triangles and tetrahedra: one thread per cell, the whole local matrix
// ((DIM+1)^2 entries) in registers. Same per-entry operation order as
// k_assemble_local (the reference's), so both give the same bits; the
// Laplacian of a P1 basis function is zero.
template <int DIM>
__global__ __launch_bounds__(256) void k_assemble_p1(AsmArgs a) {
  constexpr int NL = DIM + 1;
  const int cell = blockIdx.x * 256 + threadIdx.x;
  if(cell >= a.n_cells) return;
  Geo g;
  cell_geometry(a.shape, a.cell_xyz + (size_t)cell * a.nv * 3, g);
  const bool supg = a.supg != 0;
  const ExampleParams& ex = cell_params(a, cell);
  double acc[NL * NL], rhs[NL];
#pragma unroll
  for(int e = 0; e < NL * NL; ++e) acc[e] = 0.0;
#pragma unroll
  for(int i = 0; i < NL; ++i) rhs[i] = 0.0;
  for(int q = 0; q < a.nq; ++q) {
    const double xi = a.qp[3 * q], eta = a.qp[3 * q + 1], zeta = a.qp[3 * q + 2];
    double x[3];
    for(int d = 0; d < 3; ++d) {
      x[d] = g.c0[d] + g.C[d][0] * xi + g.C[d][1] * eta;
      if(DIM == 3) x[d] = x[d] + g.C[d][2] * zeta;
    }
    const double Mult = a.qw[q] * fabs(g.det);
    double cf[6];
    example_coeffs(ex, x[0], x[1], DIM == 3 ? x[2] : 0.0, cf);
    const double eps = cf[0], b1 = cf[1], b2 = cf[2], b3 = cf[3], c = cf[4], f = cf[5];
    const double delta = supg ? sdfem_delta(DIM, g.hK, eps, b1, b2, b3) : 0.0;
    double P[4][NL];
#pragma unroll
    for(int i = 0; i < NL; ++i) {
      double o[5];
      phys_values(a, g, q, i, false, o);
      for(int k = 0; k < 4; ++k) P[k][i] = o[k];
    }
#pragma unroll
    for(int i = 0; i < NL; ++i) {
      const double t0 = P[0][i], tx = P[1][i], ty = P[2][i], tz = P[3][i];
#pragma unroll
      for(int j = 0; j < NL; ++j) {
        const double a0 = P[0][j], ax = P[1][j], ay = P[2][j], az = P[3][j];
        double& A = acc[i * NL + j];
        if(DIM == 3) {
          double val = eps * (tx * ax + ty * ay + tz * az);
          val += (b1 * ax + b2 * ay + b3 * az) * t0;
          val += c * a0 * t0;
          A += val * Mult;
          if(supg) {
            const double bgradv = b1 * tx + b2 * ty + b3 * tz;
            double v2 = -eps * 0.0;
            v2 += b1 * ax + b2 * ay + b3 * az + c * a0;
            A += v2 * delta * bgradv * Mult;
          }
        } else {
          double val = eps * (tx * ax + ty * ay);
          val += (b1 * ax + b2 * ay) * t0;
          val += c * a0 * t0;
          val *= Mult;
          A += val;
          if(supg) {
            const double bgradv = b1 * tx + b2 * ty;
            double v2 = delta * (-eps * 0.0 + b1 * ax + b2 * ay + c * a0) * bgradv;
            v2 *= Mult;
            A += v2;
          }
        }
      }
      if(DIM == 3) {
        rhs[i] += Mult * t0 * f;
        if(supg) {
          const double bgradv = b1 * tx + b2 * ty + b3 * tz;
          rhs[i] += Mult * (delta * bgradv) * f;
        }
      } else {
        rhs[i] += Mult * t0 * f;
        if(supg) {
          const double bgradv = b1 * tx + b2 * ty;
          rhs[i] += Mult * delta * bgradv * f;
        }
      }
    }
  }
  double* A = a.local_A + (size_t)cell * NL * NL;
#pragma unroll
  for(int e = 0; e < NL * NL; ++e) A[e] = acc[e];
#pragma unroll
  for(int i = 0; i < NL; ++i) a.local_rhs[(size_t)cell * NL + i] = rhs[i];
}

// cell_err[5 * cell + k]: L2^2, H1-semi^2, SD^2, DG (volume part), Linf.
__global__ __launch_bounds__(64) void k_cell_errors(AsmArgs a, const int* __restrict__ dof,
                                                     const double* __restrict__ u,
                                                     double* __restrict__ cell_err) {
  const int cell = blockIdx.x, lane = threadIdx.x;
  __shared__ double s_U[32];
  __shared__ double s_e[kMaxQ][7];
  Geo g;
  cell_geometry(a.shape, a.cell_xyz + (size_t)cell * a.nv * 3, g);
  if(lane < a.nl) s_U[lane] = u[dof[(size_t)cell * a.nl + lane]];
  __syncthreads();
  if(lane < a.nq) {
    const int q = lane;
    double v[4] = {0, 0, 0, 0};
    for(int l = 0; l < a.nl; ++l) {
      double o[5];
      phys_values(a, g, q, l, false, o);
      for(int k = 0; k < 4; ++k) v[k] += s_U[l] * o[k];
    }
    const double xi = a.qp[3 * q], eta = a.qp[3 * q + 1], zeta = a.qp[3 * q + 2];
    double x[3];
    for(int d = 0; d < 3; ++d) {
      x[d] = g.c0[d] + g.C[d][0] * xi + g.C[d][1] * eta;
      if(a.dim == 3) x[d] = x[d] + g.C[d][2] * zeta;
    }
    const double w = a.qw[q] * fabs(g.det);
    double ex[5], cf[6];
    const double z = a.dim == 3 ? x[2] : 0.0;
    example_exact(cell_params(a, cell), x[0], x[1], z, ex);
    example_coeffs(cell_params(a, cell), x[0], x[1], z, cf);
    double nu, b1, b2, b3, c;
    if(a.dim == 3) { nu = cf[0]; b1 = cf[1]; b2 = cf[2]; b3 = cf[3]; c = cf[4]; }
    else { nu = cf[0]; b1 = cf[1]; b2 = cf[2]; b3 = 0; c = cf[4]; }
    const double delta = sdfem_delta(a.dim, g.hK, nu, b1, b2, b3);
    const double e0 = v[0] - ex[0];
    const double e1 = v[1] - ex[1], e2 = v[2] - ex[2];
    const double e3 = a.dim == 2 ? 0. : (v[3] - ex[3]);
    double e4 = b1 * e1 + b2 * e2;
    if(a.dim == 3) e4 += b3 * e3;
    s_e[q][0] = w * e0 * e0;
    s_e[q][1] = w * e1 * e1;
    s_e[q][2] = w * e2 * e2;
    s_e[q][3] = w * e3 * e3;
    s_e[q][4] = w * (nu * (e1 * e1 + e2 * e2 + e3 * e3) + c * e0 * e0 + delta * e4 * e4);
    s_e[q][5] = w * (nu * (e1 * e1 + e2 * e2 + e3 * e3) + (e4 + c * e0) * e0);
    s_e[q][6] = fabs(e0);
  }
  __syncthreads();
  if(lane == 0) {
    double L[5] = {0, 0, 0, 0, 0};
    for(int q = 0; q < a.nq; ++q) {
      L[0] += s_e[q][0];
      if(s_e[q][6] > L[4]) L[4] = s_e[q][6];
      L[1] += s_e[q][1];
      L[1] += s_e[q][2];
      L[1] += s_e[q][3];
      L[2] += s_e[q][4];
      L[3] += s_e[q][5];
    }
    for(int k = 0; k < 5; ++k) cell_err[(size_t)cell * 5 + k] = L[k];
  }
}

}  // namespace

bool p1_simplex(const AsmArgs& a) {
  return !a.tcd && ((a.shape == (int)Shape::Triangle && a.nl == 3) || (a.shape == (int)Shape::Tetrahedron && a.nl == 4));
}

void launch_assemble_local(const AsmArgs& a, hipStream_t s) {
  PMCD_CHECK(a.nq <= kMaxQ, "too many quadrature points ", a.nq);
  if(a.n_cells == 0) return;
  if(p1_simplex(a)) {
    const int blocks = (a.n_cells + 255) / 256;
    if(a.dim == 3) k_assemble_p1<3><<<blocks, 256, 0, s>>>(a);
    else k_assemble_p1<2><<<blocks, 256, 0, s>>>(a);
    PMCD_HIP(hipGetLastError());
    return;
  }
  const int blocks = (a.n_cells + kCellsPerBlock - 1) / kCellsPerBlock;
  if(blocks == 0) return;
  switch(a.nl) {
    case 3: k_assemble_local<3><<<blocks, 256, 0, s>>>(a); break;
    case 4: k_assemble_local<4><<<blocks, 256, 0, s>>>(a); break;
    case 6: k_assemble_local<6><<<blocks, 256, 0, s>>>(a); break;
    case 8: k_assemble_local<8><<<blocks, 256, 0, s>>>(a); break;
    case 9: k_assemble_local<9><<<blocks, 256, 0, s>>>(a); break;
    case 10: k_assemble_local<10><<<blocks, 256, 0, s>>>(a); break;
    case 27: k_assemble_local<27><<<blocks, 256, 0, s>>>(a); break;
    default: PMCD_FAIL("unsupported local dimension ", a.nl);
  }
  PMCD_HIP(hipGetLastError());
}

void launch_gather_matrix(int n_rows, int n_active, int nl, const int* row_ptr, const int* col,
                          const int* contrib_ptr, const int* contrib, const uint8_t* slot,
                          const double* local_A, double* val, hipStream_t s, const int* sell_slot,
                          const long long* sell_off, double* sell_val) {
  if(n_rows == 0) return;
  k_gather_matrix<<<n_rows, 64, 0, s>>>(n_rows, n_active, nl, row_ptr, col, contrib_ptr, contrib,
                                        slot, local_A, val, sell_slot, sell_off, sell_val);
  PMCD_HIP(hipGetLastError());
}

void launch_gather_rhs(int n_rows, int n_active, const int* contrib_ptr, const int* contrib,
                       const double* local_rhs, const int* extra_ptr, const double* extra,
                       const double* dirichlet_values, double* rhs, hipStream_t s) {
  if(n_rows == 0) return;
  k_gather_rhs<<<(n_rows + 255) / 256, 256, 0, s>>>(n_rows, n_active, contrib_ptr, contrib,
                                                     local_rhs, extra_ptr, extra, dirichlet_values,
                                                     rhs);
  PMCD_HIP(hipGetLastError());
}

void launch_cell_errors(const AsmArgs& a, const int* dof, const double* u, double* cell_err,
                        hipStream_t s) {
  if(a.n_cells == 0) return;
  k_cell_errors<<<a.n_cells, 64, 0, s>>>(a, dof, u, cell_err);
  PMCD_HIP(hipGetLastError());
}

}  // namespace pmcd
