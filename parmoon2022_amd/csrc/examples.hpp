// PDE data of the convection-diffusion examples, evaluable on host and device.
//
// The reference passes host callbacks (CoeffFct2D/3D, BoundCondFunct,
// BoundValueFunct; include/General/Constants.h:36-54) per cell. A device
// kernel cannot call them, so the examples are enumerated with parameter
// structs and evaluated inside the kernels. Coefficients follow the
// reference layout (src/AssembleRoutines/ConvDiff.C:1797-1801,1858-1863):
// eps, b, c, f and |b|.
#pragma once
#include <cmath>

#ifndef __HIP_DEVICE_COMPILE__
#ifndef __host__
#define __host__
#define __device__
#endif
#endif

namespace pmcd {

enum ExampleKind : int {
  // include/Examples/CD_2D/SineLaplace.h (reference 2D example 0): u =
  // sin(pi x) sin(pi y), eps = nu, b = c = 0, Neumann on component 1 (x = 1)
  EX_SINE_LAPLACE_2D = 0,
  // u = sin(pi x) sin(pi y), eps, b = (b1, b2), c = 0, Dirichlet everywhere
  EX_SINE_2D = 1,
  // u = sin(pi x) sin(pi y) sin(pi z), eps, b, c = 0, Dirichlet everywhere
  EX_SINE_3D = 2,
  // include/Examples/CD_3D/test_p0.h (example -1): u = 1
  EX_TEST_P0_3D = 3,
  // include/Examples/CD_3D/test_p1.h (example -2): u = 1 + 2x + 3y + 4z
  EX_TEST_P1_3D = 4,
  // include/Examples/CD_3D/test_p2.h (example -3): u = x^2 + y z
  EX_TEST_P2_3D = 5,
  // eps, b, c given, f = f0, homogeneous Dirichlet; no known solution
  EX_CONST_RHS = 6,
  // time-dependent examples (the time is ExampleParams::t):
  // include/Examples/TCD_3D/linear_space_time.h (reference TCD3D example -2):
  // u = 1 + 2x + 3ty + 4tz, eps = 1, b = (1, 2, 3), c = 0, Dirichlet
  EX_TCD_LINEAR_3D = 7,
  // include/Examples/TCD_3D/quadratic_space_time.h (example -1):
  // u = t^2 (x^2 + y z), eps = 1, b = (4, 3, 2), c = 1, Dirichlet
  EX_TCD_QUAD_3D = 8,
  // include/Examples/TCD_2D/exp.h (reference TCD2D example 0): u =
  // exp(sin(2 pi t)) sin(2 pi x) sin(2 pi y), eps (1 / Reynolds number),
  // b = (1, -1), c = 1, homogeneous Dirichlet
  EX_TCD_EXP_2D = 9,
  // include/Examples/CD_2D/SharpBoundaryLayer.h (reference 2D example 3):
  // eps = 1e-3, b = (cos 10 deg, sin 10 deg), c = f = 0; Neumann (0) on
  // component 2 (y = 1), Dirichlet 1 on x = 0 above y = 0.5, else 0; no known
  // solution (errors are norms of the discrete solution)
  EX_SHARP_LAYER_2D = 10,
  // include/Examples/CD_3D/smooth_solution_3d.h (reference 3D example 3):
  // u = x y z (1-x) (1-y) (1-z), eps, b = (1, 1, 1), c = 1, Dirichlet 0
  EX_SMOOTH_3D = 11,
};

struct ExampleParams {
  int kind;
  int dim;
  double eps;
  double b[3];
  double c;
  double f0;
  double t;  // time of the time-dependent examples
  // functions of t alone, evaluated on the host by set_example_time (the
  // exp example's exp(sin(2 pi t)) and cos(2 pi t)): evaluating them per
  // quadrature point inside the assembly kernels lengthened every example's
  // assembly by 60%
  double tf[2];
};

inline void set_example_time(ExampleParams& ex, double t) {
  const double q = 2 * M_PI;
  ex.t = t;
  ex.tf[0] = std::exp(std::sin(q * t));
  ex.tf[1] = std::cos(q * t);
}

// u = x y z (1-x) (1-y) (1-z): value, gradient, Laplacian
__host__ __device__ inline void example_smooth3d(double x, double y, double z, double* e) {
  e[0] = x * y * z * (1 - x) * (1 - y) * (1 - z);
  e[1] = y * z * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - y) * (1 - z);
  e[2] = x * z * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - x) * (1 - z);
  e[3] = x * y * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - x) * (1 - y);
  e[4] = -2 * y * z * (1 - y) * (1 - z) - 2 * x * z * (1 - x) * (1 - z) - 2 * x * y * (1 - x) * (1 - y);
}

// whether eps, b and c of the example are constant in space (only f may
// vary): the local matrices are then combinations of reference tensors
__host__ __device__ inline bool example_constant_matrix_coeffs(int kind) {
  switch(kind) {
    case EX_SINE_LAPLACE_2D: case EX_SINE_2D: case EX_SINE_3D: case EX_CONST_RHS: case EX_TEST_P0_3D:
    case EX_TEST_P1_3D: case EX_TEST_P2_3D: case EX_TCD_LINEAR_3D: case EX_TCD_QUAD_3D: case EX_TCD_EXP_2D:
    case EX_SHARP_LAYER_2D: case EX_SMOOTH_3D:
      return true;
    default:
      return false;
  }
}

// coefficients at a point: out = {eps, b1, b2, b3, c, f}
__host__ __device__ inline void example_coeffs(const ExampleParams& ex, double x, double y,
                                               double z, double* out) {
  const double p = M_PI;
  switch(ex.kind) {
    case EX_SINE_LAPLACE_2D: {
      // SineLaplace.h:89-108: f = -eps * laplace(u)
      double lap = -2 * p * p * sin(p * x) * sin(p * y);
      out[0] = ex.eps; out[1] = 0; out[2] = 0; out[3] = 0; out[4] = 0;
      out[5] = -out[0] * lap;
      break;
    }
    case EX_SINE_2D: {
      double u = sin(p * x) * sin(p * y);
      double ux = p * cos(p * x) * sin(p * y), uy = p * sin(p * x) * cos(p * y);
      double lap = -2 * p * p * u;
      out[0] = ex.eps; out[1] = ex.b[0]; out[2] = ex.b[1]; out[3] = 0; out[4] = 0;
      out[5] = -out[0] * lap + out[1] * ux + out[2] * uy;
      break;
    }
    case EX_SINE_3D: {
      double sx = sin(p * x), sy = sin(p * y), sz = sin(p * z);
      double u = sx * sy * sz;
      double ux = p * cos(p * x) * sy * sz, uy = p * sx * cos(p * y) * sz,
             uz = p * sx * sy * cos(p * z);
      out[0] = ex.eps; out[1] = ex.b[0]; out[2] = ex.b[1]; out[3] = ex.b[2]; out[4] = 0;
      out[5] = 3 * p * p * ex.eps * u + ex.b[0] * ux + ex.b[1] * uy + ex.b[2] * uz;
      break;
    }
    case EX_TEST_P0_3D:
      out[0] = 1; out[1] = 1; out[2] = 1; out[3] = 1; out[4] = 1; out[5] = 1;
      break;
    case EX_TEST_P1_3D:
      out[0] = 1; out[1] = 1; out[2] = 2; out[3] = 3; out[4] = 0; out[5] = 1 * 2 + 2 * 3 + 3 * 4;
      break;
    case EX_TEST_P2_3D: {
      double eps = 1;
      out[0] = eps; out[1] = 4; out[2] = 3; out[3] = 2; out[4] = 1;
      out[5] = -eps * 2 + 4 * 2 * x + 3 * z + 2 * y + 1 * (x * x + y * z);
      break;
    }
    case EX_TCD_LINEAR_3D: {
      // linear_space_time.h BilinearCoeffs: f = u_t + b . grad u
      const double t = ex.t;
      out[0] = 1; out[1] = 1; out[2] = 2; out[3] = 3; out[4] = 0;
      out[5] = 3 * y + 4 * z + 2 + 6 * t + 12 * t;
      break;
    }
    case EX_TCD_QUAD_3D: {
      // quadratic_space_time.h BilinearCoeffs: f = u_t - eps lap u + b . grad u + c u
      const double t = ex.t, eps = 1;
      out[0] = eps; out[1] = 4; out[2] = 3; out[3] = 2; out[4] = 1;
      out[5] = 2 * t * (x * x + y * z) - eps * 2 * t * t + 4 * 2 * x * t * t + 3 * z * t * t + 2 * y * t * t +
               1 * t * t * (x * x + y * z);
      break;
    }
    case EX_SHARP_LAYER_2D:
      out[0] = 0.001; out[1] = cos(p / 18); out[2] = sin(p / 18); out[3] = 0; out[4] = 0; out[5] = 0;
      break;
    case EX_SMOOTH_3D: {
      // smooth_solution_3d.h BilinearCoeffs: f = -eps lap u + b . grad u + c u
      double e[5];
      example_smooth3d(x, y, z, e);
      out[0] = ex.eps; out[1] = 1; out[2] = 1; out[3] = 1; out[4] = 1;
      out[5] = -out[0] * e[4];
      out[5] += out[1] * e[1] + out[2] * e[2] + out[3] * e[3];
      out[5] += out[4] * e[0];
      break;
    }
    case EX_TCD_EXP_2D: {
      // exp.h BilinearCoeffs: f = u_t - eps lap u + b . grad u + c u
      const double q = 2 * p, et = ex.tf[0], sx = sin(q * x), sy = sin(q * y), cx = cos(q * x),
                   cy = cos(q * y);
      const double s = sx * sy;
      out[0] = ex.eps; out[1] = 1; out[2] = -1; out[3] = 0; out[4] = 1;
      out[5] = et * q * ex.tf[1] * s - ex.eps * et * q * q * (-s - s) + 1 * et * q * cx * sy +
               (-1) * et * q * sx * cy + 1 * et * s;
      break;
    }
    default:  // EX_CONST_RHS
      out[0] = ex.eps; out[1] = ex.b[0]; out[2] = ex.b[1]; out[3] = ex.b[2];
      out[4] = ex.c; out[5] = ex.f0;
      break;
  }
}

// exact solution: out = {u, u_x, u_y, u_z, laplace u}; returns false if unknown
__host__ __device__ inline bool example_exact(const ExampleParams& ex, double x, double y,
                                              double z, double* out) {
  const double p = M_PI;
  switch(ex.kind) {
    case EX_SINE_LAPLACE_2D:
    case EX_SINE_2D:
      out[0] = sin(p * x) * sin(p * y);
      out[1] = p * cos(p * x) * sin(p * y);
      out[2] = p * sin(p * x) * cos(p * y);
      out[3] = 0;
      out[4] = -2 * p * p * sin(p * x) * sin(p * y);
      return true;
    case EX_SINE_3D: {
      double sx = sin(p * x), sy = sin(p * y), sz = sin(p * z);
      out[0] = sx * sy * sz;
      out[1] = p * cos(p * x) * sy * sz;
      out[2] = p * sx * cos(p * y) * sz;
      out[3] = p * sx * sy * cos(p * z);
      out[4] = -3 * p * p * sx * sy * sz;
      return true;
    }
    case EX_TEST_P0_3D:
      out[0] = 1; out[1] = 0; out[2] = 0; out[3] = 0; out[4] = 0;
      return true;
    case EX_TEST_P1_3D:
      out[0] = 1 + 2 * x + 3 * y + 4 * z; out[1] = 2; out[2] = 3; out[3] = 4; out[4] = 0;
      return true;
    case EX_TEST_P2_3D:
      out[0] = x * x + y * z; out[1] = 2 * x; out[2] = z; out[3] = y; out[4] = 2;
      return true;
    case EX_TCD_LINEAR_3D: {
      const double t = ex.t;
      out[0] = 1 + 2 * x + 3 * t * y + 4 * t * z; out[1] = 2; out[2] = 3 * t; out[3] = 4 * t; out[4] = 0;
      return true;
    }
    case EX_TCD_QUAD_3D: {
      const double t = ex.t;
      out[0] = t * t * (x * x + y * z); out[1] = 2 * x * t * t; out[2] = z * t * t; out[3] = y * t * t;
      out[4] = 2 * t * t;
      return true;
    }
    case EX_SMOOTH_3D:
      example_smooth3d(x, y, z, out);
      return true;
    case EX_TCD_EXP_2D: {
      const double q = 2 * p, et = ex.tf[0];
      out[0] = et * sin(q * x) * sin(q * y);
      out[1] = et * q * cos(q * x) * sin(q * y);
      out[2] = et * q * sin(q * x) * cos(q * y);
      out[3] = 0;
      out[4] = -2 * q * q * out[0];
      return true;
    }
    default:
      out[0] = out[1] = out[2] = out[3] = out[4] = 0;
      return false;
  }
}

// boundary condition kind (0 Dirichlet, 1 Neumann) of component `comp`
__host__ __device__ inline int example_bc(const ExampleParams& ex, int comp) {
  if(ex.kind == EX_SINE_LAPLACE_2D && comp == 1) return 1;
  if(ex.kind == EX_SHARP_LAYER_2D && comp == 2) return 1;
  return 0;
}

// boundary value at a boundary point of component `comp` (SineLaplace.h:67-84:
// Dirichlet value u, Neumann value eps * n . grad u)
__host__ __device__ inline double example_bvalue(const ExampleParams& ex, int comp, double x,
                                                 double y, double z) {
  double e[5];
  if(ex.kind == EX_TCD_EXP_2D || ex.kind == EX_SMOOTH_3D) return 0.0;  // BoundValue = 0
  // SharpBoundaryLayer.h BoundValue: 1 on x = 0 for boundary parameter < 0.5
  // (y > 0.5), 0 elsewhere and on the Neumann side
  if(ex.kind == EX_SHARP_LAYER_2D) return (comp != 2 && fabs(x) < 1e-12 && y > 0.5) ? 1.0 : 0.0;
  if(!example_exact(ex, x, y, z, e)) return 0.0;
  if(ex.kind == EX_SINE_LAPLACE_2D && comp == 1) {
    // outer normal of x = 1 is (1, 0)
    return ex.eps * (1.0 * e[1] + 0.0 * e[2]);
  }
  return e[0];
}

}  // namespace pmcd
