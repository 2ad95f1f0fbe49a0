// Conforming Lagrange elements of order 1 and 2 on triangles, quadrilaterals,
// tetrahedra and hexahedra: local DOF layout, reference basis functions and
// their derivatives up to second order.
//
// Reference layouts: include/FE2D/FE_C_T_P1_2D.h, FE_C_T_P2_2D.h,
// FE_C_Q_Q1_2D.h, FE_C_Q_Q2_2D.h, include/FE3D/FE_C_T_P1_3D.h, FE_C_T_P2_3D.h,
// FE_C_H_Q1_3D.h, FE_C_H_Q2_3D.h; basis functions in the matching BF_*.h.
#pragma once
#include "common.hpp"

namespace pmcd {

constexpr int kMaxLocalDof = 27;
// Derivative components, in the reference's multi-index order
// (D00,D10,D01,D20,D11,D02 in 2D; D000,D100,D010,D001,D200,D110,D101,D020,D011,D002 in 3D).
constexpr int kNDeriv2D = 6;
constexpr int kNDeriv3D = 10;

struct ElementInfo {
  Element type;
  Shape shape;
  int order;       // polynomial degree used for the quadrature rule
  int n_dof;
  // reference coordinates of the nodal point of each local DOF
  std::vector<std::array<double, 3>> nodes;
  // for each local DOF, the cell vertices spanning the smallest mesh entity
  // (vertex, edge, face or cell) whose barycentre is the nodal point
  std::vector<std::vector<int>> entity_vertices;
  // local DOFs on the closure of each joint (edge in 2D, face in 3D)
  std::vector<std::vector<int>> joint_dofs;
};

const ElementInfo& element_info(Element e);

// Evaluate all derivative components at reference point p.
// out[c * n_dof + i] = D_c phi_i(p), c over the multi-index list above.
void eval_basis(Element e, const double* p, double* out);

}  // namespace pmcd
