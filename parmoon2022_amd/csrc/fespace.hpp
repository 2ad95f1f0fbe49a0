// Continuous Lagrange FE space on a mesh: local-to-global DOF map with the
// reference's global numbering.
//
// Reference numbering (src/FE/FESpace3D.C:186-804 / FESpace2D.C ConstructSpace,
// src/FE/FEMapper.C:456-552 map_single_dof, src/FE/FESpace.C:1863-2099
// compute_various_numbers): local DOFs of face-neighbouring cells are merged
// into classes whose representative is the class member with the smallest
// (cell, local dof) position; classes are then numbered in order of that first
// appearance, inner DOFs first, then the non-Dirichlet boundary kinds in the
// order of the BoundCond enumeration, Dirichlet DOFs last
// (include/FE/FESpace.h:71-92). A class lying on boundary joints of several
// kinds takes the kind that comes first in BoundCond.
//
// This build identifies the classes directly by the mesh entity carrying each
// nodal point (vertex, edge, face, cell interior) instead of by union-find
// over faces; for conforming meshes the two give the same classes and the
// same first appearances, hence the same numbers.
#pragma once
#include "common.hpp"
#include "mesh.hpp"
#include "elements.hpp"
#include <functional>

namespace pmcd {

// boundary condition of joint with component `comp` and centroid x
using BoundCondFn = std::function<BoundCond(int comp, const double* x)>;

struct FESpace {
  Element element;
  int n_local = 0;
  int n_dof = 0;
  int n_active = 0;      // inner + non-Dirichlet boundary + hanging DOFs
  int n_dirichlet = 0;
  int n_inner = 0;
  std::vector<int> dof;  // n_cells * n_local, global DOF numbers
  // hanging DOFs n_active - n_hanging .. n_active - 1: hanging DOF h couples
  // with hang_n[h] DOFs hang_cpl[kMaxCoupling * h + k], weights
  // hang_w[kMaxCoupling * h + k]; hang_order lists the hanging DOFs (0-based)
  // so that each comes before the hanging DOFs it couples with
  // (sort_hanging_nodes, src/FE/FESpace.C:2101-2160)
  static constexpr int kMaxCoupling = 3;
  int n_hanging = 0;
  std::vector<int> hang_n;
  std::vector<int> hang_cpl;
  std::vector<double> hang_w;
  std::vector<int> hang_order;
  const int* cell_dof(int c) const { return &dof[(size_t)c * n_local]; }
  int n_active_non_hanging() const { return n_active - n_hanging; }
};

// `hanging` (2D, order 1 or 2): hanging vertices of the mesh. Their edges'
// hanging DOFs follow the non-Dirichlet boundary DOFs, in the list's order
// (FESpace.C:1863-2099 numbers them between the active and the Dirichlet
// DOFs), as the one-regular mappers make them (FE2DMapper1Reg.C:43-78):
// order 1: the vertex m of the coarse edge (a, b) hangs, coupled to a and b
// with weights 1/2 (HN_C_P1_2D_0, HNDesc.C:9-11); order 2: the coarse edge's
// middle DOF is the fine vertex m, and the middle DOFs of the fine edges
// (a, m), (m, b) hang, coupled to (a, m, b) with the quadratic Lagrange
// values at 1/4 and 3/4 of the edge, {3/8, 3/4, -1/8} and {-1/8, 3/4, 3/8}
// (HN_C_P2_2D_1 / _0, HNDesc.C:12-17).
FESpace build_fespace(const Mesh& mesh, int order, const BoundCondFn& bc,
                      const std::vector<HangingVertex>* hanging = nullptr);

// hang_order from hang_n / hang_cpl: sort_hanging_nodes (FESpace.C:2101-2160),
// depth-first post-order over the coupling DOFs that hang themselves,
// reversed
void sort_hanging(FESpace& sp);

}  // namespace pmcd
