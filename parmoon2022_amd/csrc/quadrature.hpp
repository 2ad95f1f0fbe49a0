// Quadrature rules on the reference cells used by the convection-diffusion
// path, and the degree -> rule map of the reference.
//
// Reference: src/QuadFormulas/QuadFormula.C (rule data),
//            src/QuadFormulas/QuadratureFormulaDatabase.C:22-151 (degree map),
//            src/FE/FEDatabase.C:1554-1556 (degree = N_terms * p).
#pragma once
#include "common.hpp"

namespace pmcd {

enum class QuadRule : int {
  Gauss1Line, Gauss2Line, Gauss3Line, Gauss4Line,
  Gauss3Tria, Degree9Tria,
  Gauss2Quad, Gauss3Quad, Gauss4Quad,
  P2Tetra, P4Tetra, P5Tetra,
  Gauss2Hexa, Gauss3Hexa, Gauss4Hexa
};

struct QuadFormula {
  QuadRule type;
  int dim;                      // 1, 2 or 3
  std::vector<double> w;        // weights on the reference cell
  std::vector<double> xi, eta, zeta;
  int n() const { return (int)w.size(); }
};

const QuadFormula& quad_formula(QuadRule r);
const char* quad_rule_name(QuadRule r);

// Rule used for cell integrals of an element of polynomial degree p.
// `linear_mode` mirrors INTERNAL_PROBLEM_LINEAR (N_terms = 2 instead of 3).
QuadRule cell_rule_for(Shape s, int poly_degree, bool linear_mode);
// Line rule used for Neumann edge integrals (degree 2p, Assemble2D.C:3232-3234).
QuadRule line_rule_for_degree(int degree);

}  // namespace pmcd
