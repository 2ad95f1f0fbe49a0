// Shared host-side definitions for the convection-diffusion assemble-and-solve
// library: element/shape enumerations, error type, small helpers.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include <array>
#include <sstream>

namespace pmcd {

// Errors cross the C boundary as status codes; inside the library they are
// thrown as pmcd::Error (a std::runtime_error, like the reference's ErrThrow,
// include/General/MooNMD_Io.h:435-447).
struct Error : public std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

template <class... A>
[[noreturn]] inline void fail(const char* file, int line, A&&... a) {
  std::ostringstream os;
  os << file << ":" << line << ": ";
  (os << ... << a);
  throw Error(os.str());
}
#define PMCD_FAIL(...) ::pmcd::fail(__FILE__, __LINE__, __VA_ARGS__)
#define PMCD_CHECK(c, ...) do { if(!(c)) PMCD_FAIL(__VA_ARGS__); } while(0)

// Cell shapes handled on this path (reference: include/Geometry/Enumerations_geometry.h).
enum class Shape : int { Triangle = 0, Quadrangle = 1, Tetrahedron = 2, Hexahedron = 3 };

inline int shape_dim(Shape s) { return (s == Shape::Triangle || s == Shape::Quadrangle) ? 2 : 3; }
inline int shape_n_vertices(Shape s) {
  switch(s) { case Shape::Triangle: return 3; case Shape::Quadrangle: return 4;
              case Shape::Tetrahedron: return 4; default: return 8; }
}
inline int shape_n_faces(Shape s) {  // joints: edges in 2D, faces in 3D
  switch(s) { case Shape::Triangle: return 3; case Shape::Quadrangle: return 4;
              case Shape::Tetrahedron: return 4; default: return 6; }
}
inline bool shape_is_simplex(Shape s) { return s == Shape::Triangle || s == Shape::Tetrahedron; }

// Conforming Lagrange elements on this path (reference FE types C_P1_2D_T_A,
// C_P2_2D_T_A, C_Q1_2D_Q_A, C_Q2_2D_Q_A, C_P1_3D_T_A, C_P2_3D_T_A,
// C_Q1_3D_H_A, C_Q2_3D_H_A; src/FE/FESpace.C:240-255).
enum class Element : int {
  P1Tri = 0, P2Tri = 1, Q1Quad = 2, Q2Quad = 3,
  P1Tet = 4, P2Tet = 5, Q1Hex = 6, Q2Hex = 7
};

Element element_for(Shape s, int order);
int element_n_dof(Element e);
int element_order(Element e);
Shape element_shape(Element e);

// Boundary condition kinds, same order as the reference enum BoundCond
// (include/General/Constants.h:17-18); only the first two occur on this path.
enum BoundCond : int { DIRICHLET = 0, NEUMANN = 1, ROBIN = 2 };

// Local joint vertex tables of each shape (reference ShapeDesc face/edge
// vertex tables). For quadrangle/hexahedron faces the vertex order is the
// order used to parametrise face points.
const std::vector<std::vector<int>>& shape_faces(Shape s);

}  // namespace pmcd
