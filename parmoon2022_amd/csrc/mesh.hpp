// Conforming meshes of triangles, quadrilaterals, tetrahedra or hexahedra,
// generated from the reference's default coarse meshes by uniform regular
// refinement. Cells of a refined level are stored in the order of a
// depth-first walk over the refinement tree (src/Geometry/It_EQ.C:16-51,
// src/Geometry/Domain.C:1961-2000), i.e. the children of each coarse cell,
// in child order, in the coarse cell order.
#pragma once
#include <functional>
#include "common.hpp"

namespace pmcd {

struct Mesh {
  Shape shape = Shape::Triangle;
  int n_cells = 0;
  int nv = 0;                      // vertices per cell
  int nj = 0;                      // joints (edges/faces) per cell
  std::vector<double> coords;      // 3 per vertex
  std::vector<int> cell_vertices;  // nv per cell, global vertex ids
  std::vector<int> joint_comp;     // nj per cell: boundary component id, -1 if interior
  std::vector<int> parent;         // parent cell in the coarser level, -1 for a root
  std::vector<int> child_index;    // index of the cell among its parent's children

  int dim() const { return shape_dim(shape); }
  int n_vertices() const { return (int)(coords.size() / 3); }
  const int* cell(int c) const { return &cell_vertices[(size_t)c * nv]; }
};

// Default coarse meshes (src/Geometry/Tests.C:95-188 TwoTriangles(2),
// :473-517 UnitSquare, :1893-1935 unit-cube hexahedron, :1937-1995 unit-cube
// tetrahedra). Boundary component ids follow the default boundary files
// (Tests.C:1704-1736, :1799-1845).
enum class DefaultGeometry : int {
  UnitSquare = 0, TwoTriangles = 1, TwoTriangles2 = 2, CubeHexa = 3, CubeTetra = 4
};
Mesh default_mesh(DefaultGeometry g);

// Structured N x N x N cube meshes as a mesh file would list them
// (lexicographic cells, x fastest): hexahedra, or each cube split into the
// six tetrahedra of the default unit-cube tetrahedral mesh.
Mesh structured_cube(int n, bool tetra);

// 3D mesh in the medit .mesh format, read as TDomain::InitFromMesh without a
// .PRM file reads it (src/Geometry/Mesh.C:48-260, Domain.C:2912-3330):
// vertices and cells (Tetrahedra or Hexahedra) in file order, each cell's
// vertices in file order (1-based ids). Every boundary face is its own
// boundary component, numbered in the order of the file's boundary faces
// (Triangles / Quadrilaterals); boundary faces the file does not list follow
// in cell and local face order. Cells with 2D shapes need a .PRM boundary
// description and are rejected.
Mesh read_mesh_file(const std::string& path);
// 2D mesh in the .GEO format as TDomain::ReadGeo reads it
// (src/Geometry/ReadGeo.C:43-140): counts, DCORVG, KVERT, KNPR; root cells
// and their vertices in file order. Boundary vertices (KNPR > 0) give the
// boundary parameter, evaluated on the unit square's boundary
// (Default_UnitSquare); other boundary descriptions are rejected.
Mesh read_geo_file(const std::string& path);
// Writes a 3D mesh in the same format: vertices, the boundary faces (one per
// boundary joint, in cell and local face order, marker = component + 1) and
// the cells.
void write_mesh_file(const Mesh& m, const std::string& path);

// One step of regular refinement of every cell (RefTriRegDesc.C,
// RefQuadRegDesc.C, RefTetraRegDesc.C, RefHexaRegDesc.C child tables).
Mesh refine_regular(const Mesh& coarse);

// A hanging vertex of a mesh refined by indicator: the midpoint of the edge
// (a, b) of a leaf cell whose neighbour across that edge is refined.
struct HangingVertex {
  int vertex, a, b;
};

// indicator value at a point (x, y, z)
using IndicatorFn = std::function<double(const double* x)>;

// `steps` times TDomain::RefineByIndicator (src/Geometry/Domain.C:1094-1207,
// RefinementStrategy.C:260-289): a cell whose corner values of the indicator
// lie on both sides of zero (a value within 1e-10 of zero counts for both) is
// refined regularly, then every leaf with a neighbour refined two levels
// deeper across one of its edges is refined until the mesh is 1-irregular
// (Domain.C:1818-1835, GridCell.C:348-383, 462-491). Leaves stay in
// depth-first tree order (a refined cell is replaced by its children, in
// child order). 2D meshes only. Without `conforming`, `hanging` receives the
// hanging vertices in the order their coarse edges are met over (cell, edge).
// With `conforming` (conforming_closure, triangles; MakeConfClosure,
// Domain.C:1219-1360, GridCell.C:525-575) the closure bisections of the last
// step are removed before each step (a bisected cell is marked when one of
// its halves is), triangles with two or three refined edges are refined
// regularly until none is left, and a triangle with one refined edge is
// bisected on it: the mesh is conforming and `hanging` stays empty.
Mesh refine_by_indicator(const Mesh& base, int steps, const IndicatorFn& f, std::vector<HangingVertex>& hanging,
                         bool conforming = false);

// Cell diameter as used for h_K with CELL_MEASURE=0 (Triangle.C:25-40,
// Quadrangle.C:26-37, Tetrahedron.C:64-89, longest space diagonal for
// hexahedra).
double cell_diameter(const Mesh& m, int c);

}  // namespace pmcd
