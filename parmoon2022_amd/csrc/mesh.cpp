#include "mesh.hpp"
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

namespace pmcd {

namespace {

// Refinement table of one shape: the parent-local vertices spanning each
// refined vertex (its barycentre is the new vertex) and the refined-vertex
// lists of the children, in child order.
struct RefTable {
  std::vector<std::vector<int>> span;
  std::vector<std::vector<int>> child;
};

const RefTable& ref_table(Shape s) {
  static const RefTable tri = {
      {{0}, {1}, {2}, {0, 1}, {1, 2}, {2, 0}},
      {{0, 3, 5}, {1, 4, 3}, {2, 5, 4}, {3, 4, 5}}};
  static const RefTable quad = {
      {{0}, {1}, {2}, {3}, {0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 1, 2, 3}},
      {{0, 4, 8, 7}, {1, 5, 8, 4}, {2, 6, 8, 5}, {3, 7, 8, 6}}};
  static const RefTable tet = {
      {{0}, {1}, {2}, {3}, {0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}},
      {{0, 4, 6, 7}, {4, 1, 5, 8}, {6, 5, 2, 9}, {7, 8, 9, 3},
       {4, 6, 7, 8}, {5, 6, 4, 8}, {6, 7, 8, 9}, {8, 5, 6, 9}}};
  static const RefTable hex = {
      {{0}, {1}, {2}, {3}, {4}, {5}, {6}, {7},
       {0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 1, 2, 3},
       {1, 5}, {5, 4}, {4, 0}, {0, 1, 5, 4},
       {2, 6}, {6, 5}, {1, 2, 6, 5},
       {3, 7}, {7, 6}, {2, 3, 7, 6},
       {4, 7}, {3, 0, 4, 7}, {4, 5, 6, 7},
       {0, 1, 2, 3, 4, 5, 6, 7}},
      {{0, 8, 12, 11, 15, 16, 26, 24}, {1, 9, 12, 8, 13, 19, 26, 16},
       {2, 10, 12, 9, 17, 22, 26, 19}, {3, 11, 12, 10, 20, 24, 26, 22},
       {4, 23, 25, 14, 15, 24, 26, 16}, {5, 14, 25, 18, 13, 16, 26, 19},
       {6, 18, 25, 21, 17, 19, 26, 22}, {7, 21, 25, 23, 20, 22, 26, 24}}};
  switch(s) {
    case Shape::Triangle: return tri;
    case Shape::Quadrangle: return quad;
    case Shape::Tetrahedron: return tet;
    default: return hex;
  }
}

struct KeyHash {
  size_t operator()(const std::array<int, 4>& k) const {
    uint64_t h = 1469598103934665603ull;
    for(int v : k) { h ^= (uint64_t)(uint32_t)v; h *= 1099511628211ull; }
    return (size_t)h;
  }
};

// boundary component of a joint of the unit square / unit cube from its
// centroid (Default_UnitSquare: 0 y=0, 1 x=1, 2 y=1, 3 x=0;
// Default_UnitCube: 0 z=0, 1 x=0, 2 y=0, 3 z=1, 4 x=1, 5 y=1)
int unit_box_component(int dim, const double* c) {
  const double t = 1e-12;
  if(dim == 2) {
    if(std::fabs(c[1]) < t) return 0;
    if(std::fabs(c[0] - 1) < t) return 1;
    if(std::fabs(c[1] - 1) < t) return 2;
    if(std::fabs(c[0]) < t) return 3;
  } else {
    if(std::fabs(c[2]) < t) return 0;
    if(std::fabs(c[0]) < t) return 1;
    if(std::fabs(c[1]) < t) return 2;
    if(std::fabs(c[2] - 1) < t) return 3;
    if(std::fabs(c[0] - 1) < t) return 4;
    if(std::fabs(c[1] - 1) < t) return 5;
  }
  return -1;
}

// mark joints on the boundary of the unit box
void set_box_boundary(Mesh& m) {
  const auto& faces = shape_faces(m.shape);
  m.nj = (int)faces.size();
  m.joint_comp.assign((size_t)m.n_cells * m.nj, -1);
  for(int c = 0; c < m.n_cells; ++c)
    for(int j = 0; j < m.nj; ++j) {
      double cen[3] = {0, 0, 0};
      for(int lv : faces[j])
        for(int k = 0; k < 3; ++k) cen[k] += m.coords[3 * m.cell(c)[lv] + k] / faces[j].size();
      // a joint is on the boundary only if all its vertices share the plane
      int comp = unit_box_component(m.dim(), cen);
      if(comp >= 0) {
        bool all = true;
        for(int lv : faces[j]) {
          const double* x = &m.coords[3 * m.cell(c)[lv]];
          if(unit_box_component(m.dim(), x) < 0) all = false;
        }
        if(!all) comp = -1;
      }
      m.joint_comp[(size_t)c * m.nj + j] = comp;
    }
}

}  // namespace

Mesh default_mesh(DefaultGeometry g) {
  Mesh m;
  auto add_v = [&](double x, double y, double z) {
    m.coords.push_back(x); m.coords.push_back(y); m.coords.push_back(z);
  };
  switch(g) {
    case DefaultGeometry::UnitSquare:
      m.shape = Shape::Quadrangle; m.nv = 4;
      add_v(0, 0, 0); add_v(1, 0, 0); add_v(1, 1, 0); add_v(0, 1, 0);
      m.cell_vertices = {0, 1, 2, 3};
      break;
    case DefaultGeometry::TwoTriangles:
      m.shape = Shape::Triangle; m.nv = 3;
      add_v(0, 0, 0); add_v(1, 0, 0); add_v(1, 1, 0); add_v(0, 1, 0);
      m.cell_vertices = {0, 1, 2, 0, 2, 3};
      break;
    case DefaultGeometry::TwoTriangles2:
      m.shape = Shape::Triangle; m.nv = 3;
      add_v(0, 0, 0); add_v(1, 0, 0); add_v(1, 1, 0); add_v(0, 1, 0);
      m.cell_vertices = {0, 1, 3, 1, 2, 3};
      break;
    case DefaultGeometry::CubeHexa:
      m.shape = Shape::Hexahedron; m.nv = 8;
      add_v(0, 0, 0); add_v(1, 0, 0); add_v(1, 1, 0); add_v(0, 1, 0);
      add_v(0, 0, 1); add_v(1, 0, 1); add_v(1, 1, 1); add_v(0, 1, 1);
      m.cell_vertices = {0, 1, 2, 3, 4, 5, 6, 7};
      break;
    case DefaultGeometry::CubeTetra:
      m.shape = Shape::Tetrahedron; m.nv = 4;
      add_v(0, 0, 0); add_v(1, 0, 0); add_v(1, 1, 0); add_v(0, 1, 0);
      add_v(0, 0, 1); add_v(1, 0, 1); add_v(1, 1, 1); add_v(0, 1, 1);
      m.cell_vertices = {0, 1, 2, 6, 0, 5, 1, 6, 0, 2, 3, 6,
                         0, 3, 7, 6, 0, 4, 5, 6, 0, 7, 4, 6};
      break;
  }
  m.n_cells = (int)(m.cell_vertices.size() / m.nv);
  m.parent.assign(m.n_cells, -1);
  m.child_index.assign(m.n_cells, 0);
  set_box_boundary(m);
  return m;
}

Mesh structured_cube(int n, bool tetra) {
  PMCD_CHECK(n >= 1, "structured_cube needs n >= 1");
  Mesh m;
  m.shape = tetra ? Shape::Tetrahedron : Shape::Hexahedron;
  m.nv = tetra ? 4 : 8;
  const int n1 = n + 1;
  m.coords.resize((size_t)3 * n1 * n1 * n1);
  for(int k = 0; k < n1; ++k)
    for(int j = 0; j < n1; ++j)
      for(int i = 0; i < n1; ++i) {
        size_t v = (size_t)i + (size_t)n1 * (j + (size_t)n1 * k);
        m.coords[3 * v] = (double)i / n;
        m.coords[3 * v + 1] = (double)j / n;
        m.coords[3 * v + 2] = (double)k / n;
      }
  // local cube corners in the default unit-cube vertex order
  const int co[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                        {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const int tets[6][4] = {{0, 1, 2, 6}, {0, 5, 1, 6}, {0, 2, 3, 6},
                          {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 7, 4, 6}};
  for(int k = 0; k < n; ++k)
    for(int j = 0; j < n; ++j)
      for(int i = 0; i < n; ++i) {
        int g[8];
        for(int a = 0; a < 8; ++a)
          g[a] = (i + co[a][0]) + n1 * ((j + co[a][1]) + n1 * (k + co[a][2]));
        if(tetra) {
          for(auto& t : tets)
            for(int a = 0; a < 4; ++a) m.cell_vertices.push_back(g[t[a]]);
        } else {
          for(int a = 0; a < 8; ++a) m.cell_vertices.push_back(g[a]);
        }
      }
  m.n_cells = (int)(m.cell_vertices.size() / m.nv);
  m.parent.assign(m.n_cells, -1);
  m.child_index.assign(m.n_cells, 0);
  set_box_boundary(m);
  return m;
}

Mesh refine_regular(const Mesh& cm) {
  const RefTable& rt = ref_table(cm.shape);
  const auto& faces = shape_faces(cm.shape);
  Mesh fm;
  fm.shape = cm.shape;
  fm.nv = cm.nv;
  fm.nj = cm.nj;
  fm.coords = cm.coords;  // old vertices keep their ids
  const int n_children = (int)rt.child.size();
  const int n_ref = (int)rt.span.size();
  fm.n_cells = cm.n_cells * n_children;
  fm.cell_vertices.resize((size_t)fm.n_cells * fm.nv);
  fm.joint_comp.assign((size_t)fm.n_cells * fm.nj, -1);
  fm.parent.resize(fm.n_cells);
  fm.child_index.resize(fm.n_cells);

  // refined vertices shared between cells are keyed by the sorted global ids
  // of the vertices spanning them; cell-interior vertices are never shared
  std::unordered_map<std::array<int, 4>, int, KeyHash> shared;
  shared.reserve((size_t)cm.n_cells * 8);

  // which refined vertices lie on each parent joint
  std::vector<std::vector<bool>> on_joint(faces.size(), std::vector<bool>(n_ref, false));
  for(size_t j = 0; j < faces.size(); ++j)
    for(int r = 0; r < n_ref; ++r) {
      bool in = true;
      for(int v : rt.span[r])
        if(std::find(faces[j].begin(), faces[j].end(), v) == faces[j].end()) in = false;
      on_joint[j][r] = in;
    }

  std::vector<int> gid(n_ref);
  for(int c = 0; c < cm.n_cells; ++c) {
    const int* cv = cm.cell(c);
    for(int r = 0; r < n_ref; ++r) {
      const auto& sp = rt.span[r];
      if(sp.size() == 1) { gid[r] = cv[sp[0]]; continue; }
      double x[3] = {0, 0, 0};
      for(int v : sp)
        for(int k = 0; k < 3; ++k) x[k] += cm.coords[3 * cv[v] + k];
      for(int k = 0; k < 3; ++k) x[k] *= 1.0 / sp.size();
      bool interior = (int)sp.size() == cm.nv;
      int id = -1;
      if(!interior) {
        std::array<int, 4> key = {-1, -1, -1, -1};
        std::vector<int> g;
        for(int v : sp) g.push_back(cv[v]);
        std::sort(g.begin(), g.end());
        for(size_t a = 0; a < g.size(); ++a) key[a] = g[a];
        auto it = shared.find(key);
        if(it != shared.end()) id = it->second;
        else {
          id = (int)(fm.coords.size() / 3);
          fm.coords.push_back(x[0]); fm.coords.push_back(x[1]); fm.coords.push_back(x[2]);
          shared.emplace(key, id);
        }
      } else {
        id = (int)(fm.coords.size() / 3);
        fm.coords.push_back(x[0]); fm.coords.push_back(x[1]); fm.coords.push_back(x[2]);
      }
      gid[r] = id;
    }
    for(int ch = 0; ch < n_children; ++ch) {
      int fc = c * n_children + ch;
      for(int a = 0; a < fm.nv; ++a) fm.cell_vertices[(size_t)fc * fm.nv + a] = gid[rt.child[ch][a]];
      fm.parent[fc] = c;
      fm.child_index[fc] = ch;
      // a child joint inherits the component of the parent joint containing it
      for(int jj = 0; jj < fm.nj; ++jj) {
        for(size_t j = 0; j < faces.size(); ++j) {
          int comp = cm.joint_comp[(size_t)c * cm.nj + j];
          if(comp < 0) continue;
          bool in = true;
          for(int lv : faces[jj])
            if(!on_joint[j][rt.child[ch][lv]]) in = false;
          if(in) { fm.joint_comp[(size_t)fc * fm.nj + jj] = comp; break; }
        }
      }
    }
  }
  return fm;
}

namespace {

// edge midpoints created so far, keyed by the edge's sorted end vertices
using EdgeMids = std::unordered_map<uint64_t, int>;
uint64_t edge_key(int u, int v) {
  if(u > v) std::swap(u, v);
  return ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
}
int edge_mid(const EdgeMids& mids, int u, int v) {
  auto it = mids.find(edge_key(u, v));
  return it == mids.end() ? -1 : it->second;
}

// one regular refinement of the marked cells; unmarked cells stay, children
// replace their parent in place
Mesh refine_marked(const Mesh& cm, const std::vector<char>& mark, EdgeMids& mids) {
  const RefTable& rt = ref_table(cm.shape);
  const auto& faces = shape_faces(cm.shape);
  const int n_children = (int)rt.child.size();
  const int n_ref = (int)rt.span.size();
  Mesh fm;
  fm.shape = cm.shape;
  fm.nv = cm.nv;
  fm.nj = cm.nj;
  fm.coords = cm.coords;
  std::vector<std::vector<bool>> on_joint(faces.size(), std::vector<bool>(n_ref, false));
  for(size_t j = 0; j < faces.size(); ++j)
    for(int r = 0; r < n_ref; ++r) {
      bool in = true;
      for(int v : rt.span[r])
        if(std::find(faces[j].begin(), faces[j].end(), v) == faces[j].end()) in = false;
      on_joint[j][r] = in;
    }
  std::vector<int> gid(n_ref);
  for(int c = 0; c < cm.n_cells; ++c) {
    const int* cv = cm.cell(c);
    if(!mark[c]) {
      fm.cell_vertices.insert(fm.cell_vertices.end(), cv, cv + cm.nv);
      fm.joint_comp.insert(fm.joint_comp.end(), cm.joint_comp.begin() + (size_t)c * cm.nj,
                           cm.joint_comp.begin() + (size_t)(c + 1) * cm.nj);
      fm.parent.push_back(c);
      fm.child_index.push_back(0);
      continue;
    }
    for(int r = 0; r < n_ref; ++r) {
      const auto& sp = rt.span[r];
      if(sp.size() == 1) { gid[r] = cv[sp[0]]; continue; }
      double x[3] = {0, 0, 0};
      for(int v : sp)
        for(int k = 0; k < 3; ++k) x[k] += cm.coords[3 * cv[v] + k];
      for(int k = 0; k < 3; ++k) x[k] *= 1.0 / sp.size();
      int id = -1;
      if(sp.size() == 2) {
        auto it = mids.find(edge_key(cv[sp[0]], cv[sp[1]]));
        if(it != mids.end()) id = it->second;
      }
      if(id < 0) {
        id = (int)(fm.coords.size() / 3);
        fm.coords.push_back(x[0]); fm.coords.push_back(x[1]); fm.coords.push_back(x[2]);
        if(sp.size() == 2) mids.emplace(edge_key(cv[sp[0]], cv[sp[1]]), id);
      }
      gid[r] = id;
    }
    for(int ch = 0; ch < n_children; ++ch) {
      for(int a = 0; a < fm.nv; ++a) fm.cell_vertices.push_back(gid[rt.child[ch][a]]);
      fm.parent.push_back(c);
      fm.child_index.push_back(ch);
      for(int jj = 0; jj < fm.nj; ++jj) {
        int comp = -1;
        for(size_t j = 0; j < faces.size() && comp < 0; ++j) {
          const int pc = cm.joint_comp[(size_t)c * cm.nj + j];
          if(pc < 0) continue;
          bool in = true;
          for(int lv : faces[jj])
            if(!on_joint[j][rt.child[ch][lv]]) in = false;
          if(in) comp = pc;
        }
        fm.joint_comp.push_back(comp);
      }
    }
  }
  fm.n_cells = (int)fm.parent.size();
  return fm;
}

// children of a triangle bisected on edge j (vertices v_j, v_j+1, midpoint
// m): {v_j, m, v_j+2} and {v_j+1, v_j+2, m} (RefTriBis0Desc.C:21-22,
// RefTriBis1Desc.C:22-23, RefTriBis2Desc.C:22-23)
void bisect_children(const int* v, int j, int m, int ch[2][3]) {
  ch[0][0] = v[j]; ch[0][1] = m; ch[0][2] = v[(j + 2) % 3];
  ch[1][0] = v[(j + 1) % 3]; ch[1][1] = v[(j + 2) % 3]; ch[1][2] = m;
}

}  // namespace

Mesh refine_by_indicator(const Mesh& base, int steps, const IndicatorFn& f, std::vector<HangingVertex>& hanging,
                         bool conforming) {
  PMCD_CHECK(base.dim() == 2, "refinement by indicator is implemented for 2D meshes");
  PMCD_CHECK(!conforming || base.shape == Shape::Triangle,
             "the conforming closure is implemented for triangles (for quadrilaterals the reference makes "
             "mixed meshes)");
  EdgeMids mids;
  Mesh cur = base;
  const auto& faces = shape_faces(cur.shape);
  auto straddles = [&](const int* v, int n) {
    int inner = 0, outer = 0;
    for(int a = 0; a < n; ++a) {
      const double val = f(&cur.coords[3 * v[a]]);
      if(val <= 1e-10) inner++;
      if(val >= -1e-10) outer++;
    }
    return inner > 0 && outer > 0;
  };
  auto refined_edges = [&](int c, int* last) {
    int n = 0;
    for(int j = 0; j < cur.nj; ++j)
      if(edge_mid(mids, cur.cell(c)[faces[j][0]], cur.cell(c)[faces[j][1]]) >= 0) {
        n++;
        if(last) *last = j;
      }
    return n;
  };
  // refine the marked leaves, then every leaf with an edge (u, v) whose
  // midpoint m has a half (u, m) or (m, v) split again (a neighbour two
  // levels finer), until the mesh is 1-irregular
  auto refine_1irregular = [&](std::vector<char>& mark) {
    for(;;) {
      cur = refine_marked(cur, mark, mids);
      mark.assign(cur.n_cells, 0);
      bool any = false;
      for(int c = 0; c < cur.n_cells; ++c)
        for(int j = 0; j < cur.nj; ++j) {
          const int u = cur.cell(c)[faces[j][0]], v = cur.cell(c)[faces[j][1]];
          const int m = edge_mid(mids, u, v);
          if(m >= 0 && (edge_mid(mids, u, m) >= 0 || edge_mid(mids, m, v) >= 0)) {
            mark[c] = 1;
            any = true;
          }
        }
      if(!any) break;
    }
  };
  std::vector<int> bis;  // conforming: bisected edge of each leaf of cur, or -1
  for(int step = 0; step < steps; ++step) {
    // RefinementStrategy::apply_indicator: corner values on both sides of
    // zero; a bisected cell (its closure removed first) is marked when one
    // of its halves is
    std::vector<char> mark(cur.n_cells, 0);
    for(int c = 0; c < cur.n_cells; ++c) {
      const int* v = cur.cell(c);
      if(!bis.empty() && bis[c] >= 0) {
        int ch[2][3];
        bisect_children(v, bis[c], edge_mid(mids, v[bis[c]], v[(bis[c] + 1) % 3]), ch);
        mark[c] = straddles(ch[0], 3) || straddles(ch[1], 3);
      } else {
        mark[c] = straddles(v, cur.nv);
      }
    }
    refine_1irregular(mark);
    bis.clear();
    if(!conforming) continue;
    // TDomain::MakeConfClosure (Domain.C:1219-1360, GridCell.C:525-575):
    // triangles with two or three refined edges are refined regularly until
    // none is left; a triangle with one refined edge is bisected on it
    for(;;) {
      mark.assign(cur.n_cells, 0);
      bool any = false;
      for(int c = 0; c < cur.n_cells; ++c)
        if(refined_edges(c, nullptr) >= 2) mark[c] = 1, any = true;
      if(!any) break;
      refine_1irregular(mark);
    }
    bis.assign(cur.n_cells, -1);
    for(int c = 0; c < cur.n_cells; ++c) {
      int j = -1;
      if(refined_edges(c, &j) == 1) bis[c] = j;
    }
  }
  hanging.clear();
  if(!bis.empty()) {
    // the leaves with each bisected cell replaced by its halves; a half's
    // joints along the parent's edges keep their components
    Mesh out;
    out.shape = cur.shape;
    out.nv = cur.nv;
    out.nj = cur.nj;
    out.coords = cur.coords;
    for(int c = 0; c < cur.n_cells; ++c) {
      const int* v = cur.cell(c);
      const int* jc = &cur.joint_comp[(size_t)c * cur.nj];
      if(bis[c] < 0) {
        out.cell_vertices.insert(out.cell_vertices.end(), v, v + 3);
        out.joint_comp.insert(out.joint_comp.end(), jc, jc + 3);
        out.parent.push_back(c);
        out.child_index.push_back(0);
        continue;
      }
      const int j = bis[c];
      int ch[2][3];
      bisect_children(v, j, edge_mid(mids, v[j], v[(j + 1) % 3]), ch);
      // edges of {v_j, m, v_j+2}: half of edge j, interior, edge j+2;
      // of {v_j+1, v_j+2, m}: edge j+1, interior, half of edge j
      const int comps[2][3] = {{jc[j], -1, jc[(j + 2) % 3]}, {jc[(j + 1) % 3], -1, jc[j]}};
      for(int h = 0; h < 2; ++h) {
        out.cell_vertices.insert(out.cell_vertices.end(), ch[h], ch[h] + 3);
        out.joint_comp.insert(out.joint_comp.end(), comps[h], comps[h] + 3);
        out.parent.push_back(c);
        out.child_index.push_back(h);
      }
    }
    out.n_cells = (int)out.parent.size();
    return out;
  }
  std::vector<char> seen(cur.n_vertices(), 0);
  for(int c = 0; c < cur.n_cells; ++c)
    for(int j = 0; j < cur.nj; ++j) {
      const int u = cur.cell(c)[faces[j][0]], v = cur.cell(c)[faces[j][1]];
      const int m = edge_mid(mids, u, v);
      if(m < 0 || seen[m]) continue;
      seen[m] = 1;
      hanging.push_back({m, u, v});
    }
  return cur;
}

double cell_diameter(const Mesh& m, int c) {
  const int* v = m.cell(c);
  auto d2 = [&](int a, int b) {
    const double* x = &m.coords[3 * v[a]];
    const double* y = &m.coords[3 * v[b]];
    double s = 0;
    for(int k = 0; k < 3; ++k) s += (x[k] - y[k]) * (x[k] - y[k]);
    return s;
  };
  double r = 0;
  switch(m.shape) {
    case Shape::Triangle: r = std::max(d2(0, 1), std::max(d2(1, 2), d2(2, 0))); break;
    case Shape::Quadrangle: r = std::max(d2(0, 2), d2(1, 3)); break;
    case Shape::Tetrahedron: {
      const int e[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
      for(auto& p : e) r = std::max(r, d2(p[0], p[1]));
      break;
    }
    case Shape::Hexahedron:
      r = std::max(std::max(d2(0, 6), d2(1, 7)), std::max(d2(2, 4), d2(3, 5)));
      break;
  }
  return std::sqrt(r);
}

namespace {

// whitespace-separated tokens of a .mesh file, '#' comments removed
std::vector<std::string> mesh_tokens(const std::string& path) {
  std::ifstream in(path);
  PMCD_CHECK(in.good(), "cannot open mesh file ", path);
  std::vector<std::string> tok;
  std::string line;
  while(std::getline(in, line)) {
    const size_t h = line.find('#');
    if(h != std::string::npos) line.erase(h);
    std::istringstream ls(line);
    std::string t;
    while(ls >> t) tok.push_back(t);
  }
  return tok;
}

}  // namespace

Mesh read_geo_file(const std::string& path) {
  std::ifstream in(path);
  PMCD_CHECK(in.good(), "cannot open .GEO file ", path);
  std::string line;
  std::getline(in, line);  // title
  std::getline(in, line);  // parametrisation comment
  int nel = 0, nvt = 0, nmt = 0, nve = 0, nbct = 0;
  in >> nel >> nvt >> nmt >> nve >> nbct;
  PMCD_CHECK(!in.fail() && nel > 0 && nvt > 0, ".GEO file ", path, ": bad counts line");
  PMCD_CHECK(nve == 3 || nve == 4, ".GEO file ", path, ": ", nve,
             " vertices per cell; 2D triangles (3) or quadrilaterals (4) are read");
  std::getline(in, line);  // rest of the counts line
  std::getline(in, line);  // DCORVG
  std::vector<double> dc(2 * (size_t)nvt);
  for(int i = 0; i < nvt; ++i) {
    in >> dc[2 * i] >> dc[2 * i + 1];
    std::getline(in, line);
  }
  std::getline(in, line);  // KVERT
  std::vector<int> kv((size_t)nel * nve);
  for(int c = 0; c < nel; ++c) {
    for(int j = 0; j < nve; ++j) in >> kv[(size_t)c * nve + j];
    std::getline(in, line);
  }
  std::getline(in, line);  // KNPR
  std::vector<int> knpr(nvt);
  for(int i = 0; i < nvt; ++i) in >> knpr[i];
  PMCD_CHECK(!in.fail(), ".GEO file ", path, " ends before its KNPR section is complete");
  Mesh m;
  m.shape = nve == 3 ? Shape::Triangle : Shape::Quadrangle;
  m.nv = nve;
  // boundary vertices carry the boundary parameter T: component floor(T),
  // local parameter T - floor(T) on the unit square's line segments
  // (Default_UnitSquare: (0,0)-(1,0), (1,0)-(1,1), (1,1)-(0,1), (0,1)-(0,0))
  static const double seg[4][4] = {{0, 0, 1, 0}, {1, 0, 1, 1}, {1, 1, 0, 1}, {0, 1, 0, 0}};
  m.coords.resize(3 * (size_t)nvt, 0.0);
  for(int i = 0; i < nvt; ++i) {
    if(knpr[i] == 0) {
      m.coords[3 * i] = dc[2 * i];
      m.coords[3 * i + 1] = dc[2 * i + 1];
      continue;
    }
    PMCD_CHECK(knpr[i] == 1, ".GEO file ", path, ": boundary part ", knpr[i],
               "; only the unit square boundary (one part, 4 components) is known");
    const double T = dc[2 * i];
    const int comp = (int)std::floor(T);
    PMCD_CHECK(comp >= 0 && comp < 4, ".GEO file ", path, ": boundary parameter ", T, " outside [0, 4)");
    const double t = T - comp;
    m.coords[3 * i] = seg[comp][0] + t * (seg[comp][2] - seg[comp][0]);
    m.coords[3 * i + 1] = seg[comp][1] + t * (seg[comp][3] - seg[comp][1]);
  }
  m.cell_vertices.resize(kv.size());
  for(size_t p = 0; p < kv.size(); ++p) {
    PMCD_CHECK(kv[p] >= 1 && kv[p] <= nvt, ".GEO file ", path, ": vertex id ", kv[p], " out of range");
    m.cell_vertices[p] = kv[p] - 1;
  }
  m.n_cells = nel;
  m.parent.assign(m.n_cells, -1);
  m.child_index.assign(m.n_cells, 0);
  set_box_boundary(m);
  return m;
}

Mesh read_mesh_file(const std::string& path) {
  const std::vector<std::string> tok = mesh_tokens(path);
  size_t p = 0;
  auto next = [&]() -> const std::string& {
    PMCD_CHECK(p < tok.size(), "unexpected end of mesh file ", path);
    return tok[p++];
  };
  auto num_i = [&]() {
    const std::string& t = next();
    char* end = nullptr;
    const long v = std::strtol(t.c_str(), &end, 10);
    PMCD_CHECK(end && *end == 0, "mesh file ", path, ": expected an integer, got '", t, "'");
    return (int)v;
  };
  // an element's reference number; like the reference's stream reader, a
  // missing one at the end of a section (the next token is not a number)
  // reads as 0
  auto ref_i = [&]() {
    if(p >= tok.size()) return 0;
    char* end = nullptr;
    const long v = std::strtol(tok[p].c_str(), &end, 10);
    if(!(end && *end == 0)) return 0;
    ++p;
    return (int)v;
  };
  auto num_d = [&]() {
    const std::string& t = next();
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    PMCD_CHECK(end && *end == 0, "mesh file ", path, ": expected a number, got '", t, "'");
    return v;
  };
  int dim = 0;
  std::vector<double> coords;
  std::vector<int> tets, hexes, tris, quads;  // 1-based ids
  while(p < tok.size()) {
    const std::string key = next();
    if(key == "MeshVersionFormatted") {
      num_i();
    } else if(key == "Dimension") {
      dim = num_i();
      PMCD_CHECK(dim == 2 || dim == 3, "mesh file ", path, ": dimension ", dim);
    } else if(key == "Vertices") {
      PMCD_CHECK(dim != 0, "mesh file ", path, ": Vertices before Dimension");
      const int n = num_i();
      PMCD_CHECK(n >= 0, "mesh file ", path, ": negative vertex count");
      coords.assign((size_t)3 * n, 0.0);
      for(int v = 0; v < n; ++v) {
        for(int k = 0; k < dim; ++k) coords[3 * (size_t)v + k] = num_d();
        num_i();  // vertex reference
      }
    } else if(key == "Edges" || key == "Triangles" || key == "Quadrilaterals" || key == "Tetrahedra" ||
              key == "Hexahedra") {
      const int nn = key == "Edges" ? 2 : key == "Triangles" ? 3 : key == "Tetrahedra" || key == "Quadrilaterals" ? 4 : 8;
      std::vector<int>* dst = key == "Triangles" ? &tris : key == "Quadrilaterals" ? &quads
                              : key == "Tetrahedra" ? &tets : key == "Hexahedra" ? &hexes : nullptr;
      const int n = num_i();
      PMCD_CHECK(n >= 0, "mesh file ", path, ": negative count in ", key);
      for(int e = 0; e < n; ++e) {
        for(int k = 0; k < nn; ++k) {
          const int id = num_i();
          if(dst) dst->push_back(id);
        }
        const int ref = ref_i();
        if(dst) dst->push_back(ref);
      }
    } else if(key == "End") {
      break;
    } else {
      PMCD_FAIL("mesh file ", path, ": unsupported section '", key, "'");
    }
  }
  PMCD_CHECK(!(tets.empty() && hexes.empty()), "mesh file ", path,
             ": no Tetrahedra or Hexahedra (2D meshes need a .PRM boundary description, not supported)");
  PMCD_CHECK(tets.empty() || hexes.empty(), "mesh file ", path, ": mixed tetrahedra and hexahedra");
  Mesh m;
  const bool tetra = !tets.empty();
  m.shape = tetra ? Shape::Tetrahedron : Shape::Hexahedron;
  m.nv = tetra ? 4 : 8;
  m.coords = std::move(coords);
  const int nvert = m.n_vertices();
  const std::vector<int>& cells = tetra ? tets : hexes;
  m.n_cells = (int)(cells.size() / (m.nv + 1));
  m.cell_vertices.resize((size_t)m.n_cells * m.nv);
  for(int c = 0; c < m.n_cells; ++c)
    for(int k = 0; k < m.nv; ++k) {
      const int id = cells[(size_t)c * (m.nv + 1) + k];
      PMCD_CHECK(id >= 1 && id <= nvert, "mesh file ", path, ": vertex id ", id, " out of range");
      m.cell_vertices[(size_t)c * m.nv + k] = id - 1;
    }
  // faces: a face of one cell is a boundary joint
  const auto& faces = shape_faces(m.shape);
  m.nj = (int)faces.size();
  const int fn = tetra ? 3 : 4;
  auto key_of = [&](const int* v) {
    std::array<int, 4> k = {v[0], v[1], v[2], fn == 4 ? v[3] : -1};
    std::sort(k.begin(), k.begin() + fn);
    return k;
  };
  std::map<std::array<int, 4>, int> count;
  for(int c = 0; c < m.n_cells; ++c)
    for(int j = 0; j < m.nj; ++j) {
      int v[4] = {-1, -1, -1, -1};
      for(int a = 0; a < fn; ++a) v[a] = m.cell(c)[faces[j][a]];
      ++count[key_of(v)];
    }
  // boundary components: file faces on the boundary, in file order
  std::map<std::array<int, 4>, int> comp_of;
  int n_comp = 0;
  const std::vector<int>& bf = tetra ? tris : quads;
  for(size_t f = 0; f < bf.size(); f += fn + 1) {
    int v[4] = {-1, -1, -1, -1};
    for(int a = 0; a < fn; ++a) {
      PMCD_CHECK(bf[f + a] >= 1 && bf[f + a] <= nvert, "mesh file ", path, ": face vertex out of range");
      v[a] = bf[f + a] - 1;
    }
    const auto k = key_of(v);
    auto it = count.find(k);
    if(it != count.end() && it->second == 1 && !comp_of.count(k)) comp_of[k] = n_comp++;
  }
  m.joint_comp.assign((size_t)m.n_cells * m.nj, -1);
  for(int c = 0; c < m.n_cells; ++c)
    for(int j = 0; j < m.nj; ++j) {
      int v[4] = {-1, -1, -1, -1};
      for(int a = 0; a < fn; ++a) v[a] = m.cell(c)[faces[j][a]];
      const auto k = key_of(v);
      const int cnt = count[k];
      PMCD_CHECK(cnt <= 2, "mesh file ", path, ": a face shared by ", cnt, " cells");
      if(cnt == 2) continue;
      auto it = comp_of.find(k);
      if(it == comp_of.end()) it = comp_of.emplace(k, n_comp++).first;
      m.joint_comp[(size_t)c * m.nj + j] = it->second;
    }
  m.parent.assign(m.n_cells, -1);
  m.child_index.assign(m.n_cells, 0);
  return m;
}

void write_mesh_file(const Mesh& m, const std::string& path) {
  PMCD_CHECK(m.dim() == 3, "only 3D meshes are written");
  std::ofstream out(path);
  PMCD_CHECK(out.good(), "cannot write mesh file ", path);
  out.precision(17);
  out << "MeshVersionFormatted 1\n\nDimension 3\n\nVertices\n" << m.n_vertices() << "\n";
  for(int v = 0; v < m.n_vertices(); ++v)
    out << m.coords[3 * v] << " " << m.coords[3 * v + 1] << " " << m.coords[3 * v + 2] << " 0\n";
  const auto& faces = shape_faces(m.shape);
  const bool tetra = m.shape == Shape::Tetrahedron;
  std::ostringstream bf;
  int n_bf = 0;
  for(int c = 0; c < m.n_cells; ++c)
    for(int j = 0; j < m.nj; ++j) {
      const int comp = m.joint_comp[(size_t)c * m.nj + j];
      if(comp < 0) continue;
      for(int lv : faces[j]) bf << m.cell(c)[lv] + 1 << " ";
      bf << comp + 1 << "\n";
      ++n_bf;
    }
  out << "\n" << (tetra ? "Triangles" : "Quadrilaterals") << "\n" << n_bf << "\n" << bf.str();
  out << "\n" << (tetra ? "Tetrahedra" : "Hexahedra") << "\n" << m.n_cells << "\n";
  for(int c = 0; c < m.n_cells; ++c) {
    for(int k = 0; k < m.nv; ++k) out << m.cell(c)[k] + 1 << " ";
    out << "1\n";
  }
  out << "\nEnd\n";
  PMCD_CHECK(out.good(), "error writing mesh file ", path);
}

}  // namespace pmcd
