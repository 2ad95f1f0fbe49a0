// Lagrange elements of order <= 2. Tensor-product elements live on [-1,1]^d
// with local index ix + (p+1) iy + (p+1)^2 iz (include/FE3D/BF_C_H_Q2_3D.h:5-68
// uses this order); simplicial elements live on the unit simplex with the
// reference's vertex/edge-midpoint numbering (include/FE2D/BF_C_T_P2_2D.h,
// include/FE3D/BF_C_T_P2_3D.h).
#include "elements.hpp"
#include <algorithm>

namespace pmcd {

Element element_for(Shape s, int order) {
  PMCD_CHECK(order == 1 || order == 2, "only ansatz orders 1 and 2 are supported, got ", order);
  switch(s) {
    case Shape::Triangle: return order == 1 ? Element::P1Tri : Element::P2Tri;
    case Shape::Quadrangle: return order == 1 ? Element::Q1Quad : Element::Q2Quad;
    case Shape::Tetrahedron: return order == 1 ? Element::P1Tet : Element::P2Tet;
    case Shape::Hexahedron: return order == 1 ? Element::Q1Hex : Element::Q2Hex;
  }
  PMCD_FAIL("bad shape");
}

int element_order(Element e) {
  switch(e) {
    case Element::P1Tri: case Element::Q1Quad: case Element::P1Tet: case Element::Q1Hex: return 1;
    default: return 2;
  }
}

Shape element_shape(Element e) {
  switch(e) {
    case Element::P1Tri: case Element::P2Tri: return Shape::Triangle;
    case Element::Q1Quad: case Element::Q2Quad: return Shape::Quadrangle;
    case Element::P1Tet: case Element::P2Tet: return Shape::Tetrahedron;
    default: return Shape::Hexahedron;
  }
}

int element_n_dof(Element e) { return element_info(e).n_dof; }

const std::vector<std::vector<int>>& shape_faces(Shape s) {
  // Triangle.C / Quadrangle.C edge j = (v_j, v_{j+1});
  // Tetrahedron.C:18-19 and Hexahedron.C:20-22 face vertex tables.
  static const std::vector<std::vector<int>> tri = {{0, 1}, {1, 2}, {2, 0}};
  static const std::vector<std::vector<int>> quad = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const std::vector<std::vector<int>> tet = {{0, 1, 2}, {0, 3, 1}, {2, 1, 3}, {0, 2, 3}};
  static const std::vector<std::vector<int>> hex = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2},
                                                    {2, 6, 7, 3}, {0, 3, 7, 4}, {4, 7, 6, 5}};
  switch(s) {
    case Shape::Triangle: return tri;
    case Shape::Quadrangle: return quad;
    case Shape::Tetrahedron: return tet;
    default: return hex;
  }
}

namespace {

// reference vertex index of a tensor corner (cx, cy, cz) in {0,1}^d
int tensor_corner_vertex(int cx, int cy, int cz) {
  static const int q[2][2] = {{0, 3}, {1, 2}};  // [cx][cy]
  return 4 * cz + q[cx][cy];
}

ElementInfo build_tensor(Element e, Shape s, int p) {
  ElementInfo info;
  info.type = e; info.shape = s; info.order = p;
  const int d = shape_dim(s);
  const int n1 = p + 1;
  info.n_dof = d == 2 ? n1 * n1 : n1 * n1 * n1;
  const int nz = d == 3 ? n1 : 1;
  for(int iz = 0; iz < nz; ++iz)
    for(int iy = 0; iy < n1; ++iy)
      for(int ix = 0; ix < n1; ++ix) {
        auto coord = [&](int i) { return -1.0 + 2.0 * i / p; };
        info.nodes.push_back({coord(ix), coord(iy), d == 3 ? coord(iz) : 0.0});
        // corners spanning the entity: an axis index at an end fixes the
        // corner coordinate, an interior index leaves both
        auto choices = [&](int i) {
          std::vector<int> c;
          if(i == 0) c = {0};
          else if(i == p) c = {1};
          else c = {0, 1};
          return c;
        };
        std::vector<int> verts;
        for(int cz : (d == 3 ? choices(iz) : std::vector<int>{0}))
          for(int cy : choices(iy))
            for(int cx : choices(ix)) verts.push_back(tensor_corner_vertex(cx, cy, cz));
        std::sort(verts.begin(), verts.end());
        info.entity_vertices.push_back(verts);
      }
  return info;
}

ElementInfo build_simplex(Element e, Shape s, int p) {
  ElementInfo info;
  info.type = e; info.shape = s; info.order = p;
  const int d = shape_dim(s);
  std::vector<std::array<double, 3>> vc = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  std::vector<std::vector<int>> ent;
  if(p == 1) {
    for(int v = 0; v <= d; ++v) ent.push_back({v});
  } else if(d == 2) {
    ent = {{0}, {0, 1}, {1}, {0, 2}, {1, 2}, {2}};
  } else {
    ent = {{0}, {0, 1}, {1}, {0, 2}, {1, 2}, {2}, {0, 3}, {1, 3}, {2, 3}, {3}};
  }
  info.n_dof = (int)ent.size();
  for(auto& en : ent) {
    std::array<double, 3> x = {0, 0, 0};
    for(int v : en)
      for(int k = 0; k < 3; ++k) x[k] += vc[v][k] / en.size();
    info.nodes.push_back(x);
    info.entity_vertices.push_back(en);
  }
  return info;
}

std::vector<ElementInfo> build_all() {
  std::vector<ElementInfo> all;
  all.push_back(build_simplex(Element::P1Tri, Shape::Triangle, 1));
  all.push_back(build_simplex(Element::P2Tri, Shape::Triangle, 2));
  all.push_back(build_tensor(Element::Q1Quad, Shape::Quadrangle, 1));
  all.push_back(build_tensor(Element::Q2Quad, Shape::Quadrangle, 2));
  all.push_back(build_simplex(Element::P1Tet, Shape::Tetrahedron, 1));
  all.push_back(build_simplex(Element::P2Tet, Shape::Tetrahedron, 2));
  all.push_back(build_tensor(Element::Q1Hex, Shape::Hexahedron, 1));
  all.push_back(build_tensor(Element::Q2Hex, Shape::Hexahedron, 2));
  for(auto& info : all) {
    for(auto& f : shape_faces(info.shape)) {
      std::vector<int> jd;
      for(int i = 0; i < info.n_dof; ++i) {
        bool inside = true;
        for(int v : info.entity_vertices[i])
          if(std::find(f.begin(), f.end(), v) == f.end()) inside = false;
        if(inside) jd.push_back(i);
      }
      info.joint_dofs.push_back(jd);
    }
  }
  return all;
}

// 1D Lagrange basis on [-1,1], order 1 (nodes -1,1) or 2 (nodes -1,0,1)
inline void lagrange1d(int p, double x, double* v, double* d1, double* d2) {
  if(p == 1) {
    v[0] = 0.5 * (1 - x); v[1] = 0.5 * (1 + x);
    d1[0] = -0.5; d1[1] = 0.5;
    d2[0] = 0.0; d2[1] = 0.0;
  } else {
    v[0] = 0.5 * x * (x - 1); v[1] = 1 - x * x; v[2] = 0.5 * x * (x + 1);
    d1[0] = x - 0.5; d1[1] = -2 * x; d1[2] = x + 0.5;
    d2[0] = 1.0; d2[1] = -2.0; d2[2] = 1.0;
  }
}

void eval_tensor(const ElementInfo& info, const double* pt, double* out) {
  const int d = shape_dim(info.shape), p = info.order, n1 = p + 1, n = info.n_dof;
  double v[3][3], a[3][3], b[3][3];
  for(int k = 0; k < d; ++k) lagrange1d(p, pt[k], v[k], a[k], b[k]);
  if(d == 2) {
    for(int iy = 0; iy < n1; ++iy)
      for(int ix = 0; ix < n1; ++ix) {
        int i = ix + n1 * iy;
        out[0 * n + i] = v[0][ix] * v[1][iy];
        out[1 * n + i] = a[0][ix] * v[1][iy];
        out[2 * n + i] = v[0][ix] * a[1][iy];
        out[3 * n + i] = b[0][ix] * v[1][iy];
        out[4 * n + i] = a[0][ix] * a[1][iy];
        out[5 * n + i] = v[0][ix] * b[1][iy];
      }
  } else {
    for(int iz = 0; iz < n1; ++iz)
      for(int iy = 0; iy < n1; ++iy)
        for(int ix = 0; ix < n1; ++ix) {
          int i = ix + n1 * iy + n1 * n1 * iz;
          out[0 * n + i] = v[0][ix] * v[1][iy] * v[2][iz];
          out[1 * n + i] = a[0][ix] * v[1][iy] * v[2][iz];
          out[2 * n + i] = v[0][ix] * a[1][iy] * v[2][iz];
          out[3 * n + i] = v[0][ix] * v[1][iy] * a[2][iz];
          out[4 * n + i] = b[0][ix] * v[1][iy] * v[2][iz];
          out[5 * n + i] = a[0][ix] * a[1][iy] * v[2][iz];
          out[6 * n + i] = a[0][ix] * v[1][iy] * a[2][iz];
          out[7 * n + i] = v[0][ix] * b[1][iy] * v[2][iz];
          out[8 * n + i] = v[0][ix] * a[1][iy] * a[2][iz];
          out[9 * n + i] = v[0][ix] * v[1][iy] * b[2][iz];
        }
  }
}

void eval_simplex(const ElementInfo& info, const double* pt, double* out) {
  const int d = shape_dim(info.shape), n = info.n_dof;
  double lam[4], g[4][3] = {{0}};
  lam[0] = 1.0;
  for(int k = 0; k < d; ++k) { lam[0] -= pt[k]; lam[k + 1] = pt[k]; }
  for(int k = 0; k < d; ++k) { g[0][k] = -1.0; g[k + 1][k] = 1.0; }
  // second-derivative component list (k,l) in multi-index order
  const int pairs2[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  const int pairs3[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  const int n2 = d == 2 ? 3 : 6;
  for(int i = 0; i < n; ++i) {
    const auto& en = info.entity_vertices[i];
    double val, grad[3], hess[6];
    if(info.order == 1) {
      int a = en[0];
      val = lam[a];
      for(int k = 0; k < d; ++k) grad[k] = g[a][k];
      for(int m = 0; m < n2; ++m) hess[m] = 0.0;
    } else if(en.size() == 1) {
      int a = en[0];
      val = lam[a] * (2 * lam[a] - 1);
      for(int k = 0; k < d; ++k) grad[k] = (4 * lam[a] - 1) * g[a][k];
      for(int m = 0; m < n2; ++m) {
        int k = d == 2 ? pairs2[m][0] : pairs3[m][0], l = d == 2 ? pairs2[m][1] : pairs3[m][1];
        hess[m] = 4 * g[a][k] * g[a][l];
      }
    } else {
      int a = en[0], b = en[1];
      val = 4 * lam[a] * lam[b];
      for(int k = 0; k < d; ++k) grad[k] = 4 * (g[a][k] * lam[b] + lam[a] * g[b][k]);
      for(int m = 0; m < n2; ++m) {
        int k = d == 2 ? pairs2[m][0] : pairs3[m][0], l = d == 2 ? pairs2[m][1] : pairs3[m][1];
        hess[m] = 4 * (g[a][k] * g[b][l] + g[a][l] * g[b][k]);
      }
    }
    out[i] = val;
    for(int k = 0; k < d; ++k) out[(1 + k) * n + i] = grad[k];
    for(int m = 0; m < n2; ++m) out[(1 + d + m) * n + i] = hess[m];
  }
}

}  // namespace

const ElementInfo& element_info(Element e) {
  static std::vector<ElementInfo> all = build_all();
  return all[(int)e];
}

void eval_basis(Element e, const double* p, double* out) {
  const ElementInfo& info = element_info(e);
  if(shape_is_simplex(info.shape)) eval_simplex(info, p, out);
  else eval_tensor(info, p, out);
}

}  // namespace pmcd
