#include "fespace.hpp"
#include <algorithm>
#include <unordered_map>
#include <climits>

namespace pmcd {

namespace {
struct Key3Hash {
  size_t operator()(const std::array<int, 4>& k) const {
    uint64_t h = 1469598103934665603ull;
    for(int v : k) { h ^= (uint64_t)(uint32_t)v; h *= 1099511628211ull; }
    return (size_t)h;
  }
};
}  // namespace

FESpace build_fespace(const Mesh& mesh, int order, const BoundCondFn& bc, const std::vector<HangingVertex>* hanging) {
  FESpace sp;
  sp.element = element_for(mesh.shape, order);
  const ElementInfo& info = element_info(sp.element);
  sp.n_local = info.n_dof;
  const int nl = sp.n_local;
  const int nc = mesh.n_cells;

  // class of each (cell, local dof), classes created in first-appearance order
  std::vector<int> cls((size_t)nc * nl);
  std::vector<int> vertex_class(mesh.n_vertices(), -1);
  std::unordered_map<std::array<int, 4>, int, Key3Hash> entity_class;
  entity_class.reserve((size_t)nc * 4);
  auto edge_key = [](int a, int b) {
    return std::array<int, 4>{std::min(a, b), std::max(a, b), -1, -1};
  };
  // order 2 with hanging vertices: the middle DOF of a refined coarse edge
  // is the DOF of the fine vertex at its midpoint
  std::unordered_map<std::array<int, 4>, int, Key3Hash> edge_midpoint;
  const bool has_hanging = hanging && !hanging->empty();
  if(has_hanging) {
    PMCD_CHECK((order == 1 || order == 2) && mesh.dim() == 2, "hanging vertices are implemented for 2D, order 1 or 2");
    if(order == 2)
      for(const HangingVertex& hv : *hanging) edge_midpoint.emplace(edge_key(hv.a, hv.b), hv.vertex);
  }
  int n_classes = 0;
  const int full = shape_n_vertices(mesh.shape);
  for(int c = 0; c < nc; ++c) {
    const int* cv = mesh.cell(c);
    for(int i = 0; i < nl; ++i) {
      const auto& ev = info.entity_vertices[i];
      int k;
      if(ev.size() == 1) {
        int& vc = vertex_class[cv[ev[0]]];
        if(vc < 0) vc = n_classes++;
        k = vc;
      } else if((int)ev.size() == full) {
        k = n_classes++;  // cell-interior DOF
      } else {
        std::array<int, 4> key = {-1, -1, -1, -1};
        int g[4];
        for(size_t a = 0; a < ev.size(); ++a) g[a] = cv[ev[a]];
        std::sort(g, g + ev.size());
        for(size_t a = 0; a < ev.size(); ++a) key[a] = g[a];
        auto mid = edge_midpoint.find(key);
        if(mid != edge_midpoint.end()) {
          int& vc = vertex_class[mid->second];
          if(vc < 0) vc = n_classes++;
          k = vc;
        } else {
          auto it = entity_class.find(key);
          if(it == entity_class.end()) { k = n_classes++; entity_class.emplace(key, k); }
          else k = it->second;
        }
      }
      cls[(size_t)c * nl + i] = k;
    }
  }

  // boundary kind of each class: the smallest BoundCond over boundary joints
  const int INNER = INT_MAX;
  std::vector<int> kind(n_classes, INNER);
  const auto& faces = shape_faces(mesh.shape);
  for(int c = 0; c < nc; ++c)
    for(int j = 0; j < mesh.nj; ++j) {
      int comp = mesh.joint_comp[(size_t)c * mesh.nj + j];
      if(comp < 0) continue;
      double cen[3] = {0, 0, 0};
      for(int lv : faces[j])
        for(int k = 0; k < 3; ++k) cen[k] += mesh.coords[3 * mesh.cell(c)[lv] + k];
      for(int k = 0; k < 3; ++k) cen[k] /= faces[j].size();
      int b = (int)bc(comp, cen);
      for(int i : info.joint_dofs[j]) {
        int k = cls[(size_t)c * nl + i];
        kind[k] = std::min(kind[k], b);
      }
    }

  // hanging DOFs: their own kind, numbered in the list's order; couplings
  // and weights by class
  constexpr int MC = FESpace::kMaxCoupling;
  const int HANGING = INT_MAX - 1;
  std::vector<int> hang_index(n_classes, -1);
  std::vector<int> hang_class, hang_n;
  std::vector<int> hang_cpl_class;
  std::vector<double> hang_w;
  if(has_hanging) {
    for(const HangingVertex& hv : *hanging) {
      const int ka = vertex_class[hv.a], kb = vertex_class[hv.b], km = vertex_class[hv.vertex];
      PMCD_CHECK(ka >= 0 && kb >= 0 && km >= 0, "hanging vertex without a DOF");
      if(order == 1) {
        hang_class.push_back(km);
        hang_n.push_back(2);
        hang_cpl_class.insert(hang_cpl_class.end(), {ka, kb, -1});
        hang_w.insert(hang_w.end(), {0.5, 0.5, 0.0});
      } else {
        const int half[2][2] = {{hv.a, hv.vertex}, {hv.vertex, hv.b}};
        const double wt[2][3] = {{0.375, 0.75, -0.125}, {-0.125, 0.75, 0.375}};
        for(int q = 0; q < 2; ++q) {
          auto it = entity_class.find(edge_key(half[q][0], half[q][1]));
          PMCD_CHECK(it != entity_class.end(), "fine edge of a hanging vertex without a DOF");
          hang_class.push_back(it->second);
          hang_n.push_back(3);
          hang_cpl_class.insert(hang_cpl_class.end(), {ka, km, kb});
          hang_w.insert(hang_w.end(), {wt[q][0], wt[q][1], wt[q][2]});
        }
      }
    }
    for(size_t h = 0; h < hang_class.size(); ++h) {
      kind[hang_class[h]] = HANGING;
      hang_index[hang_class[h]] = (int)h;
    }
  }

  // numbering: classes are already in first-appearance order
  constexpr int NK = 8;
  std::vector<int> count_kind(NK + 1, 0);  // [0..6] boundary kinds, [7] inner, [8] hanging
  auto slot = [&](int kd) { return kd == INNER ? NK - 1 : (kd == HANGING ? NK : kd); };
  for(int k = 0; k < n_classes; ++k) count_kind[slot(kind[k])]++;
  sp.n_inner = count_kind[NK - 1];
  sp.n_dirichlet = count_kind[DIRICHLET];
  sp.n_hanging = count_kind[NK];
  // offsets: inner, NEUMANN, ROBIN, ..., hanging, then DIRICHLET
  std::vector<int> offset(NK + 1, 0);
  int off = 0;
  offset[NK - 1] = off; off += count_kind[NK - 1];
  for(int kd = 1; kd < NK - 1; ++kd) { offset[kd] = off; off += count_kind[kd]; }
  offset[NK] = off; off += count_kind[NK];
  sp.n_active = off;
  offset[DIRICHLET] = off; off += count_kind[DIRICHLET];
  sp.n_dof = off;
  std::vector<int> number(n_classes);
  for(int k = 0; k < n_classes; ++k)
    number[k] = kind[k] == HANGING ? offset[NK] + hang_index[k] : offset[slot(kind[k])]++;
  sp.dof.resize((size_t)nc * nl);
  for(size_t p = 0; p < sp.dof.size(); ++p) sp.dof[p] = number[cls[p]];
  if(sp.n_hanging) {
    const int nh = sp.n_hanging, base = sp.n_active_non_hanging();
    PMCD_CHECK(nh == (int)hang_class.size(), "hanging DOF shared by two coarse edges");
    sp.hang_n = hang_n;
    sp.hang_w = hang_w;
    sp.hang_cpl.assign((size_t)MC * nh, -1);
    for(int h = 0; h < nh; ++h)
      for(int k = 0; k < hang_n[h]; ++k) sp.hang_cpl[MC * h + k] = number[hang_cpl_class[MC * h + k]];
    (void)base;
    sort_hanging(sp);
  }
  return sp;
}

void sort_hanging(FESpace& sp) {
  constexpr int MC = FESpace::kMaxCoupling;
  const int nh = sp.n_hanging, base = sp.n_active_non_hanging();
  // depth-first post-order over hanging coupling DOFs, reversed
  std::vector<char> seen(nh, 0);
  std::vector<int> post;
  post.reserve(nh);
  std::vector<std::pair<int, int>> stack;  // (hanging DOF, next coupling)
  for(int r = 0; r < nh; ++r) {
    if(seen[r]) continue;
    seen[r] = 1;
    stack.push_back({r, 0});
    while(!stack.empty()) {
      auto& top = stack.back();
      if(top.second < sp.hang_n[top.first]) {
        const int c = sp.hang_cpl[MC * top.first + top.second++] - base;
        if(c >= 0 && c < nh && !seen[c]) {
          seen[c] = 1;
          stack.push_back({c, 0});
        }
      } else {
        post.push_back(top.first);
        stack.pop_back();
      }
    }
  }
  PMCD_CHECK((int)post.size() == nh, "not all hanging DOFs sorted");
  sp.hang_order.assign(post.rbegin(), post.rend());
}

}  // namespace pmcd
