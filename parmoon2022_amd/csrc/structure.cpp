#include "structure.hpp"
#include <algorithm>
#include <climits>
#include <iterator>

namespace pmcd {

bool pattern_symmetric(int n, const int* row_ptr, const int* col) {
  for(int r = 0; r < n; ++r)
    for(int k = row_ptr[r] + 1; k < row_ptr[r + 1]; ++k)
      if(col[k - 1] >= col[k]) return false;  // unsorted rows are taken as not symmetric
  for(int r = 0; r < n; ++r)
    for(int k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      if(!std::binary_search(col + row_ptr[col[k]], col + row_ptr[col[k] + 1], r)) return false;
  return true;
}

Structure structure_from_pattern(const FESpace& sp, int n_cells, const int* row_ptr, const int* col) {
  Structure st;
  const int nl = sp.n_local;
  const int n = sp.n_dof;
  PMCD_CHECK(row_ptr && col, "null structure arrays");
  PMCD_CHECK(row_ptr[0] == 0, "row_ptr[0] must be 0");
  st.n_rows = n;
  st.row_ptr.assign(row_ptr, row_ptr + n + 1);
  for(int r = 0; r < n; ++r) {
    const int len = row_ptr[r + 1] - row_ptr[r];
    PMCD_CHECK(len >= 0, "row_ptr decreases at row ", r);
    st.max_row = std::max(st.max_row, len);
  }
  st.nnz = row_ptr[n];
  PMCD_CHECK(st.max_row <= 255, "row with ", st.max_row, " entries exceeds the 8-bit slot map");
  st.col.assign(col, col + st.nnz);
  for(int64_t k = 0; k < st.nnz; ++k) PMCD_CHECK(col[k] >= 0 && col[k] < n, "column index out of range");
  st.contrib_ptr.assign(n + 1, 0);
  for(size_t p = 0; p < sp.dof.size(); ++p) st.contrib_ptr[sp.dof[p] + 1]++;
  for(int r = 0; r < n; ++r) st.contrib_ptr[r + 1] += st.contrib_ptr[r];
  st.contrib.resize(sp.dof.size());
  {
    std::vector<int> fill(st.contrib_ptr.begin(), st.contrib_ptr.end() - 1);
    for(int c = 0; c < n_cells; ++c)
      for(int m = 0; m < nl; ++m) st.contrib[fill[sp.dof[(size_t)c * nl + m]]++] = c * nl + m;
  }
  // positions by linear search in each row, as the reference's scatter finds
  // them (Assemble3D.C:258-292); a coupling missing from the pattern is an
  // error (the reference prints "not found" and drops the entry)
  st.diag.assign(n, -1);
  st.slot.resize(st.contrib.size() * (size_t)nl);
  st.symmetric = pattern_symmetric(n, row_ptr, col);
  for(int r = 0; r < n; ++r) {
    const int b = row_ptr[r], e = row_ptr[r + 1];
    for(int k = b; k < e; ++k)
      if(col[k] == r) { st.diag[r] = k; break; }
    PMCD_CHECK(st.diag[r] >= 0, "row ", r, " of the structure has no diagonal entry");
    for(int q = st.contrib_ptr[r]; q < st.contrib_ptr[r + 1]; ++q) {
      const int* d = sp.cell_dof(st.contrib[q] / nl);
      for(int k = 0; k < nl; ++k) {
        int pos = -1;
        for(int t = b; t < e; ++t)
          if(col[t] == d[k]) { pos = t - b; break; }
        PMCD_CHECK(pos >= 0, "coupling (", r, ", ", d[k], ") is not in the structure");
        st.slot[(size_t)q * nl + k] = (uint8_t)pos;
      }
    }
  }
  return st;
}

Structure build_structure(const FESpace& sp, int n_cells) {
  Structure st;
  const int nl = sp.n_local;
  const int n = sp.n_dof;
  st.n_rows = n;
  // DOF -> (cell, local) list, cells increasing
  st.contrib_ptr.assign(n + 1, 0);
  for(size_t p = 0; p < sp.dof.size(); ++p) st.contrib_ptr[sp.dof[p] + 1]++;
  for(int r = 0; r < n; ++r) st.contrib_ptr[r + 1] += st.contrib_ptr[r];
  st.contrib.resize(sp.dof.size());
  {
    std::vector<int> fill(st.contrib_ptr.begin(), st.contrib_ptr.end() - 1);
    for(int c = 0; c < n_cells; ++c)
      for(int m = 0; m < nl; ++m) st.contrib[fill[sp.dof[(size_t)c * nl + m]]++] = c * nl + m;
  }
  // rows: union of the DOFs of the cells containing the row DOF, ascending
  std::vector<int64_t> len(n, 0);
#pragma omp parallel
  {
    std::vector<int> buf;
#pragma omp for schedule(dynamic, 1024)
    for(int r = 0; r < n; ++r) {
      buf.clear();
      for(int e = st.contrib_ptr[r]; e < st.contrib_ptr[r + 1]; ++e) {
        int c = st.contrib[e] / nl;
        const int* d = sp.cell_dof(c);
        buf.insert(buf.end(), d, d + nl);
      }
      std::sort(buf.begin(), buf.end());
      len[r] = std::unique(buf.begin(), buf.end()) - buf.begin();
    }
  }
  st.row_ptr.assign(n + 1, 0);
  int64_t tot = 0;
  for(int r = 0; r < n; ++r) {
    tot += len[r];
    PMCD_CHECK(tot < INT_MAX, "matrix has more than 2^31 entries");
    st.row_ptr[r + 1] = (int)tot;
    st.max_row = std::max(st.max_row, (int)len[r]);
  }
  st.nnz = tot;
  PMCD_CHECK(st.max_row <= 255, "row with ", st.max_row, " entries exceeds the 8-bit slot map");
  st.col.resize(tot);
  st.diag.resize(n);
  st.slot.resize(st.contrib.size() * (size_t)nl);
#pragma omp parallel
  {
    std::vector<int> buf;
#pragma omp for schedule(dynamic, 1024)
    for(int r = 0; r < n; ++r) {
      buf.clear();
      for(int e = st.contrib_ptr[r]; e < st.contrib_ptr[r + 1]; ++e) {
        int c = st.contrib[e] / nl;
        const int* d = sp.cell_dof(c);
        buf.insert(buf.end(), d, d + nl);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      std::copy(buf.begin(), buf.end(), st.col.begin() + st.row_ptr[r]);
      st.diag[r] = st.row_ptr[r] + (int)(std::lower_bound(buf.begin(), buf.end(), r) - buf.begin());
      for(int e = st.contrib_ptr[r]; e < st.contrib_ptr[r + 1]; ++e) {
        int c = st.contrib[e] / nl;
        const int* d = sp.cell_dof(c);
        for(int k = 0; k < nl; ++k)
          st.slot[(size_t)e * nl + k] =
              (uint8_t)(std::lower_bound(buf.begin(), buf.end(), d[k]) - buf.begin());
      }
    }
  }
  if(sp.n_hanging) return hanging_structure(sp, n_cells, st);
  return st;
}

// Rows of coupling DOFs take the columns of the hanging rows added to them
// (FEMatrix::ModifyMatrixAccordingToCoupling, src/Matrix/FEMatrix.C:244-296,
// needs every entry it adds; Structure.C:75-230 puts them in the pattern of
// every coupling row, Dirichlet rows included, where they stay zero), in the
// sorted hanging order so that a hanging coupling DOF passes on what it
// received. For P1 triangles the reference also adds the transposed
// couplings (Structure.C:196-218), zeros only its AFC uses; they are left
// out.
Structure hanging_structure(const FESpace& sp, int n_cells, const Structure& base) {
  const int n = base.n_rows, first = sp.n_active_non_hanging();
  constexpr int MC = FESpace::kMaxCoupling;
  std::vector<std::vector<int>> rows(n);
  for(int r = 0; r < n; ++r) rows[r].assign(base.col.begin() + base.row_ptr[r], base.col.begin() + base.row_ptr[r + 1]);
  // a hanging row holds its coupling DOFs, also those outside its cells
  // (add_columns_far_hanging_coupling, Structure.C:142-170: order 2)
  for(int h = 0; h < sp.n_hanging; ++h) {
    auto& row = rows[first + h];
    for(int k = 0; k < sp.hang_n[h]; ++k) {
      const int c = sp.hang_cpl[MC * h + k];
      auto it = std::lower_bound(row.begin(), row.end(), c);
      if(it == row.end() || *it != c) row.insert(it, c);
    }
  }
  std::vector<int> merged;
  for(int h : sp.hang_order) {
    const int hd = first + h;
    for(int k = 0; k < sp.hang_n[h]; ++k) {
      const int c = sp.hang_cpl[MC * h + k];
      merged.clear();
      std::set_union(rows[c].begin(), rows[c].end(), rows[hd].begin(), rows[hd].end(), std::back_inserter(merged));
      rows[c].swap(merged);
    }
  }
  std::vector<int> row_ptr(n + 1, 0), col;
  for(int r = 0; r < n; ++r) row_ptr[r + 1] = row_ptr[r] + (int)rows[r].size();
  col.reserve(row_ptr[n]);
  for(int r = 0; r < n; ++r) col.insert(col.end(), rows[r].begin(), rows[r].end());
  return structure_from_pattern(sp, n_cells, row_ptr.data(), col.data());  // checks the symmetry
}

SweepOrder build_sweep_order(const Structure& st, int max_exact_classes) {
  const int n = st.n_rows;
  // rows c < r that hold column r (the transposed half of the adjacency,
  // needed when the pattern is not symmetric; a neighbour met twice is
  // harmless)
  std::vector<int> tp(n + 1, 0), tc;
  if(!st.symmetric) {
    for(int c = 0; c < n; ++c)
      for(int k = st.row_ptr[c]; k < st.row_ptr[c + 1]; ++k)
        if(st.col[k] > c) tp[st.col[k] + 1]++;
    for(int r = 0; r < n; ++r) tp[r + 1] += tp[r];
    tc.resize(tp[n]);
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for(int c = 0; c < n; ++c)
      for(int k = st.row_ptr[c]; k < st.row_ptr[c + 1]; ++k)
        if(st.col[k] > c) tc[fill[st.col[k]]++] = c;
  }
  // visits every row c < r adjacent to r
  auto lower = [&](int r, auto&& f) {
    for(int k = st.row_ptr[r]; k < st.row_ptr[r + 1]; ++k)
      if(st.col[k] < r) f(st.col[k]);
    for(int k = tp[r]; k < tp[r + 1]; ++k) f(tc[k]);
  };
  std::vector<int> cls(n, 0);
  int n_cls = n > 0 ? 1 : 0;
  for(int r = 0; r < n; ++r) {
    int c = 0;
    lower(r, [&](int j) { c = std::max(c, cls[j] + 1); });
    cls[r] = c;
    n_cls = std::max(n_cls, c + 1);
  }
  SweepOrder o;
  o.exact = n_cls <= max_exact_classes;
  if(!o.exact) {
    // greedy colouring in row order: the smallest colour no coloured
    // neighbour holds
    std::vector<int> stamp(1, -1);
    n_cls = 0;
    for(int r = 0; r < n; ++r) {
      lower(r, [&](int j) { stamp[cls[j]] = r; });
      int c = 0;
      while(c < (int)stamp.size() && stamp[c] == r) ++c;
      if(c == (int)stamp.size()) stamp.push_back(-1);
      cls[r] = c;
      n_cls = std::max(n_cls, c + 1);
    }
  }
  o.class_ptr.assign(n_cls + 1, 0);
  for(int r = 0; r < n; ++r) o.class_ptr[cls[r] + 1]++;
  for(int k = 0; k < n_cls; ++k) o.class_ptr[k + 1] += o.class_ptr[k];
  o.rows.resize(n);
  std::vector<int> fill(o.class_ptr.begin(), o.class_ptr.end() - 1);
  for(int r = 0; r < n; ++r) o.rows[fill[cls[r]]++] = r;
  return o;
}

SellLayout build_sell(const Structure& st, int n_active, int window) {
  PMCD_CHECK(window > 0 && window % kSellRows == 0, "sliced ELL window must be a multiple of ", kSellRows);
  SellLayout o;
  const int n_slots = (n_active + kSellRows - 1) / kSellRows * kSellRows;
  o.perm.assign(n_slots, -1);
  o.len.assign(n_slots, 0);
  auto row_len = [&](int r) { return st.row_ptr[r + 1] - st.row_ptr[r]; };
  for(int w0 = 0; w0 < n_active; w0 += window) {
    const int w1 = std::min(n_active, w0 + window);
    std::vector<int> rows(w1 - w0);
    for(int r = w0; r < w1; ++r) rows[r - w0] = r;
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) { return row_len(a) > row_len(b); });
    for(int i = 0; i < (int)rows.size(); ++i) {
      o.perm[w0 + i] = rows[i];
      o.len[w0 + i] = row_len(rows[i]);
    }
  }
  const int n_slices = n_slots / kSellRows;
  o.slice_off.assign(n_slices + 1, 0);
  for(int sl = 0; sl < n_slices; ++sl) {
    int width = 0;
    for(int l = 0; l < kSellRows; ++l) width = std::max(width, o.len[sl * kSellRows + l]);
    o.slice_off[sl + 1] = o.slice_off[sl] + (long long)width * kSellRows;
  }
  return o;
}

SellFootprint build_sell_footprint(const Structure& st, const SellLayout& o, int block_slots, int max_fp,
                                   int n_cols) {
  PMCD_CHECK(block_slots > 0 && block_slots % kSellRows == 0, "footprint blocks must be whole slices");
  PMCD_CHECK(max_fp > 0 && max_fp <= 65536, "footprint limit must be 1..65536");
  SellFootprint f;
  const int n_slots = o.n_slots();
  const int n_blocks = (n_slots + block_slots - 1) / block_slots;
  const int n = n_cols < 0 ? st.n_rows : n_cols;
  std::vector<int> mark(n, -1), pos(n, 0);
  std::vector<int> cols;
  f.fp_off.assign(n_blocks + 1, 0);
  f.lcol.assign((size_t)o.slice_off.back(), 0);
  for(int b = 0; b < n_blocks; ++b) {
    const int t0 = b * block_slots, t1 = std::min(n_slots, t0 + block_slots);
    cols.clear();
    for(int t = t0; t < t1; ++t) {
      const int r = o.perm[t];
      if(r < 0) continue;
      for(int q = st.row_ptr[r]; q < st.row_ptr[r + 1]; ++q)
        if(mark[st.col[q]] != b) {
          mark[st.col[q]] = b;
          cols.push_back(st.col[q]);
        }
    }
    if((int)cols.size() > max_fp) return SellFootprint{};
    std::sort(cols.begin(), cols.end());
    for(int i = 0; i < (int)cols.size(); ++i) pos[cols[i]] = i;
    for(int t = t0; t < t1; ++t) {
      const int r = o.perm[t];
      if(r < 0) continue;
      const long long base = o.slice_off[t / kSellRows] + t % kSellRows;
      for(int q = st.row_ptr[r]; q < st.row_ptr[r + 1]; ++q)
        f.lcol[base + (long long)(q - st.row_ptr[r]) * kSellRows] = (unsigned short)pos[st.col[q]];
    }
    f.fp.insert(f.fp.end(), cols.begin(), cols.end());
    f.fp_off[b + 1] = (int)f.fp.size();
    f.fp_max = std::max(f.fp_max, (int)cols.size());
  }
  return f;
}

}  // namespace pmcd
