// Sparsity pattern of a square FE matrix and the maps the device assembly
// uses to write into it.
//
// Reference: src/Matrix/Structure.C:226-556 (full pattern: every pair of
// DOFs sharing a cell, Dirichlet rows included), :587-618 (columns ascending
// in each row). The reference adds each cell's local matrix into the global
// rows by linear search (src/AssembleRoutines/Assemble3D.C:258-292), cell
// after cell, so every CSR entry receives its contributions in increasing
// cell order. Here the device works row by row: for every DOF the list of
// (cell, local row) pairs containing it, in increasing cell order, plus a
// precomputed slot map giving, for each such pair, the position of each of
// the cell's DOFs inside the row. Summing a row's contributions in list order
// reproduces the reference's order without atomics.
#pragma once
#include "common.hpp"
#include "fespace.hpp"
#include "kernels.hpp"

namespace pmcd {

struct Structure {
  int n_rows = 0;
  int64_t nnz = 0;
  int max_row = 0;
  std::vector<int> row_ptr;  // n_rows + 1 (int32: nnz < 2^31 required)
  std::vector<int> col;      // nnz, ascending per row
  std::vector<int> diag;     // CSR position of the diagonal entry of each row
  // DOF -> contributions: entries contrib_ptr[r]..contrib_ptr[r+1] of contrib
  // hold c * n_local + m with dof(c, m) == r, c increasing
  std::vector<int> contrib_ptr;
  std::vector<int> contrib;
  // slot[e * n_local + k] = position (relative to row_ptr[r]) of dof(c, k) in
  // row r, for contribution e = (c, m) of row r
  std::vector<uint8_t> slot;
  // (r, c) in the pattern exactly when (c, r) is: true for the cell pattern,
  // false after the hanging columns are added (or for a caller's pattern
  // that is not)
  bool symmetric = true;
};

Structure build_structure(const FESpace& sp, int n_cells);
// true when every row's columns ascend and (r, c) is in the pattern exactly
// when (c, r) is
bool pattern_symmetric(int n, const int* row_ptr, const int* col);
// a space with hanging DOFs: `base` (the cell pattern) with the columns of
// each hanging row added to the rows of its active coupling DOFs
Structure hanging_structure(const FESpace& sp, int n_cells, const Structure& base);
// the maps of a caller's pattern (row_ptr, col: n_dof rows, any column order)
// for the space's cells: diagonal positions, (cell, local row) lists and slot
// maps; every coupling of two DOFs of one cell must be in the pattern
Structure structure_from_pattern(const FESpace& sp, int n_cells, const int* row_ptr, const int* col);

// Row classes for a successive over-relaxation sweep on the device. A forward
// sweep visits the classes in order, a backward sweep in reverse; rows of one
// class are never adjacent, so a class updates in parallel. Adjacency is
// symmetric: r and c are adjacent when either row holds the other's column
// (a pattern with hanging columns is not symmetric: row r may read x_c while
// row c does not read x_r, and c must still not be updated beside r).
//
// exact: classes are the levels of the row-order dependency graph
// (class(r) = 1 + max class of the adjacent rows c < r). Every row then reads
// new values for its neighbours c < r and old ones for c > r, the values the
// serial sweep of the reference uses (src/Matrix/Matrix.C:1370-1395), so the
// result is bit-identical. Otherwise (more than max_exact_classes levels,
// e.g. Q2 in 3D) greedy colouring in row order gives a multicolour SOR: the
// same per-row arithmetic in a different row order.
struct SweepOrder {
  std::vector<int> rows;       // rows grouped by class, ascending inside a class
  std::vector<int> class_ptr;  // class k = rows[class_ptr[k] .. class_ptr[k+1])
  bool exact = false;
};

SweepOrder build_sweep_order(const Structure& st, int max_exact_classes);

// Sliced ELL layout of the active rows (r < n_active) of a CSR pattern:
// slices of kSellRows rows, one wavefront per slice; inside a slice entry k of
// the row in lane l sits at slice_off[s] + k * kSellRows + l, so that the
// wavefront reads kSellRows consecutive values per load. Rows are sorted by
// length (longest first, stable) inside windows of `window` rows so that a
// slice holds rows of about equal length. Each row keeps its entries in CSR
// (ascending column) order.
struct SellLayout {
  std::vector<int> perm;             // slot -> row, -1 for padding slots
  std::vector<int> len;              // slot -> row length (0 for padding)
  std::vector<long long> slice_off;  // first entry of each slice, n_slices + 1
  int n_slots() const { return (int)perm.size(); }
};
SellLayout build_sell(const Structure& st, int n_active, int window);

// Column footprint of each block of `block_slots` slots of a sliced ELL
// layout: the distinct columns of the block's rows, ascending (fp[fp_off[b]
// .. fp_off[b+1]]), and for every layout entry its column's position in its
// block's list (lcol, same positions as the layout's values; 0 on padding).
// A kernel then gathers each distinct x entry of a block once into LDS and
// reads it there by the 16-bit position. Empty (fp_off empty) when a block
// has more than max_fp distinct columns.
struct SellFootprint {
  std::vector<unsigned short> lcol;
  std::vector<int> fp_off;
  std::vector<int> fp;
  int fp_max = 0;
};
// n_cols: columns of the matrix (-1: st.n_rows, a square matrix)
SellFootprint build_sell_footprint(const Structure& st, const SellLayout& o, int block_slots, int max_fp,
                                   int n_cols = -1);

}  // namespace pmcd
