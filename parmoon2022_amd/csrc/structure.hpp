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
};

Structure build_structure(const FESpace& sp, int n_cells);

}  // namespace pmcd
