// Assembly on a caller-owned FE space and structure: the Assemble3D /
// Assemble2D entry (include/AssembleRoutines/Assemble3D.h:28-34) with the
// convection-diffusion LocalAssembling (LocalAssembling.h:229-231), adding
// into the caller's matrix entries and right-hand side.
//
// The caller's space is its cells (vertex coordinates in the element's local
// vertex order), its local-to-global map (GlobalNumbers with BeginIndex) and
// its active bound (inner and non-Dirichlet DOFs first, Dirichlet DOFs last,
// include/FE/FESpace.h:71-92); its structure is a CSR pattern holding every
// coupling of two DOFs of one cell. The local matrices and rhs come from the
// same kernels as a system's assembly, and each matrix entry and rhs entry
// receives its cell contributions in increasing cell order, the order of the
// reference's cell loop (Assemble3D.C:183-549).
#pragma once
#include "system.hpp"

namespace pmcd {

class CallerAssembler {
 public:
  // n_hanging > 0: the hanging DOFs n_active - n_hanging .. n_active - 1,
  // hanging[4i] = the hanging DOF, hanging[4i+1..4i+3] its coupling DOFs (-1
  // beyond their count), weights[3i..3i+2] their weights
  CallerAssembler(Shape shape, int order, int n_cells, int n_dof, int n_active, const double* vertices,
                  const int* global_numbers, const int* begin_index, const int* row_ptr, const int* col,
                  const ProblemDesc& forms, hipStream_t stream, int n_hanging = 0, const int* hanging = nullptr,
                  const double* weights = nullptr);
  // local matrices and rhs of every cell summed into the device copies (CSR
  // values with identity Dirichlet rows; rhs with the Dirichlet values on the
  // tail); dirichlet_values: n_dof - n_active values (host or device memory,
  // null: zero)
  void assemble(const double* dirichlet_values, bool values_on_device);
  // the reference's update of the caller's arrays (Assemble3D.C:258-300,
  // 346-360 and its boundary loop): active rows get the sums added, each
  // Dirichlet row's diagonal is set to 1 (the rest of the row is left as it
  // is), the rhs gets the sums added on active DOFs and the Dirichlet values
  // set on the others; with hanging DOFs the coupled sums are added and the
  // hanging rows and rhs entries are set to the constraints and 0
  // (Assemble2D.C:3350-3427 on the caller's reset arrays)
  void add_into_host(double* entries, double* rhs);
  void add_into_device(double* entries, double* rhs, hipStream_t caller_stream);
  int n_dof() const { return L_.space.n_dof; }
  int64_t nnz() const { return L_.st.nnz; }
  hipStream_t stream() const { return stream_; }

 private:
  ProblemDesc forms_;
  hipStream_t stream_;
  Shape shape_;
  int n_cells_;
  Level L_;
  QuadRule rule_;
  int ncomp_ = 0;
  DevBuf<double> ref_, qw_, qp_, tensor_, local_A_, local_rhs_;
};

}  // namespace pmcd
