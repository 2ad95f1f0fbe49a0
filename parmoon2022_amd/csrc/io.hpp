// Vector and matrix files in the reference's formats (see io.cpp).
#pragma once
#include <string>

namespace pmcd {

// "<n>\n" followed by n raw doubles (BlockVector::write_to_stream)
void write_vector_stream(const std::string& path, long long n, const double* v);
// reads that format into v; the length must be n_expected
long long read_vector_stream(const std::string& path, long long n_expected, double* v);
// MatrixMarket array format (BlockVector::write)
void write_vector_market(const std::string& path, long long n, const double* v);
// MatrixMarket coordinate format, 1-based (TMatrix::write)
void write_matrix_market(const std::string& path, int n_rows, int n_cols, const int* rp, const int* col,
                         const double* val);

}  // namespace pmcd
