// On-disk formats of vectors and matrices, as the reference writes them:
// BlockVector::write_to_stream / read_from_stream (src/Matrix/BlockVector.C:
// 1064-1110: the length on a line of its own, then the raw doubles; used by
// CheckpointIO for restarts), BlockVector::write (:598-615, MatrixMarket
// array format) and TMatrix::write (src/Matrix/Matrix.C:97-130, MatrixMarket
// coordinate format, 1-based indices). Host arrays only.
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include "common.hpp"
#include "io.hpp"

namespace pmcd {

void write_vector_stream(const std::string& path, long long n, const double* v) {
  std::ofstream os(path, std::ios::binary);
  PMCD_CHECK(os.good(), "cannot open file '", path, "' to save data");
  os << n << '\n';
  os.write(reinterpret_cast<const char*>(v), (std::streamsize)(sizeof(double) * n));
  PMCD_CHECK(os.good(), "error writing ", path);
}

long long read_vector_stream(const std::string& path, long long n_expected, double* v) {
  std::ifstream is(path, std::ios::binary);
  PMCD_CHECK(is.good(), "cannot open file '", path, "' to read data");
  std::string line;
  std::getline(is, line);
  std::istringstream parser(line);
  long long n = 0;
  const bool got = static_cast<bool>(parser >> n);
  parser >> std::ws;
  // the first line holds the number of entries and nothing else
  PMCD_CHECK(got && !parser.good() && n >= 0,
             "formatting error, the first line in the file should contain only a number indicating the number "
             "of entries to be read");
  PMCD_CHECK(n == n_expected, "unexpected size to be read. Expected: ", n_expected, "\tin file: ", n);
  is.read(reinterpret_cast<char*>(v), (std::streamsize)(sizeof(double) * n));
  PMCD_CHECK(is.gcount() == (std::streamsize)(sizeof(double) * n), "file ", path, " ends before ", n,
             " entries");
  return n;
}

void write_vector_market(const std::string& path, long long n, const double* v) {
  std::ofstream os(path);
  PMCD_CHECK(os.good(), "cannot open file '", path, "'");
  os << "%%MatrixMarket matrix array real general \n";
  os << n << "\t" << 1 << "\t" << "\n";
  for(long long i = 0; i < n; ++i) os << std::setprecision(16) << v[i] << "\n";
  PMCD_CHECK(os.good(), "error writing ", path);
}

void write_matrix_market(const std::string& path, int n_rows, int n_cols, const int* rp, const int* col,
                         const double* val) {
  std::ofstream os(path);
  PMCD_CHECK(os.good(), "cannot open file '", path, "'");
  os << "%%MatrixMarket matrix coordinate real general \n";
  os << n_rows << "\t" << n_cols << "\t" << rp[n_rows] << "\n";
  for(int i = 0; i < n_rows; ++i)
    for(int k = rp[i]; k < rp[i + 1]; ++k)
      os << std::setprecision(16) << i + 1 << "\t" << col[k] + 1 << "\t" << val[k] << "\n";
  PMCD_CHECK(os.good(), "error writing ", path);
}

}  // namespace pmcd
