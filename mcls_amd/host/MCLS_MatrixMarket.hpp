// MatrixMarket exchange-format input and output for the facade's CrsMatrix
// and Vector (R. F. Boisvert, R. Pozo, K. Remington, "The Matrix Market
// Exchange Formats: Initial Design", NISTIR 5935, 1996).
//
// Matrices: "%%MatrixMarket matrix coordinate <field> <symmetry>" with
// field real | integer | pattern (pattern entries are 1) and symmetry
// general | symmetric | skew-symmetric; indices are 1-based; the mirrored
// entry of a symmetric (negated for skew-symmetric) off-diagonal entry is
// added; repeated entries are summed (CrsMatrix::fillComplete).  The solver
// needs a square matrix.  Vectors: "%%MatrixMarket matrix array real
// general" with one column, or a coordinate file with one column.
//
// The reference reads its example systems through Trilinos' own readers
// (EpetraExt / Tpetra MatrixMarket); this header is the build's loader.
#pragma once

#include <cctype>
#include <cstdint>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "MCLS_Core.hpp"

namespace MCLS {

struct MatrixMarketInfo {
  std::string format;    // "coordinate" or "array"
  std::string field;     // "real", "integer", "pattern"
  std::string symmetry;  // "general", "symmetric", "skew-symmetric"
  int64_t rows = 0, cols = 0, entries = 0;  // as declared in the size line
};

namespace detail {

inline std::string mm_lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

// Reads the banner and the size line; leaves `in` at the first data line.
inline MatrixMarketInfo mm_header(std::istream& in, const std::string& path) {
  std::string line;
  MCLS_INSIST((bool)std::getline(in, line), "empty MatrixMarket file: " + path);
  std::istringstream banner(line);
  std::string tag, object;
  MatrixMarketInfo info;
  banner >> tag >> object >> info.format >> info.field >> info.symmetry;
  MCLS_INSIST(tag == "%%MatrixMarket", "not a MatrixMarket file (missing banner): " + path);
  object = mm_lower(object);
  info.format = mm_lower(info.format);
  info.field = mm_lower(info.field);
  info.symmetry = mm_lower(info.symmetry);
  MCLS_INSIST(object == "matrix", "MatrixMarket object '" + object + "' is not supported");
  MCLS_INSIST(info.format == "coordinate" || info.format == "array",
              "MatrixMarket format '" + info.format + "' is not supported");
  MCLS_INSIST(info.field == "real" || info.field == "integer" || info.field == "pattern" ||
                  info.field == "double",
              "MatrixMarket field '" + info.field + "' is not supported (real, integer, pattern)");
  MCLS_INSIST(info.symmetry == "general" || info.symmetry == "symmetric" || info.symmetry == "skew-symmetric",
              "MatrixMarket symmetry '" + info.symmetry + "' is not supported");
  MCLS_INSIST(!(info.format == "array" && info.field == "pattern"), "array format cannot be pattern");
  while (std::getline(in, line)) {
    size_t p = line.find_first_not_of(" \t\r");
    if (p == std::string::npos || line[p] == '%') continue;
    std::istringstream sz(line);
    if (info.format == "coordinate") {
      MCLS_INSIST((bool)(sz >> info.rows >> info.cols >> info.entries), "bad MatrixMarket size line: " + line);
    } else {
      MCLS_INSIST((bool)(sz >> info.rows >> info.cols), "bad MatrixMarket size line: " + line);
      info.entries = info.rows * info.cols;
    }
    MCLS_INSIST(info.rows >= 0 && info.cols >= 0 && info.entries >= 0, "negative MatrixMarket sizes");
    return info;
  }
  MCLS_INSIST(false, "MatrixMarket file without a size line: " + path);
  return info;
}

// Next non-comment, non-blank data line.
inline bool mm_next(std::istream& in, std::string& line) {
  while (std::getline(in, line)) {
    size_t p = line.find_first_not_of(" \t\r");
    if (p != std::string::npos && line[p] != '%') return true;
  }
  return false;
}

}  // namespace detail

// Square sparse matrix from a coordinate MatrixMarket file.
template <class Scalar = double>
RCP<CrsMatrix<Scalar>> readMatrixMarket(const std::string& path, MatrixMarketInfo* info_out = nullptr) {
  std::ifstream in(path);
  MCLS_INSIST(in.good(), "cannot open " + path);
  MatrixMarketInfo info = detail::mm_header(in, path);
  MCLS_INSIST(info.format == "coordinate", "sparse matrices must use the coordinate format: " + path);
  MCLS_INSIST(info.rows == info.cols, "the solver needs a square matrix: " + path);
  const int64_t n = info.rows;
  auto A = std::make_shared<CrsMatrix<Scalar>>(n);
  const bool pattern = info.field == "pattern";
  const bool sym = info.symmetry == "symmetric", skew = info.symmetry == "skew-symmetric";
  std::string line;
  for (int64_t e = 0; e < info.entries; ++e) {
    MCLS_INSIST(detail::mm_next(in, line), "MatrixMarket file ends after " + std::to_string(e) + " of " +
                                               std::to_string(info.entries) + " entries: " + path);
    std::istringstream ls(line);
    int64_t i = 0, j = 0;
    double v = 1.0;
    MCLS_INSIST((bool)(ls >> i >> j), "bad MatrixMarket entry: " + line);
    if (!pattern) MCLS_INSIST((bool)(ls >> v), "bad MatrixMarket value: " + line);
    MCLS_INSIST(i >= 1 && i <= n && j >= 1 && j <= n, "MatrixMarket index out of range: " + line);
    MCLS_INSIST(!(skew && i == j), "skew-symmetric matrix with a diagonal entry: " + line);
    A->insertGlobalValues(i - 1, {j - 1}, {(Scalar)v});
    if ((sym || skew) && i != j) A->insertGlobalValues(j - 1, {i - 1}, {(Scalar)(skew ? -v : v)});
  }
  A->fillComplete();
  if (info_out) *info_out = info;
  return A;
}

// Dense vector from a one-column MatrixMarket file (array or coordinate).
inline RCP<Vector> readMatrixMarketVector(const std::string& path) {
  std::ifstream in(path);
  MCLS_INSIST(in.good(), "cannot open " + path);
  MatrixMarketInfo info = detail::mm_header(in, path);
  MCLS_INSIST(info.cols == 1, "a vector file has one column: " + path);
  MCLS_INSIST(info.field != "pattern", "a vector file needs values: " + path);
  auto v = std::make_shared<Vector>(info.rows);
  std::string line;
  for (int64_t e = 0; e < info.entries; ++e) {
    MCLS_INSIST(detail::mm_next(in, line), "MatrixMarket vector ends early: " + path);
    std::istringstream ls(line);
    if (info.format == "array") {
      double x = 0.0;
      MCLS_INSIST((bool)(ls >> x), "bad MatrixMarket value: " + line);
      (*v)[e] = x;
    } else {
      int64_t i = 0, j = 0;
      double x = 0.0;
      MCLS_INSIST((bool)(ls >> i >> j >> x), "bad MatrixMarket entry: " + line);
      MCLS_INSIST(i >= 1 && i <= info.rows && j == 1, "MatrixMarket index out of range: " + line);
      (*v)[i - 1] += x;
    }
  }
  return v;
}

// One-column "array real general" file; values at 17 significant digits,
// so a read gives back the same doubles.
inline void writeMatrixMarketVector(const std::string& path, const Vector& v) {
  std::ofstream out(path);
  MCLS_INSIST(out.good(), "cannot write " + path);
  out << "%%MatrixMarket matrix array real general\n" << v.size() << " 1\n";
  out.precision(17);
  for (int64_t i = 0; i < v.size(); ++i) out << v[i] << "\n";
  MCLS_INSIST(out.good(), "write failed: " + path);
}

}  // namespace MCLS
