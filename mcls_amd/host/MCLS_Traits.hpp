// MatrixTraits / VectorTraits: the compile-time plugin API of MCLS, every
// member of the reference's base templates, specialised for the host
// facade's CrsMatrix<Scalar> and Vector.  The managers reach matrices and
// vectors only through these traits, so binding another matrix type
// (Tpetra::CrsMatrix, say) is a pair of specialisations and nothing else.
//
// Mirrors (paths relative to the MCLS source root):
//   src/MCLS_MatrixTraits.hpp:75-406   (unspecialised use fails to compile
//                                      through this_type_is_missing_a_specialization)
//   src/MCLS_VectorTraits.hpp:76-353
//   tpetra/src/MCLS_TpetraCrsMatrixAdapter.hpp:73-498, tpetra/src/MCLS_TpetraVectorAdapter.hpp:62-374
//
// One build extension, MatrixTraits::getLocalCrs: the local rows as plain
// CSR arrays (int64 row offsets, int32 columns, FP64 or FP32 values), which
// the device path uploads in one copy.  The reference's domain build reads
// rows one at a time (getLocalRowCopy); a specialisation that has no CSR
// storage of its own can return localCrsFromRows(), which assembles the
// arrays through getLocalNumRows / getLocalRowCopy.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "MCLS_Core.hpp"

namespace MCLS {

template <class T>
struct UndefinedMatrixTraits {
  static inline void notDefined() { return T::this_type_is_missing_a_specialization(); }
};

// The local rows of a matrix as CSR arrays.  The pointers either point into
// the matrix's own storage (valid while it lives and is not modified) or
// into the owned vectors below.
struct LocalCrs {
  int64_t num_rows = 0;
  int64_t nnz = 0;
  const int64_t* row_ptr = nullptr;
  const int32_t* col = nullptr;
  const void* val = nullptr;
  bool f32 = false;  // values are float (else double)
  std::vector<int64_t> rp_store;
  std::vector<int32_t> col_store;
  std::vector<double> val_store;
  double value(int64_t k) const { return f32 ? (double)((const float*)val)[k] : ((const double*)val)[k]; }
};

// Unspecialised traits: every member a compile error, as in the reference.
template <class Vector, class Matrix>
class MatrixTraits {
 public:
  static void getLocalNumRows(const Matrix&) { UndefinedMatrixTraits<Matrix>::notDefined(); }
  static void getLocalCrs(const Matrix&) { UndefinedMatrixTraits<Matrix>::notDefined(); }
  static void apply(const Matrix&, const Vector&, Vector&) { UndefinedMatrixTraits<Matrix>::notDefined(); }
};

template <class Vector>
class VectorTraits {
 public:
  static void getLocalLength(const Vector&) { UndefinedMatrixTraits<Vector>::notDefined(); }
};

// One process holds every row: the communicator the traits return.
struct SerialComm {
  int getRank() const { return 0; }
  int getSize() const { return 1; }
  void barrier() const {}
};

// LocalCrs assembled through getLocalNumRows / getLocalRowCopy.
template <class Vector, class Matrix>
LocalCrs localCrsFromRows(const Matrix& A) {
  using MT = MatrixTraits<Vector, Matrix>;
  LocalCrs c;
  c.num_rows = (int64_t)MT::getLocalNumRows(A);
  c.rp_store.assign((size_t)c.num_rows + 1, 0);
  std::vector<int64_t> cols;
  std::vector<double> vals;
  for (int64_t i = 0; i < c.num_rows; ++i) {
    size_t m = 0;
    MT::getLocalRowCopy(A, i, cols, vals, m);
    for (size_t k = 0; k < m; ++k) {
      c.col_store.push_back((int32_t)cols[k]);
      c.val_store.push_back(vals[k]);
    }
    c.rp_store[(size_t)i + 1] = (int64_t)c.col_store.size();
  }
  c.nnz = (int64_t)c.col_store.size();
  c.row_ptr = c.rp_store.data();
  c.col = c.col_store.data();
  c.val = c.val_store.data();
  c.f32 = false;
  return c;
}

// true when every stored entry of the CSR lies on the diagonal; d receives it
inline bool crsIsDiagonal(const LocalCrs& c, std::vector<double>* d) {
  d->assign((size_t)c.num_rows, 0.0);
  for (int64_t i = 0; i < c.num_rows; ++i)
    for (int64_t k = c.row_ptr[i]; k < c.row_ptr[i + 1]; ++k) {
      if (c.col[k] != i) return false;
      (*d)[(size_t)i] = c.value(k);
    }
  return true;
}

template <>
class VectorTraits<Vector> {
 public:
  using vector_type = Vector;
  using scalar_type = double;
  using local_ordinal_type = int64_t;
  using global_ordinal_type = int64_t;
  static RCP<Vector> clone(const Vector& v) { return std::make_shared<Vector>(v.size()); }
  static RCP<const SerialComm> getComm(const Vector&) { return std::make_shared<SerialComm>(); }
  // one process owns every row: the given rows are all of them
  static RCP<Vector> createFromRows(const SerialComm&, const std::vector<int64_t>& rows) {
    return std::make_shared<Vector>((int64_t)rows.size());
  }
  static RCP<Vector> deepCopy(const Vector& v) { return std::make_shared<Vector>(v); }
  // column k of a multivector: a Vector is its own only column
  static RCP<Vector> getVectorNonConst(const RCP<Vector>& v, int k) {
    MCLS_INSIST(k == 0, "a Vector has one column");
    return v;
  }
  static int64_t getGlobalLength(const Vector& v) { return v.size(); }
  static int64_t getLocalLength(const Vector& v) { return v.size(); }
  static bool isGlobalRow(const Vector& v, int64_t row) { return row >= 0 && row < v.size(); }
  static bool isLocalRow(const Vector& v, int64_t row) { return row >= 0 && row < v.size(); }
  static int64_t getGlobalRow(const Vector&, int64_t local_row) { return local_row; }
  static int64_t getLocalRow(const Vector&, int64_t global_row) { return global_row; }
  static void replaceGlobalValue(Vector& v, int64_t row, double a) { v[row] = a; }
  static void replaceLocalValue(Vector& v, int64_t row, double a) { v[row] = a; }
  static void sumIntoGlobalValue(Vector& v, int64_t row, double a) { v[row] += a; }
  static void sumIntoLocalValue(Vector& v, int64_t row, double a) { v[row] += a; }
  static void putScalar(Vector& v, double a) { std::fill(v.raw().begin(), v.raw().end(), a); }
  // uniform [-1, 1), as Tpetra::MultiVector::randomize
  static void randomize(Vector& v) {
    std::mt19937_64 g(0x5eedULL + (uint64_t)v.size());
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto& x : v.raw()) x = u(g);
  }
  static const double* view(const Vector& v) { return v.data(); }
  static double* viewNonConst(Vector& v) { return v.data(); }
  static double dot(const Vector& a, const Vector& b) {
    double s = 0.0;
    for (int64_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
  }
  static double norm1(const Vector& v) {
    double s = 0.0;
    for (double x : v.raw()) s += std::fabs(x);
    return s;
  }
  static double norm2(const Vector& v) { return std::sqrt(dot(v, v)); }
  static double normInf(const Vector& v) {
    double s = 0.0;
    for (double x : v.raw()) s = std::max(s, std::fabs(x));
    return s;
  }
  static double meanValue(const Vector& v) {
    double s = 0.0;
    for (double x : v.raw()) s += x;
    return v.size() ? s / (double)v.size() : 0.0;
  }
  // A = |B|
  static void abs(Vector& A, const Vector& B) {
    for (int64_t i = 0; i < A.size(); ++i) A[i] = std::fabs(B[i]);
  }
  static void scale(Vector& v, double a) {
    for (auto& x : v.raw()) x *= a;
  }
  // A = a B
  static void scaleCopy(Vector& A, double a, const Vector& B) {
    for (int64_t i = 0; i < A.size(); ++i) A[i] = a * B[i];
  }
  // A = 1 / B
  static void reciprocal(Vector& A, const Vector& B) {
    for (int64_t i = 0; i < A.size(); ++i) A[i] = 1.0 / B[i];
  }
  // A = alpha A + beta B (tpetra/src/MCLS_TpetraVectorAdapter.hpp update);
  // with alpha = 0 A is not read, as Tpetra's update
  static void update(Vector& A, double alpha, const Vector& B, double beta) {
    for (int64_t i = 0; i < A.size(); ++i) A[i] = (alpha == 0.0 ? 0.0 : alpha * A[i]) + beta * B[i];
  }
  // A = alpha A + beta B + gamma C
  static void update(Vector& A, double alpha, const Vector& B, double beta, const Vector& C, double gamma) {
    for (int64_t i = 0; i < A.size(); ++i)
      A[i] = (alpha == 0.0 ? 0.0 : alpha * A[i]) + beta * B[i] + gamma * C[i];
  }
  // A(i) = alpha A(i) + beta B(i) C(i) (src/MCLS_VectorTraits.hpp:345-352)
  static void elementWiseMultiply(Vector& A, double alpha, const Vector& B, const Vector& C, double beta) {
    for (int64_t i = 0; i < A.size(); ++i) A[i] = alpha * A[i] + beta * B[i] * C[i];
  }
};

template <class Scalar>
class MatrixTraits<Vector, CrsMatrix<Scalar>> {
 public:
  using matrix_type = CrsMatrix<Scalar>;
  using vector_type = Vector;
  using scalar_type = double;
  using local_ordinal_type = int64_t;
  using global_ordinal_type = int64_t;

  // a new matrix with the same rows and no entries
  static RCP<matrix_type> clone(const matrix_type& A) { return std::make_shared<matrix_type>(A.numRows()); }
  // one process owns every row: the export is a copy of those rows
  static RCP<matrix_type> exportFromRows(const matrix_type& A, const std::vector<int64_t>& rows) {
    auto B = std::make_shared<matrix_type>(A.numRows());
    for (int64_t r : rows) {
      std::vector<int64_t> c;
      std::vector<Scalar> v;
      for (int64_t k = A.rowPtr()[(size_t)r]; k < A.rowPtr()[(size_t)r + 1]; ++k) {
        c.push_back(A.colInd()[(size_t)k]);
        v.push_back(A.values()[(size_t)k]);
      }
      B->insertGlobalValues(r, c, v);
    }
    B->fillComplete();
    return B;
  }
  static RCP<Vector> cloneVectorFromMatrixRows(const matrix_type& A) { return std::make_shared<Vector>(A.numRows()); }
  static RCP<Vector> cloneVectorFromMatrixCols(const matrix_type& A) { return std::make_shared<Vector>(A.numRows()); }
  static RCP<Vector> cloneVectorFromMatrixRowsAndCols(const matrix_type& A) {
    return std::make_shared<Vector>(A.numRows());
  }
  // no neighbour ranks: the nearest-neighbour copy is the matrix itself
  // (src/MCLS_AlmostOptimalDomain_impl.hpp:380 calls it for the boundary)
  static RCP<const matrix_type> copyNearestNeighbors(const RCP<const matrix_type>& A, int) { return A; }
  static RCP<const SerialComm> getComm(const matrix_type&) { return std::make_shared<SerialComm>(); }
  static int64_t getGlobalNumRows(const matrix_type& A) { return A.numRows(); }
  static int64_t getLocalNumRows(const matrix_type& A) { return A.numRows(); }
  static int64_t getLocalNumCols(const matrix_type& A) { return A.numRows(); }
  static int64_t getGlobalMaxNumRowEntries(const matrix_type& A) { return A.maxRowEntries(); }
  static int64_t getGlobalRow(const matrix_type&, int64_t local_row) { return local_row; }
  static int64_t getLocalRow(const matrix_type&, int64_t global_row) { return global_row; }
  static int64_t getGlobalCol(const matrix_type&, int64_t local_col) { return local_col; }
  static int64_t getLocalCol(const matrix_type&, int64_t global_col) { return global_col; }
  static void getGlobalRowRanks(const matrix_type&, const std::vector<int64_t>& rows, std::vector<int>& ranks) {
    ranks.assign(rows.size(), 0);
  }
  static void getMyGlobalRows(const matrix_type& A, std::vector<int64_t>& rows) {
    rows.resize((size_t)A.numRows());
    std::iota(rows.begin(), rows.end(), 0);
  }
  static void getMyGlobalCols(const matrix_type& A, std::vector<int64_t>& cols) { getMyGlobalRows(A, cols); }
  static bool isGlobalRow(const matrix_type& A, int64_t r) { return r >= 0 && r < A.numRows(); }
  static bool isLocalRow(const matrix_type& A, int64_t r) { return r >= 0 && r < A.numRows(); }
  static bool isGlobalCol(const matrix_type& A, int64_t c) { return c >= 0 && c < A.numRows(); }
  static bool isLocalCol(const matrix_type& A, int64_t c) { return c >= 0 && c < A.numRows(); }
  static bool isFillComplete(const matrix_type& A) { return A.isFillComplete(); }
  static void getGlobalRowCopy(const matrix_type& A, int64_t row, std::vector<int64_t>& cols,
                               std::vector<double>& vals, size_t& num_entries) {
    getLocalRowCopy(A, row, cols, vals, num_entries);
  }
  static void getLocalRowCopy(const matrix_type& A, int64_t row, std::vector<int64_t>& cols,
                              std::vector<double>& vals, size_t& num_entries) {
    MCLS_INSIST(isLocalRow(A, row), "row out of range");
    const auto& rp = A.rowPtr();
    num_entries = (size_t)(rp[(size_t)row + 1] - rp[(size_t)row]);
    cols.resize(num_entries);
    vals.resize(num_entries);
    for (size_t k = 0; k < num_entries; ++k) {
      cols[k] = A.colInd()[(size_t)rp[(size_t)row] + k];
      vals[k] = (double)A.values()[(size_t)rp[(size_t)row] + k];
    }
  }
  static void getLocalDiagCopy(const matrix_type& A, Vector& d) {
    const auto& rp = A.rowPtr();
    for (int64_t i = 0; i < A.numRows(); ++i) {
      d[i] = 0.0;
      for (int64_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k)
        if (A.colInd()[(size_t)k] == i) d[i] = (double)A.values()[(size_t)k];
    }
  }
  // A(i, j) = x(i) A(i, j)
  static void leftScale(matrix_type& A, const Vector& x) {
    auto& v = A.valuesNonConst();
    for (int64_t i = 0; i < A.numRows(); ++i)
      for (int64_t k = A.rowPtr()[(size_t)i]; k < A.rowPtr()[(size_t)i + 1]; ++k)
        v[(size_t)k] = (Scalar)((double)v[(size_t)k] * x[i]);
  }
  // A(i, j) = A(i, j) x(j)
  static void rightScale(matrix_type& A, const Vector& x) {
    auto& v = A.valuesNonConst();
    for (int64_t k = 0; k < A.nnz(); ++k) v[(size_t)k] = (Scalar)((double)v[(size_t)k] * x[A.colInd()[(size_t)k]]);
  }
  // y = A x on the host (the device path applies A in its residual kernel)
  static void apply(const matrix_type& A, const Vector& x, Vector& y) {
    const auto& rp = A.rowPtr();
    for (int64_t i = 0; i < A.numRows(); ++i) {
      double s = 0.0;
      for (int64_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k)
        s = s + (double)A.values()[(size_t)k] * x[A.colInd()[(size_t)k]];
      y[i] = s;
    }
  }
  // y = A^T x
  static void applyTranspose(const matrix_type& A, const Vector& x, Vector& y) {
    for (int64_t i = 0; i < A.numRows(); ++i) y[i] = 0.0;
    const auto& rp = A.rowPtr();
    for (int64_t i = 0; i < A.numRows(); ++i)
      for (int64_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k)
        y[A.colInd()[(size_t)k]] += (double)A.values()[(size_t)k] * x[i];
  }
  // op(A) op(B), then the row filter of the reference's multiply (keep
  // entries with |v| > threshold x the row's signed maximum;
  // tpetra/src/MCLS_TpetraHelpers.hpp:196-263)
  static RCP<matrix_type> multiply(const RCP<const matrix_type>& A, bool transpose_A, const RCP<const matrix_type>& B,
                                   bool transpose_B, double threshold = 0.0) {
    RCP<const matrix_type> X = transpose_A ? RCP<const matrix_type>(copyTranspose(*A)) : A;
    RCP<const matrix_type> Y = transpose_B ? RCP<const matrix_type>(copyTranspose(*B)) : B;
    const int64_t n = X->numRows();
    std::vector<int64_t> rp((size_t)n + 1, 0);
    std::vector<int32_t> col;
    std::vector<Scalar> val;
    std::vector<double> acc((size_t)n, 0.0);
    std::vector<char> used((size_t)n, 0);
    std::vector<int64_t> cols;
    for (int64_t i = 0; i < n; ++i) {
      cols.clear();
      for (int64_t k = X->rowPtr()[(size_t)i]; k < X->rowPtr()[(size_t)i + 1]; ++k) {
        const int64_t j = X->colInd()[(size_t)k];
        const double xv = (double)X->values()[(size_t)k];
        for (int64_t l = Y->rowPtr()[(size_t)j]; l < Y->rowPtr()[(size_t)j + 1]; ++l) {
          const int64_t c = Y->colInd()[(size_t)l];
          if (!used[(size_t)c]) {
            used[(size_t)c] = 1;
            cols.push_back(c);
          }
          acc[(size_t)c] += xv * (double)Y->values()[(size_t)l];
        }
      }
      std::sort(cols.begin(), cols.end());
      double rmax = -INFINITY;
      for (int64_t c : cols) rmax = std::max(rmax, acc[(size_t)c]);
      for (int64_t c : cols) {
        const double v = acc[(size_t)c];
        if (v != 0.0 && std::fabs(v) > rmax * threshold) {
          col.push_back((int32_t)c);
          val.push_back((Scalar)v);
        }
        acc[(size_t)c] = 0.0;
        used[(size_t)c] = 0;
      }
      rp[(size_t)i + 1] = (int64_t)col.size();
    }
    return std::make_shared<matrix_type>(n, std::move(rp), std::move(col), std::move(val));
  }
  // B = scalar_A op(A) + scalar_B B
  static void add(const RCP<const matrix_type>& A, bool transpose_A, double scalar_A, const RCP<matrix_type>& B,
                  double scalar_B) {
    RCP<const matrix_type> X = transpose_A ? RCP<const matrix_type>(copyTranspose(*A)) : A;
    const int64_t n = X->numRows();
    MCLS_INSIST(B->isFillComplete() && B->numRows() == n, "add: B must be fill-complete with A's rows");
    std::vector<int64_t> rp((size_t)n + 1, 0);
    std::vector<int32_t> col;
    std::vector<Scalar> val;
    for (int64_t i = 0; i < n; ++i) {
      int64_t a = X->rowPtr()[(size_t)i], ae = X->rowPtr()[(size_t)i + 1];
      int64_t b = B->rowPtr()[(size_t)i], be = B->rowPtr()[(size_t)i + 1];
      while (a < ae || b < be) {
        const int64_t ca = a < ae ? X->colInd()[(size_t)a] : INT64_MAX;
        const int64_t cb = b < be ? B->colInd()[(size_t)b] : INT64_MAX;
        double v = 0.0;
        const int64_t c = std::min(ca, cb);
        if (ca == c) v += scalar_A * (double)X->values()[(size_t)a++];
        if (cb == c) v += scalar_B * (double)B->values()[(size_t)b++];
        col.push_back((int32_t)c);
        val.push_back((Scalar)v);
      }
      rp[(size_t)i + 1] = (int64_t)col.size();
    }
    B->setCrs(std::move(rp), std::move(col), std::move(val));
  }
  static RCP<matrix_type> copyTranspose(const matrix_type& A) {
    const int64_t n = A.numRows();
    auto T = std::make_shared<matrix_type>(n);
    const auto& rp = A.rowPtr();
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k)
        T->insertGlobalValues(A.colInd()[(size_t)k], {i}, {A.values()[(size_t)k]});
    T->fillComplete();
    return T;
  }
  // build extension: the matrix's own CSR arrays, no copy
  static LocalCrs getLocalCrs(const matrix_type& A) {
    MCLS_INSIST(A.isFillComplete(), "matrix must be fill-complete");
    LocalCrs c;
    c.num_rows = A.numRows();
    c.nnz = A.nnz();
    c.row_ptr = A.rowPtr().data();
    c.col = A.colInd().data();
    c.val = A.values().data();
    c.f32 = std::is_same<Scalar, float>::value;
    return c;
  }
};

}  // namespace MCLS
