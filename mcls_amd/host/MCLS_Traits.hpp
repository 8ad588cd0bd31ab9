// MatrixTraits / VectorTraits: the compile-time plugin API of MCLS, with the
// members the adjoint MC / MCSA path uses, specialised for the host facade's
// CrsMatrix<Scalar> and Vector.
//
// Mirrors (paths relative to the MCLS source root):
//   src/MCLS_MatrixTraits.hpp:75-406   (unspecialised use fails to compile
//                                      through this_type_is_missing_a_specialization)
//   src/MCLS_VectorTraits.hpp:76-353
//   tpetra/src/MCLS_TpetraCrsMatrixAdapter.hpp:73-498, tpetra/src/MCLS_TpetraVectorAdapter.hpp:62-374
//
// A Tpetra specialisation would forward its local CSR arrays to the same C
// ABI; it is not compiled here because Trilinos is not installed.
#pragma once

#include <cmath>

#include "MCLS_Core.hpp"

namespace MCLS {

template <class T>
struct UndefinedMatrixTraits {
  static inline void notDefined() { return T::this_type_is_missing_a_specialization(); }
};

template <class Vector, class Matrix>
class MatrixTraits {
 public:
  static void getLocalNumRows(const Matrix&) { UndefinedMatrixTraits<Matrix>::notDefined(); }
};

template <class Vector>
class VectorTraits {
 public:
  static void getLocalLength(const Vector&) { UndefinedMatrixTraits<Vector>::notDefined(); }
};

template <>
class VectorTraits<Vector> {
 public:
  using vector_type = Vector;
  using scalar_type = double;
  static RCP<Vector> clone(const Vector& v) { return std::make_shared<Vector>(v.size()); }
  static RCP<Vector> deepCopy(const Vector& v) { return std::make_shared<Vector>(v); }
  static int64_t getLocalLength(const Vector& v) { return v.size(); }
  static int64_t getGlobalLength(const Vector& v) { return v.size(); }
  static const double* view(const Vector& v) { return v.data(); }
  static double* viewNonConst(Vector& v) { return v.data(); }
  static void putScalar(Vector& v, double a) { std::fill(v.raw().begin(), v.raw().end(), a); }
  static void scale(Vector& v, double a) {
    for (auto& x : v.raw()) x *= a;
  }
  // A = alpha A + beta B (tpetra/src/MCLS_TpetraVectorAdapter.hpp update)
  static void update(Vector& A, double alpha, const Vector& B, double beta) {
    for (int64_t i = 0; i < A.size(); ++i) A[i] = alpha * A[i] + beta * B[i];
  }
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
};

template <class Scalar>
class MatrixTraits<Vector, CrsMatrix<Scalar>> {
 public:
  using matrix_type = CrsMatrix<Scalar>;
  using vector_type = Vector;
  static RCP<Vector> cloneVectorFromMatrixRows(const matrix_type& A) {
    return std::make_shared<Vector>(A.numRows());
  }
  static int64_t getLocalNumRows(const matrix_type& A) { return A.numRows(); }
  static int64_t getGlobalNumRows(const matrix_type& A) { return A.numRows(); }
  static int64_t getGlobalMaxNumRowEntries(const matrix_type& A) { return A.maxRowEntries(); }
  static bool isFillComplete(const matrix_type& A) { return A.isFillComplete(); }
  static void getLocalRowCopy(const matrix_type& A, int64_t row, std::vector<int64_t>& cols,
                              std::vector<double>& vals, size_t& num_entries) {
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
};

}  // namespace MCLS
