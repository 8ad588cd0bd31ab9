// MatrixTraits / VectorTraits for Tpetra::CrsMatrix<double, LO, GO, Node> and
// Tpetra::Vector<double, LO, GO, Node>: the binding a Trilinos build adds to
// the facade.  Nothing else in the library changes; the managers and
// LinearProblem reach matrices and vectors only through these traits.
//
// Include Tpetra_CrsMatrix.hpp and Tpetra_Vector.hpp before this header (it
// includes no Trilinos header itself, so it can be compiled against any type
// with Tpetra's member API).  Written against the Tpetra API the reference
// builds with (Trilinos 12: getNodeNumRows, getNodeNumElements,
// getNodeElementList, getLocalMatrix with the Kokkos CRS arrays
// graph.row_map / graph.entries / values); Trilinos 14 renamed the Node
// members to Local ones.
//
// One process holds every row: the device path is one GPU, so the maps must
// be serial and contiguous from global index 0 (checked).  The vectors' and
// the local matrix's arrays must be host-accessible (a host Node), since
// view / viewNonConst / getLocalCrs hand raw pointers to the device uploads.
//
// Mirrors (paths relative to the MCLS source root):
//   tpetra/src/MCLS_TpetraCrsMatrixAdapter.hpp:73-498
//   tpetra/src/MCLS_TpetraVectorAdapter.hpp:62-374
// The structural products (multiply, copyTranspose) go through the facade's
// CSR routines on the local rows, which keep the reference's semantics,
// including multiply's row filter (tpetra/src/MCLS_TpetraHelpers.hpp:196-263),
// and come back as Tpetra matrices on the operand's maps; add is left out, as
// the reference's adapter leaves it.
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>

#include "MCLS_Traits.hpp"

namespace MCLS {

namespace tpetra_detail {

// The one-process condition of this binding.
template <class MapRcp>
void requireSerialMap(const MapRcp& map, const char* what) {
  MCLS_INSIST(map->getComm()->getSize() == 1, std::string(what) + ": one process must hold every row");
  MCLS_INSIST(map->isContiguous() && map->getMinGlobalIndex() == 0,
              std::string(what) + ": the row map must be contiguous from global index 0");
}

}  // namespace tpetra_detail

template <class S, class LO, class GO, class NT>
class VectorTraits<Tpetra::Vector<S, LO, GO, NT>> {
 public:
  using vector_type = Tpetra::Vector<S, LO, GO, NT>;
  using scalar_type = S;
  using local_ordinal_type = LO;
  using global_ordinal_type = GO;
  using V = vector_type;

  static RCP<V> clone(const V& v) { return std::make_shared<V>(v.getMap()); }
  static RCP<V> deepCopy(const V& v) { return std::make_shared<V>(v, Teuchos::Copy); }
  static RCP<V> getVectorNonConst(const RCP<V>& v, int k) {
    MCLS_INSIST(k == 0, "a Vector has one column");
    return v;
  }
  static auto getComm(const V& v) -> decltype(v.getMap()->getComm()) { return v.getMap()->getComm(); }
  static int64_t getGlobalLength(const V& v) { return (int64_t)v.getGlobalLength(); }
  static int64_t getLocalLength(const V& v) { return (int64_t)v.getLocalLength(); }
  static bool isGlobalRow(const V& v, GO row) { return v.getMap()->isNodeGlobalElement(row); }
  static bool isLocalRow(const V& v, LO row) { return v.getMap()->isNodeLocalElement(row); }
  static GO getGlobalRow(const V& v, LO local_row) { return v.getMap()->getGlobalElement(local_row); }
  static LO getLocalRow(const V& v, GO global_row) { return v.getMap()->getLocalElement(global_row); }
  static void replaceGlobalValue(V& v, GO row, S a) {
    MCLS_INSIST(isGlobalRow(v, row), "row out of range");
    v.replaceGlobalValue(row, a);
  }
  static void replaceLocalValue(V& v, LO row, S a) {
    MCLS_INSIST(isLocalRow(v, row), "row out of range");
    v.replaceLocalValue(row, a);
  }
  static void sumIntoGlobalValue(V& v, GO row, S a) {
    MCLS_INSIST(isGlobalRow(v, row), "row out of range");
    v.sumIntoGlobalValue(row, a);
  }
  static void sumIntoLocalValue(V& v, LO row, S a) {
    MCLS_INSIST(isLocalRow(v, row), "row out of range");
    v.sumIntoLocalValue(row, a);
  }
  static void putScalar(V& v, S a) { v.putScalar(a); }
  static void randomize(V& v) { v.randomize(); }
  // the local entries as a raw host array (the device uploads read it)
  static const S* view(const V& v) { return v.getData().getRawPtr(); }
  static S* viewNonConst(V& v) { return v.getDataNonConst().getRawPtr(); }
  static S dot(const V& a, const V& b) { return b.dot(a); }
  static S norm1(const V& v) { return v.norm1(); }
  static S norm2(const V& v) { return v.norm2(); }
  static S normInf(const V& v) { return v.normInf(); }
  static S meanValue(const V& v) { return v.meanValue(); }
  static void abs(V& A, const V& B) { A.abs(B); }                       // A = |B|
  static void scale(V& v, S a) { v.scale(a); }                          // v = a v
  static void scaleCopy(V& A, S a, const V& B) { A.scale(a, B); }       // A = a B
  static void reciprocal(V& A, const V& B) { A.reciprocal(B); }         // A = 1 / B
  // A = alpha A + beta B
  static void update(V& A, S alpha, const V& B, S beta) { A.update(beta, B, alpha); }
  // A = alpha A + beta B + gamma C
  static void update(V& A, S alpha, const V& B, S beta, const V& C, S gamma) { A.update(beta, B, gamma, C, alpha); }
  // A(i) = alpha A(i) + beta B(i) C(i)
  static void elementWiseMultiply(V& A, S alpha, const V& B, const V& C, S beta) {
    A.elementWiseMultiply(beta, B, C, alpha);
  }
};

template <class S, class LO, class GO, class NT>
class MatrixTraits<Tpetra::Vector<S, LO, GO, NT>, Tpetra::CrsMatrix<S, LO, GO, NT>> {
 public:
  using matrix_type = Tpetra::CrsMatrix<S, LO, GO, NT>;
  using vector_type = Tpetra::Vector<S, LO, GO, NT>;
  using scalar_type = S;
  using local_ordinal_type = LO;
  using global_ordinal_type = GO;
  using M = matrix_type;
  using V = vector_type;

  // a new matrix on the same rows and no entries
  static RCP<M> clone(const M& A) { return std::make_shared<M>(A.getRowMap(), (size_t)0); }
  // one process holds every row: the export is a copy of the given rows
  static RCP<M> exportFromRows(const M& A, const std::vector<int64_t>& rows) {
    auto B = std::make_shared<M>(A.getRowMap(), (size_t)A.getGlobalMaxNumRowEntries());
    std::vector<int64_t> c;
    std::vector<double> v;
    for (int64_t r : rows) {
      size_t m = 0;
      getGlobalRowCopy(A, r, c, v, m);
      insertRow(*B, (GO)r, c, v, m);
    }
    B->fillComplete(A.getDomainMap(), A.getRangeMap());
    return B;
  }
  static RCP<V> cloneVectorFromMatrixRows(const M& A) { return std::make_shared<V>(A.getRowMap()); }
  static RCP<V> cloneVectorFromMatrixCols(const M& A) {
    MCLS_INSIST(A.isFillComplete(), "matrix must be fill-complete");
    return std::make_shared<V>(A.getColMap());
  }
  static RCP<V> cloneVectorFromMatrixDomain(const M& A) { return std::make_shared<V>(A.getDomainMap()); }
  static RCP<V> cloneVectorFromMatrixRange(const M& A) { return std::make_shared<V>(A.getRangeMap()); }
  // no neighbour ranks: the nearest-neighbour copy is the matrix itself
  static RCP<const M> copyNearestNeighbors(const RCP<const M>& A, int) { return A; }
  static auto getComm(const M& A) -> decltype(A.getComm()) { return A.getComm(); }
  static int64_t getGlobalNumRows(const M& A) { return (int64_t)A.getGlobalNumRows(); }
  static int64_t getLocalNumRows(const M& A) { return (int64_t)A.getRowMap()->getNodeNumElements(); }
  static int64_t getLocalNumCols(const M& A) { return (int64_t)A.getColMap()->getNodeNumElements(); }
  static int64_t getGlobalMaxNumRowEntries(const M& A) { return (int64_t)A.getGlobalMaxNumRowEntries(); }
  static GO getGlobalRow(const M& A, LO local_row) {
    MCLS_INSIST(A.getRowMap()->isNodeLocalElement(local_row), "row out of range");
    return A.getRowMap()->getGlobalElement(local_row);
  }
  static LO getLocalRow(const M& A, GO global_row) {
    MCLS_INSIST(A.getRowMap()->isNodeGlobalElement(global_row), "row out of range");
    return A.getRowMap()->getLocalElement(global_row);
  }
  static GO getGlobalCol(const M& A, LO local_col) {
    MCLS_INSIST(A.isFillComplete() && A.getColMap()->isNodeLocalElement(local_col), "column out of range");
    return A.getColMap()->getGlobalElement(local_col);
  }
  static LO getLocalCol(const M& A, GO global_col) {
    MCLS_INSIST(A.isFillComplete() && A.getColMap()->isNodeGlobalElement(global_col), "column out of range");
    return A.getColMap()->getLocalElement(global_col);
  }
  // one process: every row lives on rank 0
  static void getGlobalRowRanks(const M&, const std::vector<int64_t>& rows, std::vector<int>& ranks) {
    ranks.assign(rows.size(), 0);
  }
  static void getMyGlobalRows(const M& A, std::vector<int64_t>& rows) {
    auto l = A.getRowMap()->getNodeElementList();
    rows.assign(l.begin(), l.end());
  }
  static void getMyGlobalCols(const M& A, std::vector<int64_t>& cols) {
    MCLS_INSIST(A.isFillComplete(), "matrix must be fill-complete");
    auto l = A.getColMap()->getNodeElementList();
    cols.assign(l.begin(), l.end());
  }
  static bool isGlobalRow(const M& A, GO r) { return A.getRowMap()->isNodeGlobalElement(r); }
  static bool isLocalRow(const M& A, LO r) { return A.getRowMap()->isNodeLocalElement(r); }
  static bool isGlobalCol(const M& A, GO c) { return A.isFillComplete() && A.getColMap()->isNodeGlobalElement(c); }
  static bool isLocalCol(const M& A, LO c) { return A.isFillComplete() && A.getColMap()->isNodeLocalElement(c); }
  static bool isFillComplete(const M& A) { return A.isFillComplete(); }
  static void getGlobalRowCopy(const M& A, int64_t row, std::vector<int64_t>& cols, std::vector<double>& vals,
                               size_t& num_entries) {
    MCLS_INSIST(A.isFillComplete() && A.getRowMap()->isNodeGlobalElement((GO)row), "row out of range");
    const size_t m = A.getNumEntriesInGlobalRow((GO)row);
    std::vector<GO> c(m);
    std::vector<S> v(m);
    A.getGlobalRowCopy((GO)row, Teuchos::ArrayView<GO>(c.data(), (int)m), Teuchos::ArrayView<S>(v.data(), (int)m),
                       num_entries);
    cols.assign(c.begin(), c.begin() + num_entries);
    vals.assign(v.begin(), v.begin() + num_entries);
  }
  static void getLocalRowCopy(const M& A, int64_t row, std::vector<int64_t>& cols, std::vector<double>& vals,
                              size_t& num_entries) {
    MCLS_INSIST(A.isFillComplete() && A.getRowMap()->isNodeLocalElement((LO)row), "row out of range");
    const size_t m = A.getNumEntriesInLocalRow((LO)row);
    std::vector<LO> c(m);
    std::vector<S> v(m);
    A.getLocalRowCopy((LO)row, Teuchos::ArrayView<LO>(c.data(), (int)m), Teuchos::ArrayView<S>(v.data(), (int)m),
                      num_entries);
    cols.assign(c.begin(), c.begin() + num_entries);
    vals.assign(v.begin(), v.begin() + num_entries);
  }
  static void getLocalDiagCopy(const M& A, V& d) { A.getLocalDiagCopy(d); }
  static void leftScale(M& A, const V& x) { A.leftScale(x); }    // A(i, j) = x(i) A(i, j)
  static void rightScale(M& A, const V& x) { A.rightScale(x); }  // A(i, j) = A(i, j) x(j)
  static void apply(const M& A, const V& x, V& y) { A.apply(x, y); }
  static void applyTranspose(const M& A, const V& x, V& y) {
    MCLS_INSIST(A.hasTransposeApply(), "matrix has no transpose apply");
    A.apply(x, y, Teuchos::TRANS);
  }
  // op(A) op(B), then multiply's row filter (keep |v| > threshold x the row's
  // signed maximum), on the local rows
  static RCP<M> multiply(const RCP<const M>& A, bool transpose_A, const RCP<const M>& B, bool transpose_B,
                         double threshold = 0.0) {
    using FT = MatrixTraits<Vector, CrsMatrix<double>>;
    auto a = toCsr(*A), b = toCsr(*B);
    auto c = FT::multiply(a, transpose_A, b, transpose_B, threshold);
    return fromCsr(*c, *A);
  }
  // B = scalar_A op(A) + scalar_B B: not provided, as in the reference's
  // Tpetra adapter (tpetra/src/MCLS_TpetraCrsMatrixAdapter.hpp:404-411
  // leaves it empty); nothing on the solve path calls it
  static void add(const RCP<const M>&, bool, double, const RCP<M>&, double) {
    MCLS_INSIST(false, "MatrixTraits<Tpetra>::add is not provided");
  }
  static RCP<M> copyTranspose(const M& A) {
    using FT = MatrixTraits<Vector, CrsMatrix<double>>;
    auto t = FT::copyTranspose(*toCsr(A));
    return fromCsr(*t, A);
  }

  // build extension: the local rows as CSR arrays with global column
  // indices (the Kokkos CRS arrays of getLocalMatrix(), columns mapped
  // through the column map)
  static LocalCrs getLocalCrs(const M& A) {
    MCLS_INSIST(A.isFillComplete(), "matrix must be fill-complete");
    tpetra_detail::requireSerialMap(A.getRowMap(), "Tpetra::CrsMatrix");
    auto lcl = A.getLocalMatrix();
    const auto colmap = A.getColMap();
    LocalCrs c;
    c.num_rows = (int64_t)A.getNodeNumRows();
    c.nnz = (int64_t)A.getNodeNumEntries();
    c.rp_store.resize((size_t)c.num_rows + 1);
    for (int64_t i = 0; i <= c.num_rows; ++i) c.rp_store[(size_t)i] = (int64_t)lcl.graph.row_map.data()[i];
    c.col_store.resize((size_t)c.nnz);
    c.val_store.resize((size_t)c.nnz);
    // each row in ascending global column (local columns need not be)
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t i = 0; i < c.num_rows; ++i) {
      const int64_t b = c.rp_store[(size_t)i], e = c.rp_store[(size_t)i + 1];
      row.clear();
      for (int64_t k = b; k < e; ++k)
        row.emplace_back((int32_t)colmap->getGlobalElement(lcl.graph.entries.data()[k]), (double)lcl.values.data()[k]);
      std::sort(row.begin(), row.end(), [](const std::pair<int32_t, double>& x, const std::pair<int32_t, double>& y) {
        return x.first < y.first;
      });
      for (int64_t k = b; k < e; ++k) {
        c.col_store[(size_t)k] = row[(size_t)(k - b)].first;
        c.val_store[(size_t)k] = row[(size_t)(k - b)].second;
      }
    }
    c.row_ptr = c.rp_store.data();
    c.col = c.col_store.data();
    c.val = c.val_store.data();
    c.f32 = false;
    return c;
  }

 private:
  static void insertRow(M& B, GO row, const std::vector<int64_t>& c, const std::vector<double>& v, size_t m) {
    std::vector<GO> cg(c.begin(), c.begin() + m);
    std::vector<S> vs(v.begin(), v.begin() + m);
    B.insertGlobalValues(row, Teuchos::ArrayView<const GO>(cg.data(), (int)m),
                         Teuchos::ArrayView<const S>(vs.data(), (int)m));
  }
  // the local rows as the facade's CSR matrix (global column indices)
  static RCP<CrsMatrix<double>> toCsr(const M& A) {
    const LocalCrs c = getLocalCrs(A);
    std::vector<int64_t> rp(c.rp_store);
    std::vector<int32_t> col(c.col_store);
    std::vector<double> val(c.val_store);
    return std::make_shared<CrsMatrix<double>>(c.num_rows, std::move(rp), std::move(col), std::move(val));
  }
  // a facade CSR matrix back on like's maps
  static RCP<M> fromCsr(const CrsMatrix<double>& C, const M& like) {
    auto B = std::make_shared<M>(like.getRowMap(), (size_t)std::max<int64_t>(C.maxRowEntries(), 1));
    for (int64_t i = 0; i < C.numRows(); ++i) {
      const int64_t b = C.rowPtr()[(size_t)i], e = C.rowPtr()[(size_t)i + 1];
      if (e == b) continue;
      std::vector<GO> cg;
      std::vector<S> vs;
      for (int64_t k = b; k < e; ++k) {
        cg.push_back((GO)C.colInd()[(size_t)k]);
        vs.push_back((S)C.values()[(size_t)k]);
      }
      B->insertGlobalValues((GO)i, Teuchos::ArrayView<const GO>(cg.data(), (int)cg.size()),
                            Teuchos::ArrayView<const S>(vs.data(), (int)vs.size()));
    }
    B->fillComplete(like.getDomainMap(), like.getRangeMap());
    return B;
  }
};

}  // namespace MCLS
