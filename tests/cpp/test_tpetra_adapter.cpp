// The Tpetra binding, mcls_amd/host/MCLS_TpetraAdapter.hpp, compiled and
// exercised against tests/cpp/tpetra_double.hpp (Trilinos is not installed
// here).  Host checks: every trait member against the facade's own CSR
// traits on the same operator, and the vector traits against the known
// answers of the reference's tests (tpetra/test/tstTpetraVector.cpp,
// tstTpetraCrsMatrix.cpp; paths relative to the MCLS source root).  With
// --gpu, both managers solve on the device through the binding and must give
// what they give on the facade's own types for the same operator, right-hand
// side and streams.  Prints one line per group and exits non-zero on the
// first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tpetra_double.hpp"
#include "MCLS_Solvers.hpp"
#include "MCLS_TpetraAdapter.hpp"

using namespace MCLS;
using LO = int;
using GO = long long;
using TMap = Tpetra::Map<LO, GO>;
using TV = Tpetra::Vector<double, LO, GO>;
using TM = Tpetra::CrsMatrix<double, LO, GO>;
using VTT = VectorTraits<TV>;
using MTT = MatrixTraits<TV, TM>;
using FM = CrsMatrix<double>;
using FT = MatrixTraits<Vector, FM>;

static int g_checks = 0;
#define CHECK(c)                                                             \
  do {                                                                       \
    ++g_checks;                                                              \
    if (!(c)) {                                                              \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);               \
      std::exit(1);                                                          \
    }                                                                        \
  } while (0)

static Teuchos::RCP<const TMap> serial_map(size_t n) {
  Teuchos::RCP<const Teuchos::Comm<int>> comm = std::make_shared<Teuchos::SerialComm<int>>();
  return std::make_shared<TMap>(n, (GO)0, comm);
}

// a nonsymmetric 2-D operator on an m x m grid: 5-point couplings with a
// drift, diagonally dominant, rows inserted in shuffled column order and with
// a duplicate entry (summed at fillComplete)
static void build_pair(int m, RCP<FM>* fa, RCP<TM>* ta) {
  const int n = m * m;
  auto A = std::make_shared<FM>(n);
  auto T = std::make_shared<TM>(serial_map((size_t)n), (size_t)6);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      const int r = i + m * j;
      std::vector<int64_t> c;
      std::vector<double> v;
      if (j + 1 < m) c.push_back(r + m), v.push_back(-0.9 - 0.001 * r);
      if (i > 0) c.push_back(r - 1), v.push_back(-1.2);
      c.push_back(r), v.push_back(4.5 + 0.01 * (r % 7));
      if (i + 1 < m) c.push_back(r + 1), v.push_back(-0.8);
      if (j > 0) c.push_back(r - m), v.push_back(-1.1);
      std::vector<int64_t> cf = c;
      std::vector<double> vf = v;
      A->insertGlobalValues(r, cf, vf);
      // the Tpetra side gets the diagonal in two halves
      std::vector<GO> cg(c.begin(), c.end());
      for (size_t k = 0; k < v.size(); ++k)
        if (c[k] == r) v[k] *= 0.5;
      cg.push_back(r);
      v.push_back(0.5 * (4.5 + 0.01 * (r % 7)));
      T->insertGlobalValues((GO)r, Teuchos::ArrayView<const GO>(cg.data(), (int)cg.size()),
                            Teuchos::ArrayView<const double>(v.data(), (int)v.size()));
    }
  A->fillComplete();
  T->fillComplete();
  *fa = A;
  *ta = T;
}

static RCP<TV> tvec(const TM& A, double base) {
  auto v = MTT::cloneVectorFromMatrixRows(A);
  auto d = v->getDataNonConst();
  for (size_t i = 0; i < v->getLocalLength(); ++i) d[i] = base + 0.37 * std::sin(1.0 + (double)i);
  return v;
}
static Vector fvec(int64_t n, double base) {
  Vector v(n);
  for (int64_t i = 0; i < n; ++i) v[i] = base + 0.37 * std::sin(1.0 + (double)i);
  return v;
}
static double maxdiff(const TV& a, const Vector& b) {
  double d = 0.0;
  for (int64_t i = 0; i < b.size(); ++i) d = std::max(d, std::fabs(a.getData()[(size_t)i] - b[i]));
  return d;
}

static void test_matrix_traits() {
  RCP<FM> A;
  RCP<TM> T;
  build_pair(9, &A, &T);
  const int64_t n = A->numRows();
  // the local CRS arrays: rows in ascending global column, duplicates summed,
  // whatever the column map's local order
  const LocalCrs c = MTT::getLocalCrs(*T);
  CHECK(c.num_rows == n && c.nnz == A->nnz() && !c.f32);
  for (int64_t i = 0; i <= n; ++i) CHECK(c.row_ptr[i] == A->rowPtr()[(size_t)i]);
  for (int64_t k = 0; k < c.nnz; ++k) {
    CHECK(c.col[k] == A->colInd()[(size_t)k]);
    CHECK(std::fabs(c.value(k) - A->values()[(size_t)k]) <= 1e-15 * std::fabs(A->values()[(size_t)k]));
  }
  // the column map is not in ascending order, so the mapping did something
  bool permuted = false;
  for (int64_t k = 0; k < (int64_t)T->getColMap()->getNodeNumElements(); ++k)
    permuted = permuted || T->getColMap()->getGlobalElement((LO)k) != k;
  CHECK(permuted);
  CHECK(MTT::isFillComplete(*T) && MTT::getLocalNumRows(*T) == n && MTT::getGlobalNumRows(*T) == n);
  CHECK(MTT::getLocalNumCols(*T) == n && MTT::getGlobalMaxNumRowEntries(*T) == 5);
  CHECK(MTT::getGlobalRow(*T, 7) == 7 && MTT::getLocalRow(*T, 7) == 7);
  CHECK(MTT::isGlobalRow(*T, n - 1) && !MTT::isGlobalRow(*T, n) && MTT::isLocalRow(*T, 0));
  const LO lc = MTT::getLocalCol(*T, 40);
  CHECK(MTT::getGlobalCol(*T, lc) == 40 && MTT::isGlobalCol(*T, 40) && MTT::isLocalCol(*T, lc));
  std::vector<int64_t> rows, cols;
  MTT::getMyGlobalRows(*T, rows);
  MTT::getMyGlobalCols(*T, cols);
  CHECK((int64_t)rows.size() == n && rows[5] == 5 && (int64_t)cols.size() == n);
  std::vector<int> ranks;
  MTT::getGlobalRowRanks(*T, {0, 3, 9}, ranks);
  CHECK(ranks.size() == 3 && ranks[0] == 0 && ranks[2] == 0);
  // row copies: the facade's row as a set of (column, value)
  for (int64_t r : {0LL, 10LL, 40LL, (long long)n - 1}) {
    std::vector<int64_t> gc, fc;
    std::vector<double> gv, fv;
    size_t gm = 0, fm = 0;
    MTT::getGlobalRowCopy(*T, r, gc, gv, gm);
    FT::getGlobalRowCopy(*A, r, fc, fv, fm);
    CHECK(gm == fm);
    for (size_t k = 0; k < gm; ++k) {
      size_t q = 0;
      while (q < fm && fc[q] != gc[k]) ++q;
      CHECK(q < fm && std::fabs(fv[q] - gv[k]) <= 1e-15 * std::fabs(fv[q]));
    }
    std::vector<int64_t> lcols;
    std::vector<double> lvals;
    size_t lm = 0;
    MTT::getLocalRowCopy(*T, r, lcols, lvals, lm);
    CHECK(lm == gm);
    for (size_t k = 0; k < lm; ++k) CHECK(MTT::getGlobalCol(*T, (LO)lcols[k]) == gc[k] && lvals[k] == gv[k]);
  }
  // apply, applyTranspose, diagonal, scaling
  auto x = tvec(*T, 0.5);
  Vector fx = fvec(n, 0.5);
  auto y = MTT::cloneVectorFromMatrixRows(*T);
  Vector fy(n);
  MTT::apply(*T, *x, *y);
  FT::apply(*A, fx, fy);
  CHECK(maxdiff(*y, fy) <= 1e-13);
  MTT::applyTranspose(*T, *x, *y);
  FT::applyTranspose(*A, fx, fy);
  CHECK(maxdiff(*y, fy) <= 1e-13);
  auto d = MTT::cloneVectorFromMatrixRows(*T);
  Vector fd(n);
  MTT::getLocalDiagCopy(*T, *d);
  FT::getLocalDiagCopy(*A, fd);
  CHECK(maxdiff(*d, fd) <= 1e-15);
  // products and transposes on the facade's CSR semantics, back on Tpetra maps
  RCP<FM> A2;
  RCP<TM> T2;
  build_pair(9, &A2, &T2);
  for (bool ta : {false, true})
    for (bool tb : {false, true})
      for (double thr : {0.0, 0.05}) {
        auto P = MTT::multiply(T, ta, T2, tb, thr);
        auto Q = FT::multiply(A, ta, A2, tb, thr);
        const LocalCrs pc = MTT::getLocalCrs(*P);
        CHECK(pc.nnz == Q->nnz());
        for (int64_t k = 0; k < pc.nnz; ++k) {
          CHECK(pc.col[k] == Q->colInd()[(size_t)k]);
          CHECK(std::fabs(pc.value(k) - Q->values()[(size_t)k]) <= 1e-13 * std::fabs(Q->values()[(size_t)k]));
        }
      }
  auto Tt = MTT::copyTranspose(*T);
  auto At = FT::copyTranspose(*A);
  const LocalCrs tc = MTT::getLocalCrs(*Tt);
  CHECK(tc.nnz == At->nnz());
  for (int64_t k = 0; k < tc.nnz; ++k) CHECK(tc.col[k] == At->colInd()[(size_t)k]);
  auto E = MTT::exportFromRows(*T, {0, 5, 6});
  CHECK(E->getNodeNumEntries() == T->getNumEntriesInGlobalRow(0) + T->getNumEntriesInGlobalRow(5) +
                                      T->getNumEntriesInGlobalRow(6));
  auto C0 = MTT::clone(*T);
  CHECK(C0->getNodeNumRows() == (size_t)n && !C0->isFillComplete());
  MTT::leftScale(*T, *x);
  FT::leftScale(*A, fx);
  MTT::rightScale(*T, *x);
  FT::rightScale(*A, fx);
  MTT::apply(*T, *x, *y);
  FT::apply(*A, fx, fy);
  CHECK(maxdiff(*y, fy) <= 1e-12);
  CHECK(MTT::copyNearestNeighbors(T, 1) == T);
  bool threw = false;
  try {
    MTT::add(T, false, 1.0, MTT::clone(*T), 1.0);
  } catch (const Assertion&) {
    threw = true;
  }
  CHECK(threw);
  // a map that does not start at 0 is refused
  auto off = std::make_shared<TM>(std::make_shared<TMap>((size_t)4, (GO)1, T->getComm()), (size_t)1);
  for (GO r = 1; r <= 4; ++r) {
    const GO cr = r;
    const double one = 1.0;
    off->insertGlobalValues(r, Teuchos::ArrayView<const GO>(&cr, 1), Teuchos::ArrayView<const double>(&one, 1));
  }
  off->fillComplete();
  threw = false;
  try {
    MTT::getLocalCrs(*off);
  } catch (const Assertion&) {
    threw = true;
  }
  CHECK(threw);
  std::printf("ok Tpetra matrix traits (%lld rows, %lld entries)\n", (long long)n, (long long)c.nnz);
}

// the reference's vector-trait known answers (tpetra/test/tstTpetraVector.cpp)
static void test_vector_traits() {
  auto map = serial_map(10);
  auto A = std::make_shared<TV>(map);
  VTT::putScalar(*A, 1.0);
  auto B = VTT::clone(*A);
  VTT::putScalar(*B, 2.0);
  auto C = VTT::clone(*A);
  VTT::putScalar(*C, 3.0);
  VTT::elementWiseMultiply(*A, 4.0, *B, *C, 5.0);  // 1*4 + 5*2*3
  for (int i = 0; i < 10; ++i) CHECK(VTT::view(*A)[i] == 34.0);
  VTT::putScalar(*A, 1.0);
  VTT::update(*A, 2.0, *B, 3.0);  // 2*1 + 3*2
  for (int i = 0; i < 10; ++i) CHECK(VTT::view(*A)[i] == 8.0);
  VTT::putScalar(*A, 1.0);
  VTT::update(*A, 2.0, *B, 3.0, *C, 4.0);  // 2 + 6 + 12
  for (int i = 0; i < 10; ++i) CHECK(VTT::view(*A)[i] == 20.0);
  VTT::putScalar(*A, -2.0);
  CHECK(VTT::norm1(*A) == 20.0 && VTT::normInf(*A) == 2.0 && std::fabs(VTT::norm2(*A) - std::sqrt(40.0)) < 1e-14);
  CHECK(VTT::dot(*A, *B) == -40.0 && VTT::meanValue(*A) == -2.0);
  VTT::abs(*C, *A);
  VTT::reciprocal(*B, *C);
  VTT::scaleCopy(*A, 3.0, *B);
  for (int i = 0; i < 10; ++i) CHECK(VTT::view(*C)[i] == 2.0 && VTT::view(*B)[i] == 0.5 && VTT::view(*A)[i] == 1.5);
  VTT::scale(*A, 2.0);
  VTT::replaceGlobalValue(*A, 3, 7.0);
  VTT::sumIntoLocalValue(*A, 3, 1.0);
  VTT::replaceLocalValue(*A, 4, -1.0);
  VTT::sumIntoGlobalValue(*A, 4, 0.5);
  CHECK(VTT::view(*A)[3] == 8.0 && VTT::view(*A)[4] == -0.5 && VTT::view(*A)[0] == 3.0);
  auto D = VTT::deepCopy(*A);
  VTT::viewNonConst(*D)[0] = 9.0;
  CHECK(VTT::view(*A)[0] == 3.0 && VTT::view(*D)[0] == 9.0);
  CHECK(VTT::getGlobalLength(*A) == 10 && VTT::getLocalLength(*A) == 10 && VTT::isGlobalRow(*A, 9) &&
        !VTT::isGlobalRow(*A, 10) && VTT::getGlobalRow(*A, 2) == 2 && VTT::getLocalRow(*A, 2) == 2);
  CHECK(VTT::getVectorNonConst(A, 0) == A && VTT::getComm(*A)->getSize() == 1);
  VTT::randomize(*A);
  CHECK(VTT::normInf(*A) <= 1.0 && VTT::norm1(*A) > 0.0);
  std::printf("ok Tpetra vector traits\n");
}

// point Jacobi on the Tpetra type, through the Preconditioner interface
class TpetraPointJacobi : public Preconditioner<TM> {
 public:
  RCP<const ParameterList> getValidParameters() const override { return parameterList(); }
  RCP<const ParameterList> getCurrentParameters() const override { return parameterList(); }
  void setParameters(const RCP<ParameterList>&) override {}
  void setOperator(const RCP<const TM>& A) override { d_A = A; }
  const TM& getOperator() const override { return *d_A; }
  void buildPreconditioner() override {
    auto d = MTT::cloneVectorFromMatrixRows(*d_A);
    MTT::getLocalDiagCopy(*d_A, *d);
    auto P = std::make_shared<TM>(d_A->getRowMap(), (size_t)1);
    for (size_t i = 0; i < d->getLocalLength(); ++i) {
      const GO r = (GO)i;
      const double v = 1.0 / d->getData()[i];
      P->insertGlobalValues(r, Teuchos::ArrayView<const GO>(&r, 1), Teuchos::ArrayView<const double>(&v, 1));
    }
    P->fillComplete();
    d_P = P;
  }
  RCP<const TM> getLeftPreconditioner() const override { return d_P; }
  RCP<const TM> getRightPreconditioner() const override { return nullptr; }

 private:
  RCP<const TM> d_A, d_P;
};

// both managers through the binding, against the facade's own types
static void test_managers_on_gpu() {
  RCP<FM> A;
  RCP<TM> T;
  build_pair(12, &A, &T);
  const int64_t n = A->numRows();
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-10);
  plist->set<int>("Maximum Iterations", 60);
  plist->set<double>("Sample Ratio", 40.0);
  plist->set<int>("Random Seed", 31);
  Vector fb = fvec(n, 1.0);
  auto b1 = std::make_shared<Vector>(fb);
  auto x1 = std::make_shared<Vector>(n);
  auto p1 = std::make_shared<LinearProblem<Vector, FM>>(A, x1, b1);
  PointJacobiPreconditioner<double> pj1;
  pj1.setOperator(A);
  pj1.buildPreconditioner();
  p1->setLeftPrec(pj1.getLeftPreconditioner());
  MCSASolverManager<Vector, FM, AdjointTag> s1(p1, plist);
  CHECK(s1.solve());
  auto b2 = tvec(*T, 1.0);
  auto x2 = MTT::cloneVectorFromMatrixRows(*T);
  auto p2 = std::make_shared<LinearProblem<TV, TM>>(T, x2, b2);
  TpetraPointJacobi pj2;
  pj2.setOperator(T);
  pj2.buildPreconditioner();
  p2->setLeftPrec(pj2.getLeftPreconditioner());
  MCSASolverManager<TV, TM, AdjointTag> s2(p2, plist);
  CHECK(s2.solve() && s2.achievedTol() <= 1e-10);
  CHECK(std::abs(s2.getNumIters() - s1.getNumIters()) <= 1);
  const double dx = maxdiff(*x2, *x1);
  CHECK(dx <= 1e-8 * VectorTraits<Vector>::normInf(*x1));
  p2->updatePrecResidual();
  CHECK(VTT::normInf(*p2->getPrecResidual()) <= 1e-9 * VTT::normInf(*b2));
  // the MC manager: same source, same solve number on a fresh context
  auto y1 = std::make_shared<Vector>(n);
  auto y2 = MTT::cloneVectorFromMatrixRows(*T);
  auto q1 = std::make_shared<LinearProblem<Vector, FM>>(A, y1, b1);
  auto q2 = std::make_shared<LinearProblem<TV, TM>>(T, y2, b2);
  q1->setLeftPrec(pj1.getLeftPreconditioner());
  q2->setLeftPrec(pj2.getLeftPreconditioner());
  MonteCarloSolverManager<Vector, FM, AdjointTag> m1(q1, plist);
  MonteCarloSolverManager<TV, TM, AdjointTag> m2(q2, plist);
  CHECK(m1.solve() && m2.solve());
  const double dy = maxdiff(*y2, *y1);
  CHECK(dy <= 1e-12 * VectorTraits<Vector>::normInf(*y1));
  std::printf("ok Tpetra binding on the GPU: MCSA %d / %d iterations, solution diff %.2e, MC diff %.2e\n",
              s2.getNumIters(), s1.getNumIters(), dx, dy);
}

int main(int argc, char** argv) {
  test_matrix_traits();
  test_vector_traits();
  if (argc > 1 && std::strcmp(argv[1], "--gpu") == 0) test_managers_on_gpu();
  std::printf("ALL %d checks passed\n", g_checks);
  return 0;
}
