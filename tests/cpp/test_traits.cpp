// Host-only checks of the facade's MatrixTraits / VectorTraits members and
// of LinearProblem's operator family, against dense arithmetic (no GPU).
// Transcribes the reference's adapter and linear-problem tests where they
// pin a value (paths relative to the MCLS source root):
//   tpetra/test/tstTpetraCrsMatrix.cpp (apply :740, transpose :785, multiply :900)
//   tpetra/test/tstTpetraVector.cpp:366-667
//   tpetra/test/tstTpetraLinearProblem.cpp:159-410
// Prints one line per group and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "MCLS_Solvers.hpp"

using namespace MCLS;
using Mat = CrsMatrix<double>;
using MT = MatrixTraits<Vector, Mat>;
using VT = VectorTraits<Vector>;
using Dense = std::vector<std::vector<double>>;

static int g_checks = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    ++g_checks;                                                    \
    if (!(c)) {                                                    \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
      std::exit(1);                                                \
    }                                                              \
  } while (0)

static RCP<Mat> random_matrix(int n, double density, unsigned seed, bool diag = true) {
  std::mt19937 g(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0), p(0.0, 1.0);
  auto A = std::make_shared<Mat>(n);
  for (int i = 0; i < n; ++i) {
    std::vector<int64_t> c;
    std::vector<double> v;
    for (int j = 0; j < n; ++j)
      if ((diag && i == j) || p(g) < density) {
        c.push_back(j);
        v.push_back(i == j && diag ? 3.0 + u(g) : u(g));
      }
    A->insertGlobalValues(i, c, v);
  }
  A->fillComplete();
  return A;
}

static Dense dense(const Mat& A) {
  Dense D((size_t)A.numRows(), std::vector<double>((size_t)A.numRows(), 0.0));
  for (int64_t i = 0; i < A.numRows(); ++i)
    for (int64_t k = A.rowPtr()[(size_t)i]; k < A.rowPtr()[(size_t)i + 1]; ++k)
      D[(size_t)i][(size_t)A.colInd()[(size_t)k]] += A.values()[(size_t)k];
  return D;
}

static Dense matmul(const Dense& X, const Dense& Y) {
  const size_t n = X.size();
  Dense Z(n, std::vector<double>(n, 0.0));
  for (size_t i = 0; i < n; ++i)
    for (size_t k = 0; k < n; ++k)
      for (size_t j = 0; j < n; ++j) Z[i][j] += X[i][k] * Y[k][j];
  return Z;
}

static Dense transpose(const Dense& X) {
  Dense T(X.size(), std::vector<double>(X.size()));
  for (size_t i = 0; i < X.size(); ++i)
    for (size_t j = 0; j < X.size(); ++j) T[j][i] = X[i][j];
  return T;
}

static std::vector<double> mv(const Dense& D, const Vector& x) {
  std::vector<double> y(D.size(), 0.0);
  for (size_t i = 0; i < D.size(); ++i)
    for (size_t j = 0; j < D.size(); ++j) y[i] += D[i][j] * x[(int64_t)j];
  return y;
}

static bool close(const Dense& a, const Dense& b, double tol) {
  for (size_t i = 0; i < a.size(); ++i)
    for (size_t j = 0; j < a.size(); ++j)
      if (std::fabs(a[i][j] - b[i][j]) > tol) return false;
  return true;
}

static bool close(const Vector& a, const std::vector<double>& b, double tol) {
  for (int64_t i = 0; i < a.size(); ++i)
    if (std::fabs(a[i] - b[(size_t)i]) > tol * (1.0 + std::fabs(b[(size_t)i]))) return false;
  return true;
}

static Vector random_vector(int n, unsigned seed) {
  std::mt19937 g(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  Vector v(n);
  for (int i = 0; i < n; ++i) v[i] = u(g);
  return v;
}

static void test_vector_traits() {
  const int n = 17;
  Vector a = random_vector(n, 1), b = random_vector(n, 2), c = random_vector(n, 3);
  auto z = VT::clone(a);
  CHECK(VT::getLocalLength(*z) == n && VT::getGlobalLength(*z) == n && VT::normInf(*z) == 0.0);
  auto d = VT::deepCopy(a);
  CHECK(VT::norm2(*d) == VT::norm2(a));
  double dot = 0.0, n1 = 0.0, ninf = 0.0, sum = 0.0;
  for (int i = 0; i < n; ++i) {
    dot += a[i] * b[i];
    n1 += std::fabs(a[i]);
    ninf = std::max(ninf, std::fabs(a[i]));
    sum += a[i];
  }
  CHECK(VT::dot(a, b) == dot && VT::norm1(a) == n1 && VT::normInf(a) == ninf);
  CHECK(std::fabs(VT::norm2(a) - std::sqrt(dot == dot ? VT::dot(a, a) : 0.0)) < 1e-15);
  CHECK(std::fabs(VT::meanValue(a) - sum / n) < 1e-15);
  Vector e(n);
  VT::abs(e, a);
  VT::reciprocal(*z, b);
  for (int i = 0; i < n; ++i) CHECK(e[i] == std::fabs(a[i]) && (*z)[i] == 1.0 / b[i]);
  VT::scaleCopy(e, 2.5, a);
  for (int i = 0; i < n; ++i) CHECK(e[i] == 2.5 * a[i]);
  Vector f = a;
  VT::update(f, 2.0, b, -3.0, c, 0.5);
  for (int i = 0; i < n; ++i) CHECK(std::fabs(f[i] - (2.0 * a[i] - 3.0 * b[i] + 0.5 * c[i])) < 1e-15);
  Vector g = a;
  VT::elementWiseMultiply(g, 2.0, b, c, 0.25);
  for (int i = 0; i < n; ++i) CHECK(std::fabs(g[i] - (2.0 * a[i] + 0.25 * b[i] * c[i])) < 1e-15);
  {
    // the reference's known answer (tpetra/test/tstTpetraVector.cpp:636-657):
    // A = 1, B = 2, C = 3, alpha = 4, beta = 5 -> 1*4 + 5*2*3 = 34
    Vector A1(10, 1.0), B1(10, 2.0), C1(10, 3.0);
    VT::elementWiseMultiply(A1, 4.0, B1, C1, 5.0);
    for (int i = 0; i < 10; ++i) CHECK(A1[i] == 34.0);
  }
  Vector h(n, std::nan(""));
  VT::update(h, 0.0, a, 1.0);  // alpha = 0 does not read A
  for (int i = 0; i < n; ++i) CHECK(h[i] == a[i]);
  VT::replaceGlobalValue(h, 3, 7.0);
  VT::sumIntoLocalValue(h, 3, 1.0);
  CHECK(h[3] == 8.0 && VT::isLocalRow(h, n - 1) && !VT::isGlobalRow(h, n));
  VT::putScalar(h, 2.0);
  VT::scale(h, 0.5);
  CHECK(VT::normInf(h) == 1.0 && VT::norm1(h) == n);
  VT::randomize(h);
  CHECK(VT::normInf(h) <= 1.0 && VT::normInf(h) > 0.0);
  std::printf("vector traits ok\n");
}

static void test_matrix_traits() {
  const int n = 23;
  auto A = random_matrix(n, 0.2, 5), B = random_matrix(n, 0.15, 6, false);
  const Dense DA = dense(*A), DB = dense(*B);
  Vector x = random_vector(n, 7), y(n), yt(n);
  MT::apply(*A, x, y);
  MT::applyTranspose(*A, x, yt);
  CHECK(close(y, mv(DA, x), 1e-14) && close(yt, mv(transpose(DA), x), 1e-14));
  CHECK(close(dense(*MT::copyTranspose(*A)), transpose(DA), 0.0));
  // multiply in all four transpose combinations, no threshold
  for (int ta = 0; ta < 2; ++ta)
    for (int tb = 0; tb < 2; ++tb) {
      auto C = MT::multiply(A, ta != 0, B, tb != 0);
      Dense want = matmul(ta ? transpose(DA) : DA, tb ? transpose(DB) : DB);
      CHECK(close(dense(*C), want, 1e-14));
    }
  // the row filter: keep |v| > threshold x (signed row maximum)
  {
    const double thr = 0.3;
    auto C = MT::multiply(A, false, B, false, thr);
    Dense full = matmul(DA, DB), got = dense(*C);
    for (int i = 0; i < n; ++i) {
      double rmax = -INFINITY;
      bool any = false;
      for (int j = 0; j < n; ++j) {
        // the row maximum over the product's structural entries (here
        // every non-zero of the dense product)
        if (full[(size_t)i][(size_t)j] != 0.0) {
          rmax = std::max(rmax, full[(size_t)i][(size_t)j]);
          any = true;
        }
      }
      for (int j = 0; j < n && any; ++j) {
        const double v = full[(size_t)i][(size_t)j];
        const double keep = (v != 0.0 && std::fabs(v) > rmax * thr) ? v : 0.0;
        CHECK(std::fabs(got[(size_t)i][(size_t)j] - keep) < 1e-14);
      }
    }
  }
  // add: B = 2 A^T - 0.5 B
  {
    auto Bc = std::make_shared<Mat>(n, B->rowPtr(), B->colInd(), B->values());
    MT::add(A, true, 2.0, Bc, -0.5);
    Dense want = transpose(DA);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) want[(size_t)i][(size_t)j] = 2.0 * want[(size_t)i][(size_t)j] - 0.5 * DB[(size_t)i][(size_t)j];
    CHECK(close(dense(*Bc), want, 1e-14));
  }
  // scaling
  {
    auto S = std::make_shared<Mat>(n, A->rowPtr(), A->colInd(), A->values());
    Vector l = random_vector(n, 8), r = random_vector(n, 9);
    MT::leftScale(*S, l);
    MT::rightScale(*S, r);
    Dense want = DA;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) want[(size_t)i][(size_t)j] *= l[i] * r[j];
    CHECK(close(dense(*S), want, 1e-14));
  }
  // rows, diagonal, maps, the local CSR views
  {
    std::vector<int64_t> c;
    std::vector<double> v;
    size_t m = 0;
    MT::getGlobalRowCopy(*A, 4, c, v, m);
    for (size_t k = 0; k < m; ++k) CHECK(DA[4][(size_t)c[k]] == v[k]);
    Vector d(n);
    MT::getLocalDiagCopy(*A, d);
    for (int i = 0; i < n; ++i) CHECK(d[i] == DA[(size_t)i][(size_t)i]);
    CHECK(MT::getGlobalNumRows(*A) == n && MT::getLocalNumCols(*A) == n && MT::getLocalRow(*A, 5) == 5);
    std::vector<int64_t> rows;
    MT::getMyGlobalRows(*A, rows);
    std::vector<int> ranks;
    MT::getGlobalRowRanks(*A, rows, ranks);
    CHECK((int)rows.size() == n && rows[(size_t)n - 1] == n - 1 && ranks[3] == 0 && MT::getComm(*A)->getSize() == 1);
    const LocalCrs view = MT::getLocalCrs(*A);
    const LocalCrs rebuilt = localCrsFromRows<Vector, Mat>(*A);
    CHECK(view.nnz == rebuilt.nnz && view.num_rows == n && !view.f32);
    for (int64_t k = 0; k < view.nnz; ++k) CHECK(view.col[k] == rebuilt.col[k] && view.value(k) == rebuilt.value(k));
    std::vector<double> dd;
    CHECK(!crsIsDiagonal(view, &dd));
    PointJacobiPreconditioner<double> pj;
    pj.setOperator(A);
    pj.buildPreconditioner();
    CHECK(crsIsDiagonal(MT::getLocalCrs(*pj.getLeftPreconditioner()), &dd));
    for (int i = 0; i < n; ++i) CHECK(dd[(size_t)i] == 1.0 / DA[(size_t)i][(size_t)i]);
    CHECK(pj.getRightPreconditioner() == nullptr && &pj.getOperator() == A.get());
    auto E = MT::exportFromRows(*A, rows);
    CHECK(close(dense(*E), DA, 0.0) && MT::clone(*A)->numRows() == n);
  }
  std::printf("matrix traits ok\n");
}

static void test_linear_problem() {
  const int n = 19;
  auto A = random_matrix(n, 0.25, 11), PL = random_matrix(n, 0.1, 12), PR = random_matrix(n, 0.1, 13);
  const Dense DA = dense(*A), DL = dense(*PL), DR = dense(*PR);
  auto x = std::make_shared<Vector>(random_vector(n, 14));
  auto b = std::make_shared<const Vector>(random_vector(n, 15));
  LinearProblem<Vector, Mat> p(A, x, b);
  Vector y(n), v = random_vector(n, 16);
  // no preconditioners: apply = A, identity preconditioners
  p.apply(v, y);
  CHECK(close(y, mv(DA, v), 1e-13));
  p.applyLeftPrec(v, y);
  CHECK(close(y, std::vector<double>(v.raw()), 0.0));
  p.setLeftPrec(PL);
  p.setRightPrec(PR);
  p.apply(v, y);
  CHECK(close(y, mv(matmul(DL, matmul(DA, DR)), v), 1e-13));
  p.applyTranspose(v, y);
  CHECK(close(y, mv(transpose(matmul(DL, matmul(DA, DR))), v), 1e-13));
  p.applyOp(v, y);
  CHECK(close(y, mv(DA, v), 1e-13));
  // rp = PL (b - A PR x)
  p.updatePrecResidual();
  std::vector<double> t = mv(matmul(DA, DR), *x);
  Vector bmt(n);
  for (int i = 0; i < n; ++i) bmt[i] = (*b)[i] - t[(size_t)i];
  CHECK(close(*p.getPrecResidual(), mv(DL, bmt), 1e-13));
  // x += PR u
  auto u = std::make_shared<Vector>(random_vector(n, 17));
  Vector x0 = *x;
  p.updateSolution(u);
  std::vector<double> pu = mv(DR, *u);
  for (int i = 0; i < n; ++i) CHECK(std::fabs((*x)[i] - (x0[i] + pu[(size_t)i])) < 1e-14);
  // tpetra/test/tstTpetraLinearProblem.cpp:354-410 (one rank): A upper
  // bidiagonal (1, 2), PL = A^T: the composite's interior rows are (2, 5, 2)
  {
    const int m = 10;
    auto Ab = std::make_shared<Mat>(m);
    for (int i = 0; i < m - 1; ++i) Ab->insertGlobalValues(i, {i, i + 1}, {1.0, 2.0});
    Ab->fillComplete();
    auto X = MT::cloneVectorFromMatrixRows(*Ab);
    VT::putScalar(*X, 2.0);
    LinearProblem<Vector, Mat> lp(Ab, X, X);
    lp.setLeftPrec(MT::copyTranspose(*Ab));
    auto D = lp.getCompositeOperator();
    std::vector<int64_t> c;
    std::vector<double> w;
    size_t ne = 0;
    for (int i = 1; i < m - 1; ++i) {
      MT::getLocalRowCopy(*D, i, c, w, ne);
      CHECK(ne == 3 && c[0] == i - 1 && c[1] == i && c[2] == i + 1);
      CHECK(w[0] == 2.0 && w[1] == 5.0 && w[2] == 2.0);
    }
    // the transposed composite of the same problem is its transpose
    CHECK(close(dense(*lp.getTransposeCompositeOperator()), transpose(dense(*D)), 0.0));
  }
  std::printf("linear problem ok\n");
}

int main() {
  test_vector_traits();
  test_matrix_traits();
  test_linear_problem();
  std::printf("ALL %d checks passed\n", g_checks);
  return 0;
}
