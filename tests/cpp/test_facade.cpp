// C++ facade tests on the GPU, transcribing the reference's manager tests
// (paths relative to the MCLS source root):
//   tpetra/test/tstTpetraMCSASolverManager.cpp:72-203        (adjoint)
//   tpetra/test/tstTpetraMCSASolverManager.cpp:206-355       (prec_adjoint)
//   tpetra/test/tstTpetraMCSASolverManager.cpp:494-...        (prec_forward)
//   tpetra/test/tstTpetraMonteCarloSolverManager.cpp:200-340 (adjoint_prec)
//   tpetra/test/tstTpetraMonteCarloSolverManager.cpp:73-197  (adjoint MC)
//   src/MCLS_SolverFactory_impl.hpp:58-146                   (names)
// Prints one line per check and exits non-zero on the first failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "MCLS_MatrixMarket.hpp"
#include "MCLS_Solvers.hpp"

using namespace MCLS;
using VectorType = Vector;
using MatrixType = CrsMatrix<double>;
using VT = VectorTraits<VectorType>;
using MT = MatrixTraits<VectorType, MatrixType>;

static int g_checks = 0;
#define CHECK(c)                                                             \
  do {                                                                       \
    ++g_checks;                                                              \
    if (!(c)) {                                                              \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);               \
      std::exit(1);                                                          \
    }                                                                        \
  } while (0)

// tridiagonal 1 / -0.14 with explicit zeros in the corner rows
static RCP<MatrixType> build_operator(int n, double diag, double off) {
  auto A = std::make_shared<MatrixType>(n);
  A->insertGlobalValues(0, {0, 1, 2}, {diag, off, 0.0});
  for (int i = 1; i < n - 1; ++i) A->insertGlobalValues(i, {i - 1, i, i + 1}, {off, diag, off});
  A->insertGlobalValues(n - 1, {n - 3, n - 2, n - 1}, {0.0, off, diag});
  A->fillComplete();
  return A;
}

static bool all_sign(const VectorType& x, double s) {
  for (int64_t i = 0; i < x.size(); ++i)
    if (!(x[i] * s > 0.0)) return false;
  return true;
}

static void test_mcsa_adjoint(bool prec) {
  const int n = 10;
  auto A = build_operator(n, prec ? 2.0 : 1.0, prec ? -0.28 : -0.14);
  auto x = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*x, 0.0);
  auto b = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*b, -1.0);
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-8);
  plist->set<int>("Maximum Iterations", 20);
  plist->set<int>("MC Check Frequency", 50);
  plist->set<double>("Sample Ratio", 1.0);
  plist->set<std::string>("Transport Type", "Global");
  plist->set<int>("Random Seed", 12345);
  auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
  if (prec) {
    PointJacobiPreconditioner<double> pj;
    pj.setOperator(A);
    pj.buildPreconditioner();
    problem->setLeftPrec(pj.getLeftPreconditioner());
  }
  auto ms = std::make_shared<MultiSetLinearProblem<VectorType, MatrixType>>(1, 0, problem);
  MCSASolverManager<VectorType, MatrixType, AdjointTag> solver(ms, plist);

  bool conv = solver.solve();
  CHECK(conv && solver.getConvergedStatus());
  CHECK(solver.getNumIters() < 20);
  CHECK(solver.achievedTol() > 0.0 && solver.achievedTol() <= 1e-8);
  CHECK(all_sign(*x, -1.0));

  VT::putScalar(*const_cast<VectorType*>(b.get()), 2.0);
  VT::putScalar(*x, 0.0);
  problem->setLHS(x);
  CHECK(solver.solve() && solver.getNumIters() < 20);
  CHECK(all_sign(*x, 1.0));

  VT::putScalar(*x, 0.0);
  solver.setProblem(problem);
  CHECK(solver.solve() && solver.getNumIters() < 20);
  CHECK(all_sign(*x, 1.0));

  VT::putScalar(*const_cast<VectorType*>(b.get()), -2.0);
  VT::putScalar(*x, 0.0);
  CHECK(solver.solve() && solver.getNumIters() < 20);
  CHECK(all_sign(*x, -1.0));
  std::printf("ok MCSASolverManager.%s iters=%d tol=%.2e\n", prec ? "prec_adjoint" : "adjoint",
              solver.getNumIters(), solver.achievedTol());
}

static void test_mc_adjoint() {
  const int n = 10;
  auto A = build_operator(n, 1.0, -0.14);
  auto x = MT::cloneVectorFromMatrixRows(*A);
  auto b = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*b, -1.0);
  auto plist = parameterList();
  plist->set<double>("Sample Ratio", 10.0);
  plist->set<int>("History Length", 10);
  auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
  MonteCarloSolverManager<VectorType, MatrixType, AdjointTag> solver(problem, plist, 0);
  CHECK(solver.solve() && solver.getConvergedStatus());
  CHECK(solver.getNumIters() == 0);
  CHECK(all_sign(*x, -1.0));
  CHECK(solver.achievedTol() < 0.5);
  std::printf("ok MonteCarloSolverManager.adjoint tol=%.3f\n", solver.achievedTol());
}

// tpetra/test/tstTpetraMonteCarloSolverManager.cpp:360-467 (forward):
// signs of the forward MC estimate for sources -1, 2, 2 (after setProblem), -2
static void test_mc_forward() {
  const int n = 10;
  auto A = build_operator(n, 1.0, -0.14);
  auto x = MT::cloneVectorFromMatrixRows(*A);
  auto b = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*b, -1.0);
  auto plist = parameterList();
  plist->set<double>("Sample Ratio", 10.0);
  plist->set<int>("History Length", 10);
  auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
  MonteCarloSolverManager<VectorType, MatrixType, ForwardTag> solver(problem, plist, 0);
  CHECK(solver.solve() && solver.getConvergedStatus() && solver.getNumIters() == 0);
  CHECK(all_sign(*x, -1.0));
  VT::putScalar(*const_cast<VectorType*>(b.get()), 2.0);
  CHECK(solver.solve() && all_sign(*x, 1.0));
  solver.setProblem(problem);
  CHECK(solver.solve() && all_sign(*x, 1.0));
  VT::putScalar(*const_cast<VectorType*>(b.get()), -2.0);
  CHECK(solver.solve() && all_sign(*x, -1.0));
  // forward MCSA converges like the adjoint one
  auto mplist = parameterList();
  mplist->set<double>("Convergence Tolerance", 1.0e-8);
  mplist->set<int>("Maximum Iterations", 20);
  mplist->set<int>("Random Seed", 12345);
  VT::putScalar(*x, 0.0);
  MCSASolverManager<VectorType, MatrixType, ForwardTag> mcsa(problem, mplist);
  CHECK(mcsa.solve() && mcsa.getNumIters() < 20 && mcsa.achievedTol() <= 1e-8);
  CHECK(all_sign(*x, -1.0));
  std::printf("ok MonteCarloSolverManager.forward, MCSASolverManager.forward iters=%d\n", mcsa.getNumIters());
}

// A system read from MatrixMarket files (run from the repository root).
static void test_matrix_market_system() {
  auto A = readMatrixMarket<double>("tests/data/mm/symmetric.mtx");
  auto b = readMatrixMarketVector("tests/data/mm/vector.mtx");
  auto x = MT::cloneVectorFromMatrixRows(*A);
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-10);
  plist->set<int>("Maximum Iterations", 40);
  plist->set<double>("Sample Ratio", 200.0);
  plist->set<int>("Random Seed", 1);
  auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
  PointJacobiPreconditioner<double> pj;
  pj.setOperator(A);
  pj.buildPreconditioner();
  problem->setLeftPrec(pj.getLeftPreconditioner());
  MCSASolverManager<VectorType, MatrixType, AdjointTag> solver(problem, plist);
  CHECK(solver.solve());
  // residual of the unpreconditioned system
  double r2 = 0.0, b2 = 0.0;
  for (int64_t i = 0; i < A->numRows(); ++i) {
    double ax = 0.0;
    for (int64_t k = A->rowPtr()[(size_t)i]; k < A->rowPtr()[(size_t)i + 1]; ++k)
      ax += A->values()[(size_t)k] * (*x)[A->colInd()[(size_t)k]];
    r2 += ((*b)[i] - ax) * ((*b)[i] - ax);
    b2 += (*b)[i] * (*b)[i];
  }
  CHECK(std::sqrt(r2 / b2) <= 1e-9);
  std::printf("ok MatrixMarket system iters=%d\n", solver.getNumIters());
}

static void test_factory_and_errors() {
  SolverFactory<VectorType, MatrixType> factory;
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-8);
  plist->set<int>("Maximum Iterations", 20);
  plist->set<int>("Random Seed", 7);
  auto mcsa = factory.create("Adjoint MCSA", plist);
  auto mc = factory.create("Adjoint MC", plist);
  CHECK(mcsa && mc);
  auto fmcsa = factory.create("Forward MCSA", plist);
  auto fmc = factory.create("Forward MC", plist);
  CHECK(fmcsa && fmc);
  bool threw = false;
  try {
    factory.create("Adjoint Anderson", plist);
  } catch (const Assertion&) {
    threw = true;
  }
  CHECK(threw);
  threw = false;
  try {
    factory.create("Nope", plist);
  } catch (const Assertion&) {
    threw = true;
  }
  CHECK(threw);

  const int n = 10;
  auto A = build_operator(n, 1.0, -0.14);
  auto x = MT::cloneVectorFromMatrixRows(*A);
  auto b = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*b, 1.0);
  auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
  mcsa->setProblem(problem);
  CHECK(mcsa->solve() && all_sign(*x, 1.0));

  // bad parameter value -> MCLS::Assertion through the C ABI status
  auto bad = parameterList();
  bad->set<std::string>("Estimator Type", "Track Length");
  mcsa->setParameters(bad);
  threw = false;
  try {
    mcsa->solve();
  } catch (const Assertion& e) {
    threw = std::string(e.what()).find("Estimator Type") != std::string::npos;
  }
  CHECK(threw);
  std::printf("ok SolverFactory names and error convention\n");
}

// 2D 5-point Laplacian with a varying coefficient per row; the diagonal
// importance through the facade, kept across a change of operator.
static RCP<MatrixType> build_varcoef(int m, double scale) {
  auto A = std::make_shared<MatrixType>((int64_t)m * m);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      const int r = j * m + i;
      const double k = scale * (1.0 + 0.5 * ((i * 7 + j * 3) % 5));
      std::vector<int64_t> c;
      std::vector<double> v;
      if (j > 0) { c.push_back(r - m); v.push_back(-k); }
      if (i > 0) { c.push_back(r - 1); v.push_back(-k); }
      c.push_back(r); v.push_back(4.0 * k + 0.05);
      if (i < m - 1) { c.push_back(r + 1); v.push_back(-k); }
      if (j < m - 1) { c.push_back(r + m); v.push_back(-k); }
      A->insertGlobalValues(r, c, v);
    }
  A->fillComplete();
  return A;
}

static void test_mcsa_importance() {
  const int m = 16;
  auto A = build_varcoef(m, 1.0);
  auto x = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*x, 0.0);
  auto b = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*b, 1.0);
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-8);
  plist->set<int>("Maximum Iterations", 60);
  plist->set<double>("Sample Ratio", 64.0);
  plist->set<double>("History Length", 100000);
  plist->set<double>("Weight Cutoff", 1e-4);
  plist->set<double>("Richardson Relaxation", 0.7);
  plist->set<int>("Smoother Steps", 3);
  plist->set<std::string>("Transition Sampler", "Learned");
  plist->set<int>("Random Seed", 3);
  auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
  PointJacobiPreconditioner<double> pj;
  pj.setOperator(A);
  pj.buildPreconditioner();
  problem->setLeftPrec(pj.getLeftPreconditioner());
  MCSASolverManager<VectorType, MatrixType, AdjointTag> solver(problem, plist);
  bool threw = false;
  try {
    solver.solve();  // "Learned" without an importance is a state error
  } catch (const Assertion&) {
    threw = true;
  }
  CHECK(threw);
  solver.setImportanceDiagonal();
  VT::putScalar(*x, 0.0);
  CHECK(solver.solve() && solver.achievedTol() <= 1e-8);
  CHECK(all_sign(*x, 1.0));
  const int first_iters = solver.getNumIters();
  // a new operator: the importance is recomputed for it
  auto A2 = build_varcoef(m, 2.0);
  problem->setOperator(A2);
  pj.setOperator(A2);
  pj.buildPreconditioner();
  problem->setLeftPrec(pj.getLeftPreconditioner());
  VT::putScalar(*x, 0.0);
  CHECK(solver.solve() && solver.achievedTol() <= 1e-8);
  CHECK(all_sign(*x, 1.0));
  std::printf("ok MCSASolverManager.importance iters=%d,%d\n", first_iters, solver.getNumIters());
  // autotuned launch shape: same convergence
  mcls_gpu_tuning t = solver.autotune(1 << 16);
  CHECK(t.candidates >= 10 && t.ms_best > 0.0 && t.ms_confirmed > 0.0 && t.block >= 64 && t.batch >= 64);
  VT::putScalar(*x, 0.0);
  CHECK(solver.solve() && solver.achievedTol() <= 1e-8);
  std::printf("ok MCSASolverManager.autotune block=%lld per_cu=%lld batch=%lld\n", (long long)t.block,
              (long long)t.blocks_per_cu, (long long)t.batch);
}

static RCP<MatrixType> identity(int n, double v) {
  auto I = std::make_shared<MatrixType>(n);
  for (int i = 0; i < n; ++i) I->insertGlobalValues(i, {i}, {v});
  I->fillComplete();
  return I;
}

// prec_adjoint / prec_forward and adjoint_prec / forward_prec of the
// reference: PL = PR = I, signs of x for sources -1, 2, 2 (after
// setProblem), -2, fewer than 20 MCSA iterations.
template <class Tag>
static void test_identity_prec(const char* name) {
  const int n = 10;
  auto A = build_operator(n, 1.0, -0.14);
  auto x = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*x, 0.0);
  auto b = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*b, -1.0);
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-8);
  plist->set<int>("Maximum Iterations", 20);
  plist->set<double>("Sample Ratio", 1.0);
  plist->set<int>("Random Seed", 4321);
  auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
  auto I = identity(n, 1.0);
  problem->setLeftPrec(I);
  problem->setRightPrec(I);
  MCSASolverManager<VectorType, MatrixType, Tag> solver(problem, plist);
  CHECK(solver.solve() && solver.getConvergedStatus() && solver.getNumIters() < 20 && solver.achievedTol() > 0.0);
  CHECK(all_sign(*x, -1.0));
  VT::putScalar(*const_cast<VectorType*>(b.get()), 2.0);
  VT::putScalar(*x, 0.0);
  problem->setLHS(x);
  CHECK(solver.solve() && solver.getNumIters() < 20 && all_sign(*x, 1.0));
  VT::putScalar(*x, 0.0);
  solver.setProblem(problem);
  CHECK(solver.solve() && solver.getNumIters() < 20 && all_sign(*x, 1.0));
  VT::putScalar(*const_cast<VectorType*>(b.get()), -2.0);
  VT::putScalar(*x, 0.0);
  CHECK(solver.solve() && solver.getNumIters() < 20 && all_sign(*x, -1.0));
  auto mlist = parameterList();
  mlist->set<double>("Sample Ratio", 10.0);
  mlist->set<int>("History Length", 10);
  MonteCarloSolverManager<VectorType, MatrixType, Tag> mc(problem, mlist, 0);
  CHECK(mc.solve() && all_sign(*x, -1.0));
  std::printf("ok %s with PL = PR = I iters=%d\n", name, solver.getNumIters());
}

// A sparse (block-inverse) PL and a non-diagonal PR: MCSA on PL A PR y =
// PL b, x = PR y, checked by the unpreconditioned residual on the host.
static void test_sparse_prec() {
  const int n = 12;
  auto A = build_operator(n, 2.0, -0.5);
  // PL: inverse of the 2 x 2 diagonal blocks of A
  auto PL = std::make_shared<MatrixType>(n);
  for (int k = 0; k < n; k += 2) {
    auto a = [&](int i, int j) {
      for (int64_t p = A->rowPtr()[(size_t)i]; p < A->rowPtr()[(size_t)i + 1]; ++p)
        if (A->colInd()[(size_t)p] == j) return A->values()[(size_t)p];
      return 0.0;
    };
    const double a00 = a(k, k), a01 = a(k, k + 1), a10 = a(k + 1, k), a11 = a(k + 1, k + 1);
    const double det = a00 * a11 - a01 * a10;
    PL->insertGlobalValues(k, {k, k + 1}, {a11 / det, -a01 / det});
    PL->insertGlobalValues(k + 1, {k, k + 1}, {-a10 / det, a00 / det});
  }
  PL->fillComplete();
  auto PR = std::make_shared<MatrixType>(n);
  for (int i = 0; i < n; ++i) {
    if (i + 1 < n)
      PR->insertGlobalValues(i, {i, i + 1}, {0.9 + 0.02 * i, 0.03});
    else
      PR->insertGlobalValues(i, {i}, {0.9 + 0.02 * i});
  }
  PR->fillComplete();
  auto x = MT::cloneVectorFromMatrixRows(*A);
  VT::putScalar(*x, 0.0);
  auto b = MT::cloneVectorFromMatrixRows(*A);
  for (int i = 0; i < n; ++i) (*b)[i] = 1.0 + 0.1 * i;
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-10);
  plist->set<int>("Maximum Iterations", 40);
  plist->set<double>("Sample Ratio", 200.0);
  plist->set<int>("Random Seed", 9);
  auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
  problem->setLeftPrec(PL);
  problem->setRightPrec(PR);
  MCSASolverManager<VectorType, MatrixType, AdjointTag> solver(problem, plist);
  CHECK(solver.solve() && solver.achievedTol() <= 1e-10);
  double r2 = 0.0, b2 = 0.0;
  for (int i = 0; i < n; ++i) {
    double ax = 0.0;
    for (int64_t k = A->rowPtr()[(size_t)i]; k < A->rowPtr()[(size_t)i + 1]; ++k)
      ax += A->values()[(size_t)k] * (*x)[A->colInd()[(size_t)k]];
    r2 += ((*b)[i] - ax) * ((*b)[i] - ax);
    b2 += (*b)[i] * (*b)[i];
  }
  CHECK(std::sqrt(r2 / b2) <= 1e-8);
  std::printf("ok MCSASolverManager with sparse PL and PR iters=%d\n", solver.getNumIters());
}

// The device numbers the unknowns its own way (odd ones first, reversed);
// the solution comes back in the caller's numbering and matches the
// unordered solve, also with a left preconditioner and after the order is
// dropped again.
static void test_state_order() {
  const int n = 40;
  auto A = build_operator(n, 2.0, -0.5);
  auto b = MT::cloneVectorFromMatrixRows(*A);
  for (int i = 0; i < n; ++i) (*b)[i] = 1.0 + 0.05 * i;
  std::vector<int32_t> order;
  for (int i = n - 1; i >= 0; --i)
    if (i % 2) order.push_back(i);
  for (int i = n - 1; i >= 0; --i)
    if (i % 2 == 0) order.push_back(i);
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-10);
  plist->set<int>("Maximum Iterations", 40);
  plist->set<double>("Sample Ratio", 100.0);
  plist->set<int>("Random Seed", 5);
  auto solve = [&](bool ordered, bool unorder_after) {
    auto x = MT::cloneVectorFromMatrixRows(*A);
    VT::putScalar(*x, 0.0);
    auto problem = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, x, b);
    PointJacobiPreconditioner<double> pj;
    pj.setOperator(A);
    pj.buildPreconditioner();
    problem->setLeftPrec(pj.getLeftPreconditioner());
    MCSASolverManager<VectorType, MatrixType, AdjointTag> solver(problem, plist);
    if (ordered) solver.setStateOrder(order);
    CHECK(solver.solve() && solver.achievedTol() <= 1e-10);
    if (unorder_after) {
      solver.setStateOrder({});
      VT::putScalar(*x, 0.0);
      CHECK(solver.solve() && solver.achievedTol() <= 1e-10);
    }
    return x;
  };
  auto x0 = solve(false, false), x1 = solve(true, false), x2 = solve(true, true);
  double d1 = 0.0, d2 = 0.0, nx = 0.0;
  for (int i = 0; i < n; ++i) {
    d1 = std::max(d1, std::fabs((*x1)[i] - (*x0)[i]));
    d2 = std::max(d2, std::fabs((*x2)[i] - (*x0)[i]));
    nx = std::max(nx, std::fabs((*x0)[i]));
  }
  CHECK(d1 <= 1e-8 * nx && d2 <= 1e-8 * nx);
  std::printf("ok state order: max diff %.2e / %.2e of %.2e\n", d1, d2, nx);
}

// ---------------------------------------------------------------------------
// A second matrix type, shaped like Tpetra::CrsMatrix / Tpetra::Vector: rows
// inserted by global index, fillComplete, and the local rows read back as a
// (row_map, entries, values) view the way getLocalMatrix() returns Kokkos'
// CRS arrays.  The library knows nothing of it: the two specialisations
// below are the whole binding, what a Tpetra build would add.
namespace mock {

struct LocalMatrix {
  std::vector<int64_t> row_map;
  std::vector<int32_t> entries;
  std::vector<double> values;
};

class TpetraLikeCrsMatrix {
 public:
  explicit TpetraLikeCrsMatrix(int64_t n) : d_rows((size_t)n) {}
  void insertGlobalValues(int64_t row, const std::vector<int64_t>& cols, const std::vector<double>& vals) {
    for (size_t k = 0; k < cols.size(); ++k) d_rows[(size_t)row].emplace_back(cols[k], vals[k]);
  }
  void fillComplete() {
    d_local.row_map.assign(1, 0);
    for (auto& r : d_rows) {
      std::sort(r.begin(), r.end());
      for (size_t k = 0; k < r.size(); ++k) {
        if (k && r[k].first == r[k - 1].first) {
          d_local.values.back() += r[k].second;
          continue;
        }
        d_local.entries.push_back((int32_t)r[k].first);
        d_local.values.push_back(r[k].second);
      }
      d_local.row_map.push_back((int64_t)d_local.entries.size());
    }
    d_filled = true;
  }
  bool isFillComplete() const { return d_filled; }
  int64_t getNodeNumRows() const { return (int64_t)d_rows.size(); }
  const LocalMatrix& getLocalMatrix() const { return d_local; }

 private:
  std::vector<std::vector<std::pair<int64_t, double>>> d_rows;
  LocalMatrix d_local;
  bool d_filled = false;
};

class TpetraLikeVector {
 public:
  explicit TpetraLikeVector(int64_t n) : d_v((size_t)n, 0.0) {}
  int64_t getLocalLength() const { return (int64_t)d_v.size(); }
  const double* getData() const { return d_v.data(); }
  double* getDataNonConst() { return d_v.data(); }

 private:
  std::vector<double> d_v;
};

}  // namespace mock

namespace MCLS {
template <>
class VectorTraits<mock::TpetraLikeVector> {
 public:
  using V = mock::TpetraLikeVector;
  static RCP<V> clone(const V& v) { return std::make_shared<V>(v.getLocalLength()); }
  static RCP<V> deepCopy(const V& v) { return std::make_shared<V>(v); }
  static int64_t getLocalLength(const V& v) { return v.getLocalLength(); }
  static const double* view(const V& v) { return v.getData(); }
  static double* viewNonConst(V& v) { return v.getDataNonConst(); }
  static void putScalar(V& v, double a) { std::fill(v.getDataNonConst(), v.getDataNonConst() + v.getLocalLength(), a); }
  static void scale(V& v, double a) {
    for (int64_t i = 0; i < v.getLocalLength(); ++i) v.getDataNonConst()[i] *= a;
  }
  static void update(V& A, double alpha, const V& B, double beta) {
    for (int64_t i = 0; i < A.getLocalLength(); ++i)
      A.getDataNonConst()[i] = (alpha == 0.0 ? 0.0 : alpha * A.getData()[i]) + beta * B.getData()[i];
  }
  static double normInf(const V& v) {
    double m = 0.0;
    for (int64_t i = 0; i < v.getLocalLength(); ++i) m = std::max(m, std::fabs(v.getData()[i]));
    return m;
  }
};

template <>
class MatrixTraits<mock::TpetraLikeVector, mock::TpetraLikeCrsMatrix> {
 public:
  using M = mock::TpetraLikeCrsMatrix;
  using V = mock::TpetraLikeVector;
  static RCP<V> cloneVectorFromMatrixRows(const M& A) { return std::make_shared<V>(A.getNodeNumRows()); }
  static bool isFillComplete(const M& A) { return A.isFillComplete(); }
  static int64_t getLocalNumRows(const M& A) { return A.getNodeNumRows(); }
  static void getLocalRowCopy(const M& A, int64_t row, std::vector<int64_t>& cols, std::vector<double>& vals,
                              size_t& n) {
    const auto& L = A.getLocalMatrix();
    n = (size_t)(L.row_map[(size_t)row + 1] - L.row_map[(size_t)row]);
    cols.assign(L.entries.begin() + L.row_map[(size_t)row], L.entries.begin() + L.row_map[(size_t)row + 1]);
    vals.assign(L.values.begin() + L.row_map[(size_t)row], L.values.begin() + L.row_map[(size_t)row + 1]);
  }
  static void apply(const M& A, const V& x, V& y) {
    const auto& L = A.getLocalMatrix();
    for (int64_t i = 0; i < A.getNodeNumRows(); ++i) {
      double s = 0.0;
      for (int64_t k = L.row_map[(size_t)i]; k < L.row_map[(size_t)i + 1]; ++k)
        s += L.values[(size_t)k] * x.getData()[L.entries[(size_t)k]];
      y.getDataNonConst()[i] = s;
    }
  }
  // the local CRS arrays as they are (Tpetra: getLocalMatrix().graph.row_map, entries, values)
  static LocalCrs getLocalCrs(const M& A) {
    const auto& L = A.getLocalMatrix();
    LocalCrs c;
    c.num_rows = A.getNodeNumRows();
    c.nnz = (int64_t)L.entries.size();
    c.row_ptr = L.row_map.data();
    c.col = L.entries.data();
    c.val = L.values.data();
    return c;
  }
};
}  // namespace MCLS

// a Preconditioner implementation for the second type: point Jacobi
class MockPointJacobi : public Preconditioner<mock::TpetraLikeCrsMatrix> {
 public:
  using M = mock::TpetraLikeCrsMatrix;
  RCP<const ParameterList> getValidParameters() const override { return parameterList(); }
  RCP<const ParameterList> getCurrentParameters() const override { return parameterList(); }
  void setParameters(const RCP<ParameterList>&) override {}
  void setOperator(const RCP<const M>& A) override { d_A = A; }
  const M& getOperator() const override { return *d_A; }
  void buildPreconditioner() override {
    const auto& L = d_A->getLocalMatrix();
    auto P = std::make_shared<M>(d_A->getNodeNumRows());
    for (int64_t i = 0; i < d_A->getNodeNumRows(); ++i)
      for (int64_t k = L.row_map[(size_t)i]; k < L.row_map[(size_t)i + 1]; ++k)
        if (L.entries[(size_t)k] == i) P->insertGlobalValues(i, {i}, {1.0 / L.values[(size_t)k]});
    P->fillComplete();
    d_P = P;
  }
  RCP<const M> getLeftPreconditioner() const override { return d_P; }
  RCP<const M> getRightPreconditioner() const override { return nullptr; }

 private:
  RCP<const M> d_A, d_P;
};

// Both managers on the second matrix type give what they give on the
// facade's own CrsMatrix for the same operator, right-hand side and streams.
static void test_second_matrix_type() {
  using MV = mock::TpetraLikeVector;
  using MM = mock::TpetraLikeCrsMatrix;
  const int n = 30;
  auto A1 = build_operator(n, 2.0, -0.5);
  auto A2 = std::make_shared<MM>(n);
  for (int64_t i = 0; i < n; ++i) {
    std::vector<int64_t> c;
    std::vector<double> v;
    for (int64_t k = A1->rowPtr()[(size_t)i]; k < A1->rowPtr()[(size_t)i + 1]; ++k) {
      c.push_back(A1->colInd()[(size_t)k]);
      v.push_back(A1->values()[(size_t)k]);
    }
    A2->insertGlobalValues(i, c, v);
  }
  A2->fillComplete();
  auto plist = parameterList();
  plist->set<double>("Convergence Tolerance", 1.0e-10);
  plist->set<int>("Maximum Iterations", 40);
  plist->set<double>("Sample Ratio", 50.0);
  plist->set<int>("Random Seed", 77);
  // the facade's own types
  auto b1 = MT::cloneVectorFromMatrixRows(*A1);
  for (int i = 0; i < n; ++i) (*b1)[i] = 1.0 - 0.02 * i;
  auto x1 = MT::cloneVectorFromMatrixRows(*A1);
  auto p1 = std::make_shared<LinearProblem<VectorType, MatrixType>>(A1, x1, b1);
  PointJacobiPreconditioner<double> pj1;
  pj1.setOperator(A1);
  pj1.buildPreconditioner();
  p1->setLeftPrec(pj1.getLeftPreconditioner());
  MCSASolverManager<VectorType, MatrixType, AdjointTag> s1(p1, plist);
  CHECK(s1.solve());
  // the mock types, touched by the library only through the traits
  using MTM = MatrixTraits<MV, MM>;
  auto b2 = MTM::cloneVectorFromMatrixRows(*A2);
  for (int i = 0; i < n; ++i) b2->getDataNonConst()[i] = 1.0 - 0.02 * i;
  auto x2 = MTM::cloneVectorFromMatrixRows(*A2);
  auto p2 = std::make_shared<LinearProblem<MV, MM>>(A2, x2, b2);
  MockPointJacobi pj2;
  pj2.setOperator(A2);
  pj2.buildPreconditioner();
  p2->setLeftPrec(pj2.getLeftPreconditioner());
  MCSASolverManager<MV, MM, AdjointTag> s2(p2, plist);
  CHECK(s2.solve() && s2.achievedTol() <= 1e-10);
  CHECK(std::abs(s2.getNumIters() - s1.getNumIters()) <= 1);
  double d = 0.0, nx = 0.0;
  for (int i = 0; i < n; ++i) {
    d = std::max(d, std::fabs(x2->getData()[i] - (*x1)[i]));
    nx = std::max(nx, std::fabs((*x1)[i]));
  }
  CHECK(d <= 1e-8 * nx);
  // the host residual of the second type (updatePrecResidual through MT::apply)
  p2->updatePrecResidual();
  CHECK(VectorTraits<MV>::normInf(*p2->getPrecResidual()) <= 1e-8 * VectorTraits<MV>::normInf(*b2));
  // the MC manager (same source, same solve number on a fresh context)
  auto y1 = MT::cloneVectorFromMatrixRows(*A1);
  auto y2 = MTM::cloneVectorFromMatrixRows(*A2);
  auto q1 = std::make_shared<LinearProblem<VectorType, MatrixType>>(A1, y1, b1);
  auto q2 = std::make_shared<LinearProblem<MV, MM>>(A2, y2, b2);
  q1->setLeftPrec(pj1.getLeftPreconditioner());
  q2->setLeftPrec(pj2.getLeftPreconditioner());
  MonteCarloSolverManager<VectorType, MatrixType, AdjointTag> m1(q1, plist);
  MonteCarloSolverManager<MV, MM, AdjointTag> m2(q2, plist);
  CHECK(m1.solve() && m2.solve());
  d = 0.0;
  nx = 0.0;
  for (int i = 0; i < n; ++i) {
    d = std::max(d, std::fabs(y2->getData()[i] - (*y1)[i]));
    nx = std::max(nx, std::fabs((*y1)[i]));
  }
  CHECK(d <= 1e-12 * nx);
  std::printf("ok second matrix type: MCSA %d / %d iterations, solution diff %.2e\n", s2.getNumIters(),
              s1.getNumIters(), d);
}

// An internal MC solver (MCSA's inner one) takes its right-hand side as
// already preconditioned: omega b, no left preconditioning of the source
// (src/MCLS_MonteCarloSolverManager_impl.hpp:347-380).
static void test_internal_mc_solver() {
  const int n = 20;
  auto A = build_operator(n, 2.0, -0.5);
  PointJacobiPreconditioner<double> pj;
  pj.setOperator(A);
  pj.buildPreconditioner();
  auto b = MT::cloneVectorFromMatrixRows(*A);
  for (int i = 0; i < n; ++i) (*b)[i] = 0.5 + 0.1 * i;
  auto plist = parameterList();
  plist->set<double>("Sample Ratio", 200.0);
  plist->set<int>("Random Seed", 9);
  plist->set<double>("Neumann Relaxation", 1.0);
  // internal: source b (taken as preconditioned); external with PL = I: source b
  auto xi = MT::cloneVectorFromMatrixRows(*A), xe = MT::cloneVectorFromMatrixRows(*A);
  auto pi = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, xi, b);
  pi->setLeftPrec(pj.getLeftPreconditioner());
  MonteCarloSolverManager<VectorType, MatrixType, AdjointTag> si(pi, plist, 0, true);
  CHECK(si.isInternalSolver() && si.solve());
  // the same operator's domain with the source PL^-1 b makes the external
  // solver's source PL PL^-1 b = b
  auto bs = MT::cloneVectorFromMatrixRows(*A);
  Vector d(n);
  MT::getLocalDiagCopy(*A, d);
  for (int i = 0; i < n; ++i) (*bs)[i] = (*b)[i] * d[i];
  auto pe = std::make_shared<LinearProblem<VectorType, MatrixType>>(A, xe, bs);
  pe->setLeftPrec(pj.getLeftPreconditioner());
  MonteCarloSolverManager<VectorType, MatrixType, AdjointTag> se(pe, plist);
  CHECK(se.solve());
  double diff = 0.0, nx = 0.0;
  for (int i = 0; i < n; ++i) {
    diff = std::max(diff, std::fabs((*xi)[i] - (*xe)[i]));
    nx = std::max(nx, std::fabs((*xe)[i]));
  }
  // same source up to the rounding of b d / d; same streams (solve 0)
  CHECK(diff <= 1e-12 * nx && nx > 0.0);
  std::printf("ok internal MC solver: diff %.2e of %.2e\n", diff, nx);
}

int main() {
  test_second_matrix_type();
  test_internal_mc_solver();
  test_mcsa_adjoint(false);
  test_mcsa_adjoint(true);
  test_mc_adjoint();
  test_mc_forward();
  test_factory_and_errors();
  test_matrix_market_system();
  test_mcsa_importance();
  test_identity_prec<AdjointTag>("MCSASolverManager.prec_adjoint");
  test_identity_prec<ForwardTag>("MCSASolverManager.prec_forward");
  test_sparse_prec();
  test_state_order();
  std::printf("ALL %d checks passed\n", g_checks);
  return 0;
}
