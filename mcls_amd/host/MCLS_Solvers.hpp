// MCLS solver API over the gfx950 C ABI (include/mcls_gpu.h).
//
// Same class names, constructor shapes, parameter keys and error convention
// as MCLS (paths relative to the MCLS source root):
//   LinearProblem                     src/MCLS_LinearProblem.hpp:59-186
//   SolverManager                     src/MCLS_SolverManager.hpp:60-113
//   MonteCarloSolverManager<..,Tag>   src/MCLS_MonteCarloSolverManager.hpp:67-220
//   MCSASolverManager<..,Tag>         src/MCLS_MCSASolverManager.hpp:70-181
//   (Tag = AdjointTag or ForwardTag, the reference's two methods)
//   MultiSetLinearProblem (1 set)     src/MCLS_MultiSetLinearProblem.hpp
//   SolverFactory                     src/MCLS_SolverFactory_impl.hpp:58-146
//   TpetraPointJacobiPreconditioner   tpetra/src/MCLS_TpetraPointJacobiPreconditioner_impl.hpp:105-139
//   Preconditioner                    src/MCLS_Preconditioner.hpp:59-101
// The managers and the linear problem reach the operator, the
// preconditioners and the vectors only through MatrixTraits / VectorTraits
// (MCLS_Traits.hpp): another matrix type plugs in by specialising them.
// Non-zero C ABI status is rethrown as MCLS::Assertion.  The domain is
// rebuilt only when the operator object changes (pointer identity, as
// src/MCLS_MonteCarloSolverManager_impl.hpp:161-178).  A diagonal left
// preconditioner goes to the device as its diagonal; any other left or right
// preconditioner as a sparse matrix, and the domain is then built from the
// composite operator (src/MCLS_LinearProblem_impl.hpp:139-221).  Anderson
// and the stand-alone fixed-point manager are not provided by this build and
// raise MCLS::Assertion.
#pragma once

#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "MCLS_Core.hpp"
#include "MCLS_Traits.hpp"
#include "mcls_gpu.h"

namespace MCLS {

class ForwardTag {};
class AdjointTag {};

// RAII owner of one mcls_gpu_context.
class GpuContext {
 public:
  explicit GpuContext(int device = 0) { check(mcls_gpu_create(&d_h, device, 0)); }
  ~GpuContext() { mcls_gpu_destroy(d_h); }
  GpuContext(const GpuContext&) = delete;
  GpuContext& operator=(const GpuContext&) = delete;
  mcls_gpu_context* get() const { return d_h; }
  static void check(int rc) {
    if (rc != MCLS_OK) throw Assertion(std::string("mcls_gpu error ") + std::to_string(rc) + ": " + mcls_gpu_last_error());
  }
  // Forward a ParameterList; keys the device solver does not know are
  // ignored, as Teuchos ignores unused entries.
  void setParameters(const ParameterList& pl) {
    for (const auto& kv : pl.values()) {
      const char* k = kv.first.c_str();
      int rc = MCLS_OK;
      if (auto* i = std::get_if<int>(&kv.second)) rc = mcls_gpu_set_param_i64(d_h, k, *i);
      else if (auto* l = std::get_if<int64_t>(&kv.second)) rc = mcls_gpu_set_param_i64(d_h, k, *l);
      else if (auto* b = std::get_if<bool>(&kv.second)) rc = mcls_gpu_set_param_i64(d_h, k, *b ? 1 : 0);
      else if (auto* d = std::get_if<double>(&kv.second)) rc = mcls_gpu_set_param_f64(d_h, k, *d);
      else if (auto* s = std::get_if<std::string>(&kv.second)) rc = mcls_gpu_set_param_str(d_h, k, s->c_str());
      if (rc == MCLS_ERR_INVALID_ARG && std::strncmp(mcls_gpu_last_error(), "unknown", 7) == 0) continue;
      check(rc);
    }
  }

 private:
  mcls_gpu_context* d_h = nullptr;
};

template <class Vector, class Matrix>
class LinearProblem {
 public:
  using VT = VectorTraits<Vector>;
  using MT = MatrixTraits<Vector, Matrix>;
  LinearProblem(const RCP<const Matrix>& A, const RCP<Vector>& x, const RCP<const Vector>& b)
      : d_A(A), d_x(x), d_b(b) {
    MCLS_INSIST(A && x && b, "null operator or vector");
    d_r = VT::clone(*x);
    d_rp = VT::clone(*x);
  }
  void setOperator(const RCP<const Matrix>& A) { d_A = A; }
  void setLHS(const RCP<Vector>& x) { d_x = x; }
  void setRHS(const RCP<const Vector>& b) { d_b = b; }
  void setLeftPrec(const RCP<const Matrix>& PL) { d_PL = PL; }
  void setRightPrec(const RCP<const Matrix>& PR) { d_PR = PR; }
  const RCP<const Matrix>& getOperator() const { return d_A; }
  const RCP<Vector>& getLHS() const { return d_x; }
  const RCP<const Vector>& getRHS() const { return d_b; }
  const RCP<const Matrix>& getLeftPrec() const { return d_PL; }
  const RCP<const Matrix>& getRightPrec() const { return d_PR; }
  bool isLeftPrec() const { return (bool)d_PL; }
  bool isRightPrec() const { return (bool)d_PR; }
  const RCP<Vector>& getResidual() const { return d_r; }
  const RCP<Vector>& getPrecResidual() const { return d_rp; }
  // PL (A PR) and PR^T (A^T PL^T), each product filtered by `threshold`, on
  // the host through the traits (src/MCLS_LinearProblem_impl.hpp:139-221);
  // the device builds the same operators itself for the walks.
  RCP<const Matrix> getCompositeOperator(double threshold = 0.0) const {
    if (d_PL && d_PR) return MT::multiply(d_PL, false, MT::multiply(d_A, false, d_PR, false, threshold), false, threshold);
    if (d_PL) return MT::multiply(d_PL, false, d_A, false, threshold);
    if (d_PR) return MT::multiply(d_A, false, d_PR, false, threshold);
    return d_A;
  }
  RCP<const Matrix> getTransposeCompositeOperator(double threshold = 0.0) const {
    RCP<const Matrix> AT = MT::copyTranspose(*d_A);
    if (d_PL && d_PR) {
      RCP<const Matrix> t = MT::multiply(AT, false, RCP<const Matrix>(MT::copyTranspose(*d_PL)), false, threshold);
      return MT::multiply(RCP<const Matrix>(MT::copyTranspose(*d_PR)), false, t, false, threshold);
    }
    if (d_PR) return MT::multiply(RCP<const Matrix>(MT::copyTranspose(*d_PR)), false, AT, false, threshold);
    if (d_PL) return MT::multiply(AT, false, RCP<const Matrix>(MT::copyTranspose(*d_PL)), false, threshold);
    return AT;
  }
  // y = PL x, y = PR x, y = A x; the identity when that operator is absent
  // (src/MCLS_LinearProblem_impl.hpp:332-386)
  void applyLeftPrec(const Vector& x, Vector& y) const { applyOrCopy(d_PL, x, y); }
  void applyRightPrec(const Vector& x, Vector& y) const { applyOrCopy(d_PR, x, y); }
  void applyOp(const Vector& x, Vector& y) const { applyOrCopy(d_A, x, y); }
  // y = PL A PR x, the composite operator (src/MCLS_LinearProblem_impl.hpp:252-286)
  void apply(const Vector& x, Vector& y) const {
    auto t = VT::clone(y);
    applyRightPrec(x, *t);
    auto u = VT::clone(y);
    MT::apply(*d_A, *t, *u);
    applyLeftPrec(*u, y);
  }
  // y = (PL A PR)^T x = PR^T A^T PL^T x (src/MCLS_LinearProblem_impl.hpp:292-326)
  void applyTranspose(const Vector& x, Vector& y) const {
    auto t = VT::clone(y);
    if (d_PL) MT::applyTranspose(*d_PL, x, *t);
    else VT::update(*t, 0.0, x, 1.0);
    auto u = VT::clone(y);
    MT::applyTranspose(*d_A, *t, *u);
    if (d_PR) MT::applyTranspose(*d_PR, *u, y);
    else VT::update(y, 0.0, *u, 1.0);
  }
  // x += PR update (src/MCLS_LinearProblem_impl.hpp:228-245)
  void updateSolution(const RCP<Vector>& update) {
    auto t = VT::clone(*update);
    applyRightPrec(*update, *t);
    VT::update(*d_x, 1.0, *t, 1.0);
  }
  // r = b - A x  (src/MCLS_LinearProblem_impl.hpp:390-403)
  void updateResidual() {
    MT::apply(*d_A, *d_x, *d_r);
    VT::update(*d_r, -1.0, *d_b, 1.0);
  }
  // rp = PL (b - A PR x)  (src/MCLS_LinearProblem_impl.hpp:411-442), x the
  // right-preconditioned unknown
  void updatePrecResidual() {
    if (d_PR) {
      auto t = VT::clone(*d_rp);
      MT::apply(*d_PR, *d_x, *t);
      MT::apply(*d_A, *t, *d_rp);
    } else {
      MT::apply(*d_A, *d_x, *d_rp);
    }
    VT::update(*d_rp, -1.0, *d_b, 1.0);
    if (d_PL) {
      auto tmp = VT::deepCopy(*d_rp);
      MT::apply(*d_PL, *tmp, *d_rp);
    }
  }

 private:
  static void applyOrCopy(const RCP<const Matrix>& M, const Vector& x, Vector& y) {
    if (M) MT::apply(*M, x, y);
    else VT::update(y, 0.0, x, 1.0);
  }
  RCP<const Matrix> d_A;
  RCP<Vector> d_x;
  RCP<const Vector> d_b;
  RCP<const Matrix> d_PL, d_PR;
  RCP<Vector> d_r, d_rp;
};

// One set: the build runs on one GPU, so the set sum is the identity.
template <class Vector, class Matrix>
class MultiSetLinearProblem {
 public:
  MultiSetLinearProblem(int /*num_sets*/, int /*set_id*/, const RCP<LinearProblem<Vector, Matrix>>& p)
      : d_problem(p) {}
  const RCP<LinearProblem<Vector, Matrix>>& getProblem() const { return d_problem; }
  int numSets() const { return 1; }
  int globalRank() const { return 0; }

 private:
  RCP<LinearProblem<Vector, Matrix>> d_problem;
};

template <class Vector, class Matrix>
class SolverManager {
 public:
  virtual ~SolverManager() = default;
  virtual const LinearProblem<Vector, Matrix>& getProblem() const = 0;
  virtual RCP<const ParameterList> getValidParameters() const = 0;
  virtual RCP<const ParameterList> getCurrentParameters() const = 0;
  virtual double achievedTol() const = 0;
  virtual int getNumIters() const = 0;
  virtual void setProblem(const RCP<LinearProblem<Vector, Matrix>>& problem) = 0;
  virtual void setParameters(const RCP<ParameterList>& params) = 0;
  virtual bool solve() = 0;
  virtual bool getConvergedStatus() const = 0;
};

namespace detail {

// Shared device plumbing of the two managers.
template <class Vector, class Matrix>
class DeviceProblem {
 public:
  using Problem = LinearProblem<Vector, Matrix>;
  using MT = MatrixTraits<Vector, Matrix>;
  DeviceProblem() : d_ctx(std::make_shared<GpuContext>(0)) {}

  // Upload the operator when its object changed; refresh PL and PR when
  // their objects changed.
  // Every access to a matrix goes through MT (fill state, local rows, the
  // local CSR arrays), so any matrix type with MatrixTraits works here.
  void sync(const Problem& p, const ParameterList& pl) {
    const Matrix& A = *p.getOperator();
    MCLS_INSIST(MT::isFillComplete(A), "operator must be fill-complete");
    const int64_t n = (int64_t)MT::getLocalNumRows(A);
    if (d_order_dirty) {
      // a new numbering drops the device operator and what was set for it
      MCLS_INSIST(d_order.empty() || (int64_t)d_order.size() == n, "state order has the wrong length");
      GpuContext::check(mcls_gpu_set_state_order(d_ctx->get(), (int64_t)d_order.size(),
                                                 d_order.empty() ? nullptr : d_order.data(), MCLS_HOST));
      d_op = nullptr;
      d_order_dirty = false;
    }
    if (d_op != p.getOperator().get()) {
      const LocalCrs c = MT::getLocalCrs(A);
      MCLS_INSIST(c.num_rows == n, "local CSR rows differ from getLocalNumRows");
      GpuContext::check(mcls_gpu_set_matrix_csr(d_ctx->get(), c.num_rows, c.nnz, c.row_ptr, c.col, c.val,
                                                c.f32 ? MCLS_F32 : MCLS_F64, MCLS_HOST));
      d_op = p.getOperator().get();
      d_pl = nullptr;
      d_pr = nullptr;
      d_imp_dirty = true;  // a new operator drops the context's importance
    }
    if (p.isLeftPrec()) {
      const Matrix& PL = *p.getLeftPrec();
      if (d_pl != &PL) {
        MCLS_INSIST(MT::isFillComplete(PL) && (int64_t)MT::getLocalNumRows(PL) == n,
                    "left preconditioner must match the operator");
        const LocalCrs c = MT::getLocalCrs(PL);
        std::vector<double> d;
        if (crsIsDiagonal(c, &d)) {
          GpuContext::check(mcls_gpu_set_left_prec_diag(d_ctx->get(), d.data(), MCLS_HOST));
        } else {
          const std::vector<double> v = values64(c);
          GpuContext::check(mcls_gpu_set_left_prec_matrix(d_ctx->get(), c.row_ptr, c.col, v.data(), MCLS_HOST));
        }
        d_pl = &PL;
      }
    } else if (d_pl) {
      GpuContext::check(mcls_gpu_clear_left_prec(d_ctx->get()));
      d_pl = nullptr;
    }
    if (p.isRightPrec()) {
      const Matrix& PR = *p.getRightPrec();
      if (d_pr != &PR) {
        MCLS_INSIST(MT::isFillComplete(PR) && (int64_t)MT::getLocalNumRows(PR) == n,
                    "right preconditioner must match the operator");
        const LocalCrs c = MT::getLocalCrs(PR);
        const std::vector<double> v = values64(c);
        GpuContext::check(mcls_gpu_set_right_prec_matrix(d_ctx->get(), c.row_ptr, c.col, v.data(), MCLS_HOST));
        d_pr = &PR;
      }
    } else if (d_pr) {
      GpuContext::check(mcls_gpu_clear_right_prec(d_ctx->get()));
      d_pr = nullptr;
    }
    d_ctx->setParameters(pl);
    if (d_imp_dirty && d_imp_mode != ImportanceMode::None) {
      if (d_imp_mode == ImportanceMode::Diagonal) {
        GpuContext::check(mcls_gpu_set_importance_diagonal(d_ctx->get()));
      } else {
        MCLS_INSIST((int64_t)d_imp.size() == n, "importance has the wrong length");
        GpuContext::check(mcls_gpu_set_importance(d_ctx->get(), d_imp.data(), MCLS_HOST));
      }
    }
    d_imp_dirty = false;
  }
  // Importance of the learned proposal ("Transition Sampler" = "Learned"),
  // applied at the next solve and kept across operator changes.
  void setImportanceDiagonal() {
    d_imp_mode = ImportanceMode::Diagonal;
    d_imp_dirty = true;
  }
  void setImportance(std::vector<float> g) {
    d_imp = std::move(g);
    d_imp_mode = ImportanceMode::Explicit;
    d_imp_dirty = true;
  }
  void setImportanceModel(mcls_mlp* m) { GpuContext::check(mcls_gpu_set_importance_model(d_ctx->get(), m)); }
  // The device's numbering of the unknowns (mcls_gpu_set_state_order),
  // applied at the next solve; vectors keep the caller's numbering.  An
  // empty order returns to the caller's numbering.
  void setStateOrder(std::vector<int32_t> order) {
    d_order = std::move(order);
    d_order_dirty = true;
  }
  mcls_gpu_context* ctx() const { return d_ctx->get(); }
  mcls_gpu_stats stats() const {
    mcls_gpu_stats s;
    GpuContext::check(mcls_gpu_get_stats(d_ctx->get(), &s));
    return s;
  }

 private:
  enum class ImportanceMode { None, Diagonal, Explicit };
  RCP<GpuContext> d_ctx;
  static std::vector<double> values64(const LocalCrs& c) {
    std::vector<double> v((size_t)c.nnz);
    for (int64_t k = 0; k < c.nnz; ++k) v[(size_t)k] = c.value(k);
    return v;
  }
  const void* d_op = nullptr;
  const void* d_pl = nullptr;
  const void* d_pr = nullptr;
  ImportanceMode d_imp_mode = ImportanceMode::None;
  std::vector<float> d_imp;
  bool d_imp_dirty = false;
  std::vector<int32_t> d_order;
  bool d_order_dirty = false;
};

inline RCP<ParameterList> mcValidParameters() {
  // src/MCLS_MonteCarloSolverManager_impl.hpp:99-110
  auto pl = parameterList();
  pl->set<double>("Sample Ratio", 1.0);
  pl->set<double>("History Length", 10);
  pl->set<int>("MC Check Frequency", 1000);
  pl->set<int>("MC Buffer Size", 1000);
  pl->set<double>("Neumann Relaxation", 1.0);
  return pl;
}

}  // namespace detail

template <class Vector, class Matrix, class MonteCarloTag = AdjointTag>
class MonteCarloSolverManager : public SolverManager<Vector, Matrix> {
  static_assert(std::is_same<MonteCarloTag, AdjointTag>::value || std::is_same<MonteCarloTag, ForwardTag>::value,
                "MonteCarloTag is AdjointTag or ForwardTag");
  static constexpr const char* kMethod = std::is_same<MonteCarloTag, ForwardTag>::value ? "Forward" : "Adjoint";

 public:
  using Problem = LinearProblem<Vector, Matrix>;
  using VT = VectorTraits<Vector>;
  using MT = MatrixTraits<Vector, Matrix>;

  MonteCarloSolverManager(const RCP<ParameterList>& plist, int global_rank = 0, bool internal_solver = false)
      : d_plist(plist), d_internal(internal_solver) {
    (void)global_rank;
    MCLS_INSIST(plist, "null parameter list");
  }
  MonteCarloSolverManager(const RCP<Problem>& problem, const RCP<ParameterList>& plist, int global_rank = 0,
                          bool internal_solver = false)
      : MonteCarloSolverManager(plist, global_rank, internal_solver) {
    setProblem(problem);
  }
  const Problem& getProblem() const override { return *d_problem; }
  RCP<const ParameterList> getValidParameters() const override { return detail::mcValidParameters(); }
  RCP<const ParameterList> getCurrentParameters() const override { return d_plist; }
  // ||b - A x||_inf / ||b||_inf  (src/MCLS_MonteCarloSolverManager_impl.hpp:120-133)
  double achievedTol() const override {
    d_problem->updateResidual();
    return VT::normInf(*d_problem->getResidual()) / VT::normInf(*d_problem->getRHS());
  }
  int getNumIters() const override { return 0; }
  void setProblem(const RCP<Problem>& problem) override {
    MCLS_INSIST(problem, "null problem");
    d_problem = problem;
  }
  void setParameters(const RCP<ParameterList>& params) override {
    MCLS_INSIST(params, "null parameter list");
    d_plist = params;
  }
  // x = MC estimate of the solution for the source omega PL b, and x = PR y
  // when right preconditioned.  An internal solver (the one MCSA drives)
  // takes its right-hand side as already preconditioned: source omega b, no
  // recovery (src/MCLS_MonteCarloSolverManager_impl.hpp:200-235,347-380).
  bool solve() override {
    MCLS_INSIST(d_problem, "no problem set");
    d_dev.sync(*d_problem, *d_plist);
    GpuContext::check(mcls_gpu_set_param_str(d_dev.ctx(), "Monte Carlo Type", kMethod));
    Vector& x = *d_problem->getLHS();
    if (d_internal) {
      const double omega = d_plist->get<double>("Neumann Relaxation", 1.0);
      auto src = VT::deepCopy(*d_problem->getRHS());
      VT::scale(*src, omega);
      GpuContext::check(mcls_gpu_mc_solve_raw(d_dev.ctx(), VT::view(*src), d_internal_solves++, VT::viewNonConst(x)));
    } else {
      GpuContext::check(
          mcls_gpu_mc_solve(d_dev.ctx(), VT::view(*d_problem->getRHS()), VT::viewNonConst(x), MCLS_HOST));
    }
    return true;  // a direct solve (src/MCLS_MonteCarloSolverManager.hpp:170-171)
  }
  bool isInternalSolver() const { return d_internal; }
  bool getConvergedStatus() const override { return true; }
  mcls_gpu_stats deviceStats() const { return d_dev.stats(); }
  // Learned-proposal importance (no counterpart in the reference): the
  // diagonal rule g = |a_jj|, an explicit g, or a network (include/mcls_mlp.h)
  // evaluated on every MC source.  Used with "Transition Sampler" = "Learned".
  void setImportanceDiagonal() { d_dev.setImportanceDiagonal(); }
  void setImportance(std::vector<float> g) { d_dev.setImportance(std::move(g)); }
  void setImportanceModel(mcls_mlp* m) { d_dev.setImportanceModel(m); }
  void setStateOrder(std::vector<int32_t> order) { d_dev.setStateOrder(std::move(order)); }

 private:
  RCP<Problem> d_problem;
  RCP<ParameterList> d_plist;
  bool d_internal;
  uint64_t d_internal_solves = 0;  // solve numbers (random streams) of an internal solver
  detail::DeviceProblem<Vector, Matrix> d_dev;
};

template <class Vector, class Matrix, class MonteCarloTag = AdjointTag>
class MCSASolverManager : public SolverManager<Vector, Matrix> {
  static_assert(std::is_same<MonteCarloTag, AdjointTag>::value || std::is_same<MonteCarloTag, ForwardTag>::value,
                "MonteCarloTag is AdjointTag or ForwardTag");
  static constexpr const char* kMethod = std::is_same<MonteCarloTag, ForwardTag>::value ? "Forward" : "Adjoint";

 public:
  using Problem = LinearProblem<Vector, Matrix>;
  using VT = VectorTraits<Vector>;

  // src/MCLS_MCSASolverManager.hpp:95-97 takes a MultiSetLinearProblem; the
  // plain-problem and parameter-only forms are also accepted (the
  // reference's SolverFactory path, src/MCLS_SolverFactory_impl.hpp:105-107).
  explicit MCSASolverManager(const RCP<ParameterList>& plist) : d_plist(plist) {
    MCLS_INSIST(plist, "null parameter list");
  }
  MCSASolverManager(const RCP<Problem>& problem, const RCP<ParameterList>& plist) : MCSASolverManager(plist) {
    setProblem(problem);
  }
  MCSASolverManager(const RCP<MultiSetLinearProblem<Vector, Matrix>>& ms, const RCP<ParameterList>& plist)
      : MCSASolverManager(ms->getProblem(), plist) {}

  const Problem& getProblem() const override { return *d_problem; }
  RCP<const ParameterList> getValidParameters() const override {
    auto pl = detail::mcValidParameters();
    pl->set<double>("Convergence Tolerance", 1.0);
    pl->set<int>("Maximum Iterations", 1000);
    pl->set<int>("Iteration Print Frequency", 10);
    pl->set<int>("Iteration Check Frequency", 1);
    pl->set<int>("Smoother Steps", 1);
    pl->set<std::string>("Fixed Point Type", "Richardson");
    pl->set<double>("Richardson Relaxation", 1.0);
    return pl;
  }
  RCP<const ParameterList> getCurrentParameters() const override { return d_plist; }
  double achievedTol() const override { return d_stats.achieved_tol; }
  int getNumIters() const override { return (int)d_stats.iterations; }
  void setProblem(const RCP<Problem>& problem) override {
    MCLS_INSIST(problem, "null problem");
    d_problem = problem;
  }
  void setParameters(const RCP<ParameterList>& params) override {
    MCLS_INSIST(params, "null parameter list");
    d_plist = params;
  }
  bool solve() override {
    MCLS_INSIST(d_problem, "no problem set");
    d_dev.sync(*d_problem, *d_plist);
    GpuContext::check(mcls_gpu_set_param_str(d_dev.ctx(), "Monte Carlo Type", kMethod));
    Vector& x = *d_problem->getLHS();
    GpuContext::check(mcls_gpu_mcsa_solve(d_dev.ctx(), VT::view(*d_problem->getRHS()), VT::viewNonConst(x), MCLS_HOST));
    d_stats = d_dev.stats();
    return d_stats.converged != 0;
  }
  bool getConvergedStatus() const override { return d_stats.converged != 0; }
  mcls_gpu_stats deviceStats() const { return d_stats; }
  // Learned-proposal importance (no counterpart in the reference): the
  // diagonal rule g = |a_jj|, an explicit g, or a network (include/mcls_mlp.h)
  // evaluated on every MC source.  Used with "Transition Sampler" = "Learned".
  void setImportanceDiagonal() { d_dev.setImportanceDiagonal(); }
  void setImportance(std::vector<float> g) { d_dev.setImportance(std::move(g)); }
  void setImportanceModel(mcls_mlp* m) { d_dev.setImportanceModel(m); }
  void setStateOrder(std::vector<int32_t> order) { d_dev.setStateOrder(std::move(order)); }
  // Walk launch shape by timing pilot MC solves of PL b (mcls_gpu_autotune);
  // kept by later solves unless the parameter list sets the walk keys.
  mcls_gpu_tuning autotune(int64_t pilot_histories = 0) {
    MCLS_INSIST(d_problem, "no problem set");
    d_dev.sync(*d_problem, *d_plist);
    mcls_gpu_tuning t{};
    GpuContext::check(mcls_gpu_autotune(d_dev.ctx(), VT::view(*d_problem->getRHS()), MCLS_HOST, pilot_histories, &t));
    return t;
  }

 private:
  RCP<Problem> d_problem;
  RCP<ParameterList> d_plist;
  detail::DeviceProblem<Vector, Matrix> d_dev;
  mcls_gpu_stats d_stats{};
};

// src/MCLS_SolverFactory_impl.hpp:58-146 with the same solver names.  The
// factory builds managers from a parameter list alone (setProblem before
// solve), which the reference's factory cannot do for MCSA (the constructor
// it calls leaves the multi-set problem null).
template <class Vector, class Matrix>
class SolverFactory {
 public:
  using Solver = SolverManager<Vector, Matrix>;
  RCP<Solver> create(const std::string& name, const RCP<ParameterList>& params) const {
    MCLS_INSIST(params, "null parameter list");
    if (name == "Adjoint MC") return std::make_shared<MonteCarloSolverManager<Vector, Matrix, AdjointTag>>(params);
    if (name == "Adjoint MCSA") return std::make_shared<MCSASolverManager<Vector, Matrix, AdjointTag>>(params);
    if (name == "Forward MC") return std::make_shared<MonteCarloSolverManager<Vector, Matrix, ForwardTag>>(params);
    if (name == "Forward MCSA") return std::make_shared<MCSASolverManager<Vector, Matrix, ForwardTag>>(params);
    if (name == "Adjoint Anderson" || name == "Forward Anderson" || name == "Fixed Point")
      throw Assertion("Solver type '" + name + "' is not provided by this build");
    throw Assertion("Solver type not supported!");
  }
};

// src/MCLS_Preconditioner.hpp:59-101: the preconditioner plugin interface.
template <class Matrix>
class Preconditioner {
 public:
  using matrix_type = Matrix;
  virtual ~Preconditioner() = default;
  virtual RCP<const ParameterList> getValidParameters() const = 0;
  virtual RCP<const ParameterList> getCurrentParameters() const = 0;
  virtual void setParameters(const RCP<ParameterList>& params) = 0;
  virtual void setOperator(const RCP<const Matrix>& A) = 0;
  virtual const Matrix& getOperator() const = 0;
  virtual void buildPreconditioner() = 0;
  virtual RCP<const Matrix> getLeftPreconditioner() const = 0;
  virtual RCP<const Matrix> getRightPreconditioner() const = 0;
};

// Point Jacobi (tpetra/src/MCLS_TpetraPointJacobiPreconditioner_impl.hpp:
// 105-139): PL = diag(1/a_ii) as a diagonal CrsMatrix, no right
// preconditioner, no parameters.  The diagonal is read through the traits.
// (The device solver can also build it itself: mcls_gpu_set_left_prec_point_jacobi.)
template <class Scalar>
class PointJacobiPreconditioner : public Preconditioner<CrsMatrix<Scalar>> {
 public:
  using Matrix = CrsMatrix<Scalar>;
  using MT = MatrixTraits<Vector, Matrix>;
  PointJacobiPreconditioner() : d_plist(parameterList()) {}
  RCP<const ParameterList> getValidParameters() const override { return parameterList(); }
  RCP<const ParameterList> getCurrentParameters() const override { return d_plist; }
  void setParameters(const RCP<ParameterList>& params) override {
    MCLS_INSIST(params, "null parameter list");
    d_plist = params;
  }
  void setOperator(const RCP<const Matrix>& A) override { d_A = A; }
  const Matrix& getOperator() const override {
    MCLS_INSIST(d_A, "no operator set");
    return *d_A;
  }
  void buildPreconditioner() override {
    MCLS_INSIST(d_A && MT::isFillComplete(*d_A), "operator must be set and fill-complete");
    const int64_t n = MT::getLocalNumRows(*d_A);
    auto d = MT::cloneVectorFromMatrixRows(*d_A);
    MT::getLocalDiagCopy(*d_A, *d);
    std::vector<int64_t> rp((size_t)n + 1);
    std::vector<int32_t> col((size_t)n);
    std::vector<Scalar> val((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      MCLS_INSIST((*d)[i] != 0.0, "zero diagonal in row " + std::to_string(i));
      rp[(size_t)i + 1] = i + 1;
      col[(size_t)i] = (int32_t)i;
      val[(size_t)i] = (Scalar)(1.0 / (*d)[i]);
    }
    d_P = std::make_shared<Matrix>(n, rp, col, val);
  }
  RCP<const Matrix> getLeftPreconditioner() const override { return d_P; }
  RCP<const Matrix> getRightPreconditioner() const override { return nullptr; }

 private:
  RCP<ParameterList> d_plist;
  RCP<const Matrix> d_A;
  RCP<const Matrix> d_P;
};

}  // namespace MCLS
