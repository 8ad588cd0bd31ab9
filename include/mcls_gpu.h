/*
 * mcls_gpu.h -- C ABI of the MI355X (gfx950) adjoint Monte Carlo / MCSA
 * solver.  Plain pointers and sizes only; every call returns an int status
 * (MCLS_OK = 0) and no exception crosses the boundary.  The message of the
 * last failure on the calling thread is returned by mcls_gpu_last_error().
 *
 * One context = one GPU stream holding one linear problem (matrix, optional
 * left preconditioner (diagonal or sparse) and right preconditioner
 * (sparse), the Monte Carlo domain built from them, and the work vectors).  A context is not thread-safe; separate contexts are
 * independent.  Vectors are host or device pointers, chosen by `ptr_kind`.
 *
 * Reference interfaces replaced (paths relative to the MCLS source root):
 *   LinearProblem(A, x, b), setLeftPrec,
 *   setRightPrec                              src/MCLS_LinearProblem.hpp:74-91
 *   TpetraPointJacobiPreconditioner           tpetra/src/MCLS_TpetraPointJacobiPreconditioner_impl.hpp:105-139
 *   AlmostOptimalDomain(A^T composite, x, pl) src/MCLS_AlmostOptimalDomain_impl.hpp:71-191,258-366
 *   MonteCarloSolverManager<..,AdjointTag>    src/MCLS_MonteCarloSolverManager.hpp:103-220
 *   MCSASolverManager<..,AdjointTag>::solve   src/MCLS_MCSASolverManager_impl.hpp:279-432
 *   Teuchos::ParameterList keys               read at the lines cited per key below
 *   SolverManager::{achievedTol,getNumIters,getConvergedStatus}
 *                                             src/MCLS_SolverManager.hpp:60-113
 */
#ifndef MCLS_GPU_H
#define MCLS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define MCLS_OK 0
#define MCLS_ERR_INVALID_ARG 1  /* null pointer, bad size, unknown key     */
#define MCLS_ERR_HIP 2          /* HIP runtime failure                     */
#define MCLS_ERR_MATRIX 3       /* not square, bad/duplicate column, zero
                                   diagonal under point Jacobi, NaN value  */
#define MCLS_ERR_STATE 4        /* call order (no matrix / no domain)      */
#define MCLS_ERR_UNSUPPORTED 5  /* option this build does not provide      */
#define MCLS_ERR_NUMERIC 6      /* NaN/Inf in a vector or residual         */
#define MCLS_ERR_NO_DEVICE 7    /* no HIP device / bad device index        */

/* values dtype of set_matrix_csr */
#define MCLS_F64 0
#define MCLS_F32 1

/* pointer kinds */
#define MCLS_HOST 0
#define MCLS_DEVICE 1

typedef struct mcls_gpu_context mcls_gpu_context;

typedef struct mcls_gpu_stats {
  int64_t iterations;      /* MCSA iterations of the last solve (getNumIters) */
  int64_t converged;       /* getConvergedStatus                               */
  double achieved_tol;     /* ||rp||_2 / ||PL b||_2 (achievedTol)              */
  double source_norm;      /* ||PL b||_2                                       */
  int64_t histories;       /* histories simulated in the last solve            */
  int64_t steps;           /* transitions simulated in the last solve          */
  int64_t walk_launches;   /* walk-kernel launches in the last solve           */
  double ms_total;         /* wall time of the last solve call                 */
  double ms_walk;          /* HIP-event time of the walk kernel                */
  double ms_source;        /* HIP-event time of the source build               */
  double ms_residual;      /* HIP-event time of residual / smoother / update   */
  double ms_domain;        /* HIP-event time of the last domain build          */
  int64_t domain_nnz;      /* entries of the adjoint domain (rows of H^T)      */
  int64_t solve_index;     /* MC solves performed by this context so far       */
  double last_walk_ms;     /* HIP-event time of the most recent walk launch    */
} mcls_gpu_stats;

/* ------------------------------------------------------------- lifetime */
/* Context on `device` with its own HIP stream (default priority). */
int mcls_gpu_create(mcls_gpu_context** ctx, int device, unsigned flags);
int mcls_gpu_destroy(mcls_gpu_context* ctx);
/* Message of the last failure on this thread ("" when none). */
const char* mcls_gpu_last_error(void);
int mcls_gpu_synchronize(mcls_gpu_context* ctx);

/* ------------------------------------------------------------ problem -- */
/* State order (no counterpart in the reference; MCLS walks the operator in
 * the caller's numbering): the context numbers the unknowns its own way,
 * internal state k being the caller's unknown order[k] (a permutation of
 * 0..n-1, host or device array), and works on P A P^T throughout.  An order
 * that keeps the states a walk visits together close in memory (bricks of a
 * structured grid, say) raises the walks' L2 hit rate.  Every vector passed
 * through this API stays in the caller's numbering (gathered on the way in,
 * scattered on the way out, on the context's stream), and so do the states
 * of mcls_gpu_record_paths; mcls_gpu_get_domain and the features' neighbour
 * classes (mcls_gpu_importance_features) are in the internal numbering.
 * Call it before mcls_gpu_set_matrix_csr: it drops the operator and what
 * was set for it (preconditioners, importance, mask, systems).  NULL returns
 * to the caller's numbering.  Batched systems are refused while an order is
 * set (MCLS_ERR_UNSUPPORTED). */
int mcls_gpu_set_state_order(mcls_gpu_context* ctx, int64_t n, const int32_t* order, int ptr_kind);
/* LinearProblem operator A (src/MCLS_LinearProblem.hpp:74-91).  n x n CSR,
 * int64 row pointers, int32 columns (each row's columns unique), values
 * FP64 or FP32 (`val_dtype`).  The library copies the arrays (host or
 * device, `ptr_kind`) and invalidates any built domain.  nnz < 2^31. */
int mcls_gpu_set_matrix_csr(mcls_gpu_context* ctx, int64_t n, int64_t nnz,
                            const int64_t* row_ptr, const int32_t* col,
                            const void* val, int val_dtype, int ptr_kind);
/* LinearProblem::setLeftPrec with PL = diag(dinv) (n doubles). */
int mcls_gpu_set_left_prec_diag(mcls_gpu_context* ctx, const double* dinv,
                                int ptr_kind);
/* PL = diag(1/a_ii) built on the device (point Jacobi). */
int mcls_gpu_set_left_prec_point_jacobi(mcls_gpu_context* ctx);
int mcls_gpu_clear_left_prec(mcls_gpu_context* ctx);
/* Copy PL's diagonal out (n doubles); MCLS_ERR_STATE without a diagonal PL. */
int mcls_gpu_get_left_prec_diag(mcls_gpu_context* ctx, double* dinv, int ptr_kind);

/* General (sparse) preconditioners, n x n CSR with the operator's n: int64
 * row pointers, int32 columns strictly ascending in each row, FP64 values
 * (host or device arrays, `ptr_kind`; copied).
 *   LinearProblem::setLeftPrec / setRightPrec  src/MCLS_LinearProblem_impl.hpp:115-134
 * A sparse PL replaces a diagonal one and vice versa.  With either present
 * the domain is built from the reference's composite operator
 * (getCompositeOperator / getTransposeCompositeOperator, :139-221): the
 * products PL (A PR) (forward) or PR^T (A^T PL^T) (adjoint), each filtered
 * by "Composite Operator Threshold" (tpetra/src/MCLS_TpetraHelpers.hpp:196-263).
 * The residual is PL (b - A PR x) (updatePrecResidual, :411-442), the MC
 * source omega PL b, and the solvers return x = PR y
 * (src/MCLS_MonteCarloSolverManager_impl.hpp:220-229,
 * src/MCLS_MCSASolverManager_impl.hpp:410-420).  Not with batched systems
 * (MCLS_ERR_UNSUPPORTED at solve time). */
int mcls_gpu_set_left_prec_matrix(mcls_gpu_context* ctx, const int64_t* row_ptr,
                                  const int32_t* col, const double* val, int ptr_kind);
int mcls_gpu_set_right_prec_matrix(mcls_gpu_context* ctx, const int64_t* row_ptr,
                                   const int32_t* col, const double* val, int ptr_kind);
int mcls_gpu_clear_right_prec(mcls_gpu_context* ctx);

/* ---------------------------------------------------------- parameters -- */
/* Teuchos::ParameterList equivalents, by the reference's key names:
 *   f64 "Convergence Tolerance"        (1.0)  src/MCLS_MCSASolverManager_impl.hpp:294-297
 *   i64 "Maximum Iterations"           (1000) :324-328
 *   i64 "Iteration Print Frequency"    (10)   :329-333 (printing needs Verbosity > 0)
 *   i64 "Iteration Check Frequency"    (1)    :334-338
 *   i64 "Smoother Steps"               (1)    :339-343
 *   str "Fixed Point Type"             ("Richardson" only)  :97-101
 *   f64 "Richardson Relaxation"        (1.0)  src/MCLS_RichardsonIteration_impl.hpp:61-64
 *   f64 "Sample Ratio"                 (1.0)  src/MCLS_UniformAdjointSource_impl.hpp:82-86
 *   str "Source Sampling Type"         ("Random" | "Stratified") :90-101
 *   i64 "History Length"               (10)   src/MCLS_AlmostOptimalDomain_impl.hpp:88-91
 *   f64 "Neumann Relaxation"           (1.0)  src/MCLS_AlmostOptimalDomain_impl.hpp:143-148
 *   f64 "Composite Operator Threshold" (0.0)  src/MCLS_MonteCarloSolverManager_impl.hpp:325-329
 *   str "Transport Type", i64 "MC Check Frequency", i64 "MC Buffer Size":
 *        accepted; one GPU holds the whole domain, so there is no handoff.
 * Build extensions (keys present in MCLS inputs but not read by src/):
 *   f64 "Weight Cutoff"        (0 = off) relative: stop when |w| < c |w0|
 *   str "Estimator Type"       ("Collision" | "Expected Value")
 * Build-only keys:
 *   i64 "Number of Histories"  (0 = N * Sample Ratio)
 *   str "Walk Precision"       ("FP64" | "FP32")
 *   str "Transition Sampler"   ("CDF" | "Alias" | "Learned" | "Learned Alias" |
 *        "Alias Strided"): row CDF inversion (the reference's lower_bound),
 *        Vose alias tables (rows <= 64 entries), the importance-reshaped rows
 *        of the learned proposal by CDF inversion, the same proposal by
 *        alias records (FP32 walks, rows <= 64 entries), or alias tables
 *        padded to the longest row's length, one 16-byte record per slot
 *        holding the row's own W (FP32 walks, collision or forward
 *        estimator, rows <= 64 entries: half the bytes per row of "Alias"
 *        records on 27-point rows)
 *   i64 "Random Seed"          (0)
 *   i64 "Use Graph"            (0) capture each MCSA iteration in a hipGraph
 *   i64 "Verbosity"            (0)
 *   i64 "Compute Convergence Criteria" (0) print mcls_gpu_convergence_criteria
 *        after each domain build (src/MCLS_AlmostOptimalDomain_impl.hpp:95-112)
 *   i64 "Walk Block Size"      (256), i64 "Walk Blocks Per CU" (0 = auto),
 *   i64 "Walk Batch"           (64) history ids per work-queue grab
 *   i64 "Tally Tile"           (0) LDS tally tile entries per wave (0 = off,
 *        else a power of two in [64, 2048]; adjoint collision and expected-
 *        value estimators, each expected-value scatter entry through it): tallies
 *        merge in LDS and are flushed into an FP64 tally, also for FP32 walks
 *        (these four are what mcls_gpu_autotune sets)
 *   i64 "Walks Per Lane"       (1) 2: each lane carries two histories and
 *        issues both of their record loads before waiting on either
 *        ("Alias Strided" FP32 walks, collision or forward estimator; other
 *        walks and path recording use one per lane)
 *   i64 "Strided Record Bytes" (16) 8: "Alias Strided" walks (one history
 *        per lane) read the same tables in 8-byte records {column, alias
 *        slot and sign; threshold}, the row's W from the FP32 row array and,
 *        when the draw takes the alias, the alias slot's own record for its
 *        column: the same paths in half the table footprint (rows of at most
 *        32 entries, fewer than 2^26 - 1 unknowns, else MCLS_ERR_UNSUPPORTED
 *        at the next walk); the 16-byte tables are then built only for
 *        "Walks Per Lane" 2, which has no 8-byte form
 *   str "Tally Tile Precision" ("FP64" | "Walk" | "Compensated"): the
 *        tile's partial sums in FP64 (12 bytes of LDS per entry), in the
 *        walk's precision (FP32 walks: 8 bytes, more resident waves per
 *        CU), or in FP32 with a compensation term per entry that collects
 *        the exact rounding error of every addition (FP32 walks, 12
 *        bytes); each is flushed into the FP64 tally, which every walk
 *        precision uses (FP32 walks included, tiled or not)
 *   str "Birth Order"          ("Id" | "State"): with the "Random" source,
 *        walk the histories in the order of their birth states instead of
 *        their ids - each history's birth is drawn ahead of the walk from
 *        its own stream and the histories are radix-sorted by it, so
 *        every history keeps its id, stream and path and only the walking
 *        order (the locality of the walks in flight) and the tally
 *        summation order change; not with "Use Graph" (refused), and the
 *        allocation-free solve walks in id order
 *   str "Residual Format"      ("Sliced ELL" | "CSR"): the residual
 *        PL (b - A PR y) reads a sliced-ELL copy of A (slices of 64 rows
 *        padded to their longest row, built with the operator when padding
 *        adds at most a quarter to the entries and it fits in memory; each
 *        row summed left to right, as the CSR SpMV) or A's CSR arrays with
 *        a few lanes per row
 * Unknown keys and wrong types return MCLS_ERR_INVALID_ARG. */
int mcls_gpu_set_param_i64(mcls_gpu_context* ctx, const char* key, int64_t v);
int mcls_gpu_set_param_f64(mcls_gpu_context* ctx, const char* key, double v);
int mcls_gpu_set_param_str(mcls_gpu_context* ctx, const char* key, const char* v);

/* --------------------------------------------------------------- domain -- */
/* Build the adjoint domain on the device: rows of H^T with
 * H = I - omega PL A, |h| < DBL_EPSILON dropped, per-row FP64 and FP32 CDFs,
 * row weights W_i = sum |h_i.|.  Rebuilt automatically by the solves when
 * the matrix, preconditioner, "Neumann Relaxation" or "Composite Operator
 * Threshold" changed (MonteCarloSolverManager::setProblem, :161-178). */
int mcls_gpu_build_domain(mcls_gpu_context* ctx);
int mcls_gpu_domain_nnz(mcls_gpu_context* ctx, int64_t* nnz);
/* Copy the domain out: rp (n+1), col (nnz), h (nnz), cdf (nnz), W (n).
 * Any pointer may be NULL.  Host pointers. */
int mcls_gpu_get_domain(mcls_gpu_context* ctx, int64_t* rp, int32_t* col,
                        double* h, double* cdf, double* W);

/* -------------------------------------------------------------- solves -- */
/* MonteCarloSolverManager<..,AdjointTag>::solve (not internal): the source
 * is omega * PL b; x receives the normalised adjoint tally.  Always
 * "converged" (src/MCLS_MonteCarloSolverManager.hpp:170-171). */
int mcls_gpu_mc_solve(mcls_gpu_context* ctx, const double* b, double* x,
                      int ptr_kind);
/* MCSASolverManager<..,AdjointTag>::solve: x is read as the initial guess
 * and overwritten with the solution.  Returns MCLS_OK whether or not the
 * tolerance was met; read converged from mcls_gpu_get_stats. */
int mcls_gpu_mcsa_solve(mcls_gpu_context* ctx, const double* b, double* x,
                        int ptr_kind);
/* mcls_gpu_mc_solve that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  b and x are device pointers (n doubles, caller's numbering).
 * Every per-call buffer (source tables, tally, work-queue and step
 * counters, scan partials; with a state order b and x in the internal
 * numbering) lives in a static device array of 65 MiB that each GPU the
 * library runs on holds until the process ends, so n is at most
 * MCLS_NOALLOC_MAX_N (MCLS_ERR_INVALID_ARG beyond).  The domain and the
 * Transition Sampler's tables must exist for the current operator and
 * parameters (mcls_gpu_build_domain builds both), else MCLS_ERR_STATE;
 * batched systems, sparse preconditioners and an attached importance
 * network are refused (MCLS_ERR_UNSUPPORTED).  The result equals
 * mcls_gpu_mc_solve's for the same solve number.  All calls on one GPU
 * share its array: calls that overlap on one GPU (two contexts' streams,
 * two threads) race; calls that follow each other do not, each clears
 * what it reads as zero. */
#define MCLS_NOALLOC_MAX_N 1048576
int mcls_gpu_mc_solve_noalloc(mcls_gpu_context* ctx, const double* b, double* x);
/* LinearProblem::updatePrecResidual: r = PL (b - A x) and ||r||_2. */
int mcls_gpu_prec_residual(mcls_gpu_context* ctx, const double* b,
                           const double* x, double* r, double* norm2,
                           int ptr_kind);
int mcls_gpu_get_stats(mcls_gpu_context* ctx, mcls_gpu_stats* out);
/* The relative residuals ||rp||_2 / ||PL b||_2 of the last
 * mcls_gpu_mcsa_solve: entry 0 before the first iteration, entry k after
 * iteration k -- the values MCSASolverManager::solve prints every
 * "Iteration Print Frequency" iterations
 * (src/MCLS_MCSASolverManager_impl.hpp:392-406).  Copies at most `cap`
 * entries; *count receives how many there are. */
int mcls_gpu_get_residual_history(mcls_gpu_context* ctx, double* hist, int64_t cap, int64_t* count);

/* ------------------------------------------------- convergence criteria -- */
/* AlmostOptimalDomain::computeConvergenceCriteria
 * (src/MCLS_AlmostOptimalDomain_impl.hpp:439-576) for the domain operator
 * (rows of H^T, built first if needed): the spectral radii of H, |H| and
 * H* (H*_ij = |h_ij| W_i; rho(H*) < 1 is the condition for a finite
 * estimator variance), ||H||_inf, the smallest row sum of |H|, and the
 * balance ratios max_k |h_ik| / max_k |h_ki| over non-empty rows.  The
 * reference uses Anasazi's Generalized Davidson; here power iteration on
 * the device, the growth factor averaged over windows of 32 steps, stopped
 * when two windows agree to `tol` (relative) or after `max_iters` (>= 64)
 * steps.  With the i64 parameter "Compute Convergence Criteria" = 1 they
 * are also printed after every domain build, as the reference does. */
typedef struct mcls_gpu_criteria {
  double rho_h, rho_habs, rho_hstar;   /* rho(H), rho(|H|) ("H+"), rho(H*)   */
  double inf_norm;                     /* ||H||_inf = max_i W_i              */
  double min_row_sum;                  /* min_i W_i                          */
  double balance_min, balance_max, balance_ave;
  int64_t iterations[3];               /* power steps per radius             */
  int64_t converged[3];                /* 1 when two windows agreed to tol   */
} mcls_gpu_criteria;
int mcls_gpu_convergence_criteria(mcls_gpu_context* ctx, int64_t max_iters, double tol, mcls_gpu_criteria* out);

/* ----------------------------------------------------------- autotuning -- */
/* Walk launch shape picked by timing (no counterpart in the reference, whose
 * transport loop runs on CPU ranks).  Pilot MC solves of the source PL b
 * walk `pilot_histories` consecutive histories from the middle of the full
 * source (0: 1/8 of the configured count, at least 2^20), so that every
 * pilot sees the solve's births per state; each candidate keeps the faster
 * of two pilots.  They time the block size (64,
 * 128, 256) at 8, 16 and 32 resident waves per CU, then the history ids a
 * wave takes per queue grab (64 .. 512) at the fastest shape, then
 * (adjoint collision or expected-value estimator) the LDS tally tile: off, 128, 512 or 1024 entries
 * per wave (those that fit 64 KiB of LDS); last, the best shape of each
 * block size walks the whole source once, and the fastest of those sets
 * the block size, after which the chosen batch's neighbours (half and
 * double, within 64 .. 512) walk the whole source at that shape and the
 * fastest sets the batch.  `ms_best` is the window pilot time of the configuration
 * returned and `ms_confirmed` its walk of the whole source.  Shapes whose
 * tile does not fit the LDS are skipped; an entry configuration that does
 * not fit is not timed (`ms_default` = -1) and the search starts untiled.
 * The winner is stored in "Walk
 * Block Size", "Walk Blocks Per CU", "Walk Batch" and "Tally Tile" and
 * reported in `out`.  Estimates do not
 * depend on the launch shape (each history's random stream depends only on
 * its id); the order of the floating-point tally additions does.  The solve
 * counter is not advanced. */
typedef struct mcls_gpu_tuning {
  int64_t block;          /* Walk Block Size                                 */
  int64_t blocks_per_cu;  /* Walk Blocks Per CU                              */
  int64_t batch;          /* Walk Batch                                      */
  int64_t tally_tile;     /* Tally Tile                                      */
  double ms_default;      /* pilot walk time with the configuration on entry */
  double ms_best;         /* pilot walk time with the chosen configuration   */
  int64_t candidates;     /* configurations timed                            */
  double ms_confirmed;    /* full-source walk time of the chosen configuration */
} mcls_gpu_tuning;
int mcls_gpu_autotune(mcls_gpu_context* ctx, const double* b, int ptr_kind, int64_t pilot_histories,
                      mcls_gpu_tuning* out);

/* ------------------------------------------------------ learned proposal -- */
/* Importance g (n FP32 values > 0) for "Transition Sampler" = "Learned"
 * (no counterpart in the reference, whose transitions are always
 * P_ij = |h_ij| / W_i, src/MCLS_AlmostOptimalDomain_impl.hpp:342-363):
 * the walk then samples P_ij = |h_ij| g_j / S_i, S_i = sum_k |h_ik| g_k, and
 * multiplies the weight by sign(h_ij) S_i / g_j, which keeps every
 * estimator unbiased for any g.  g = 1 reproduces the plain transition
 * probabilities.  The tables are rebuilt on the device. */
int mcls_gpu_set_importance(mcls_gpu_context* ctx, const float* g, int ptr_kind);
/* Importance g_j = |a_jj| computed on the device.  For symmetric A under
 * point Jacobi (H = D^-1 T) it makes P_ij proportional to |t_ij|, the
 * symmetric random walk on the couplings, and the weight after k steps
 * telescopes to S_0 / |a_kk| times the row ratios
 * sum_j |t_ij| / |a_ii| of the states passed, each at most one for a
 * diagonally dominant row. */
int mcls_gpu_set_importance_diagonal(mcls_gpu_context* ctx);
/* Current importance g (n floats): the one set, or the one the attached
 * network computed for the last MC source. */
int mcls_gpu_get_importance(mcls_gpu_context* ctx, float* g, int ptr_kind);
/* Attach an importance network (include/mcls_mlp.h; NULL detaches).  With
 * the "Learned" sampler every MC solve then computes g from the network on
 * its own source: features of (A, source) -> network -> g = exp(clip(o,
 * -4, 4)) -> proposal tables, all on the device (and inside the captured
 * MCSA iteration).  The context does not own the network. */
struct mcls_mlp;
int mcls_gpu_set_importance_model(mcls_gpu_context* ctx, struct mcls_mlp* m);
/* Per-state switch (n int8, 0 = use g = 1 there, the plain proposal); NULL
 * switches every state back on.  The fallback where g raises variance. */
int mcls_gpu_set_importance_mask(mcls_gpu_context* ctx, const signed char* mask, int ptr_kind);
/* The network's inputs for source v (n x 8 floats):
 *   f0..f3  sums of -a_jk / a_jj over off-diagonal k < j-1, k = j-1,
 *           k = j+1, k > j+1 (the S, W, E, N couplings of a 5-point row)
 *   f4      sum_{k != j} |a_jk| / |a_jj|
 *   f5      log(|a_jj| / mean |a_ii|)
 *   f6      log(|v_j| / mean |v| + 1e-3)
 *   f7      0 */
int mcls_gpu_importance_features(mcls_gpu_context* ctx, const double* v, float* X, int ptr_kind);
/* Training targets from the solver's own adjoint tally: one MC solve of
 * source v (number `solve_index`, plain transitions) giving x, then
 * t_j = log(clip(|a_jj x_j| / mean_i |a_ii x_i|, e^-4, e^4)) (n floats).
 * For symmetric A with H = I - D^-1 A >= 0 and v >= 0, phi = D x solves
 * phi = D v + H^T phi: it is the adjoint walk's value function for the
 * functional <D v, .>, the importance under which the weight factors
 * S_i / g_j = 1 - (D v)_i / phi_i stay below one along a walk. */
int mcls_gpu_importance_targets(mcls_gpu_context* ctx, const double* v, uint64_t solve_index, float* T,
                                int ptr_kind);

/* ------------------------------------------------------- batched systems -- */
/* Many independent systems in one launch (no counterpart in the reference,
 * which solves one system per manager; its multi-RHS Thyra adapter loops
 * over columns, thyra/src/MCLS_SolverManagerAdapter_impl.hpp:83-154).  The
 * matrix set with mcls_gpu_set_matrix_csr is block diagonal with system s
 * on rows [offsets[s], offsets[s+1]) (nsys+1 host int64 offsets, 0 and n at
 * the ends); nsys = 0 returns to one system.  mcls_gpu_mcsa_solve then runs
 * every system's MCSA iteration in the same kernels: histories are
 * allocated per system in proportion to its unknown count, each system is
 * tested against its own ||PL b_s||_2, a converged system gets no further
 * MC correction, and the solve ends when all have converged (stats:
 * achieved_tol = worst system). */
int mcls_gpu_set_systems(mcls_gpu_context* ctx, int64_t nsys, const int64_t* offsets);
/* Relative residual of each system after the last batched solve (nsys). */
int mcls_gpu_get_system_residuals(mcls_gpu_context* ctx, double* rel);

/* ------------------------------------------------------------- parity -- */
/* Run the walk kernel for histories [first, first+count) of MC solve number
 * `solve_index` with source `rhs` (already preconditioned and scaled) and
 * record each history's state indices: paths is count x (max_len+1),
 * padded with -1; weights receives each final weight.  Host pointers. */
int mcls_gpu_record_paths(mcls_gpu_context* ctx, const double* rhs,
                          uint64_t solve_index, int64_t first, int64_t count,
                          int32_t max_len, int32_t* paths, double* weights);
/* One internal MC solve (no preconditioning, no omega) of number
 * `solve_index`: x = normalised tally for source rhs.  Host pointers. */
int mcls_gpu_mc_solve_raw(mcls_gpu_context* ctx, const double* rhs,
                          uint64_t solve_index, double* x);
/* The same solve restricted to the history ids [first, first+count) of its
 * full source: x = that window's tally normalised by the full history
 * count, without the expected-value estimator's birth term rhs, so the
 * windows of a partition sum to mcls_gpu_mc_solve_raw's x (minus rhs for
 * that estimator).  Host pointers.  (MCSolver::solve over part of the
 * source, src/MCLS_MCSolver_impl.hpp:83-112, for parity checks of large
 * solves against a CPU restatement.) */
int mcls_gpu_mc_solve_window(mcls_gpu_context* ctx, const double* rhs, uint64_t solve_index,
                             int64_t first, int64_t count, double* x);

#ifdef __cplusplus
}
#endif
#endif
