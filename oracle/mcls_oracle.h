/*
 * mcls_oracle.h -- CPU restatement of the MCLS adjoint Neumann-Ulam / MCSA
 * path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * path is compared against, and the CPU baseline that bench.py times.  No
 * product code links it.
 *
 * Two families of functions live here:
 *
 *  1. "build semantics" (orc_*): the exact arithmetic the GPU library
 *     implements (per-history Xorshift streams, the integer-quantised source
 *     CDF, FP64/FP32 domain tables, collision and expected-value estimators,
 *     relative weight cutoff).  These are bit-exact twins of the HIP kernels
 *     for everything except floating-point tally accumulation order.
 *
 *  2. "reference semantics" (orc_ref_*): a faithful restatement of the
 *     MCLS algorithm itself (one Xorshift stream per set, FP64 |b| CDF with
 *     lower_bound, LIFO source stack, weight sign(b)*||b||_1, multi-set sum),
 *     used as the CPU baseline.
 *
 * Every function cites the reference file:line it follows.
 */
#ifndef MCLS_ORACLE_H
#define MCLS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ RNG -- */
uint64_t orc_splitmix64(uint64_t z);
uint64_t orc_xorshift64(uint64_t* state);
uint64_t orc_history_state(uint64_t seed, uint64_t solve_index, uint64_t history_id);
double orc_u01_f64(uint64_t r);
float orc_u01_f32(uint64_t r);

/* ------------------------------------------------------------- sampling -- */
int orc_lower_bound_f64(const double* cdf, int n, double u);
int orc_lower_bound_f32(const float* cdf, int n, float u);
int64_t orc_upper_bound_u64(const uint64_t* cdf, int64_t n, uint64_t t);

/* --------------------------------------------------------------- domain -- */
/* Build the adjoint domain rows (rows of H^T, H = I - omega*PL*A, PL =
 * diag(dinv) or identity when dinv == NULL).  Output arrays must hold
 * a_nnz entries (d_rp: n+1).  When a preconditioner is present the
 * composite row is filtered first: entries with |m| <= max(row)*threshold
 * are dropped ("Composite Operator Threshold").  Returns the domain nnz, or a
 * negative error:
 * -1 column out of range, -2 duplicate column in a row. */
int64_t orc_build_domain(int n, const int64_t* a_rp, const int* a_col,
                         const double* a_val, const double* dinv, double omega,
                         double threshold, int64_t* d_rp, int* d_col,
                         double* d_h, double* d_cdf, double* d_W);
/* The forward method's domain: rows of H = I - omega PL A (not transposed),
 * each ordered by column, with the same filter and drop
 * (src/MCLS_MonteCarloSolverManager_impl.hpp:243-248,
 * src/MCLS_LinearProblem_impl.hpp:141-172). */
int64_t orc_build_domain_forward(int n, const int64_t* a_rp, const int* a_col,
                                 const double* a_val, const double* dinv,
                                 double omega, double threshold, int64_t* d_rp,
                                 int* d_col, double* d_h, double* d_cdf,
                                 double* d_W);

/* Point-Jacobi inverse diagonal: dinv_i = 1/a_ii.  Returns 0, or -(i+1)
 * for the first row with a zero / missing diagonal. */
int orc_jacobi_dinv(int n, const int64_t* a_rp, const int* a_col,
                    const double* a_val, double* dinv);

/* Integer-quantised |b| source table.  Returns Q (total) and writes
 * q (n entries) and the inclusive prefix cdf (n entries). */
uint64_t orc_source_quantize(int n, const double* b, uint64_t* q, uint64_t* cdf);

/* Per-history birth: state and weight of history `hid`.  Returns state or
 * -1 if the source is empty.  Advances *rng by one draw. */
int orc_source_birth(int n, const double* b, const uint64_t* q,
                     const uint64_t* cdf, uint64_t Q, uint64_t* rng,
                     double* w0);

/* Per-row alias tables (Vose) over p_k = |h_k| / W_i: thr[nnz] (slot keeps
 * its own entry when u < thr) and alias[nnz] (row-local index taken
 * otherwise).  Rows longer than 64 entries return -(row+1). */
int64_t orc_build_alias(int n, const int64_t* d_rp, const double* d_h,
                        const double* d_W, double* thr, int* alias);

/* Strided alias tables (the build's "Alias Strided" sampler): stride = the
 * longest row's length (>= 1; 0 when a row has more than 64 entries).  Each row is padded with zero-probability slots to
 * `stride` entries and Vose's method runs over all of them, as
 * orc_build_alias does; thr[n*stride], alias[n*stride] (row-local index).
 * A padding slot keeps threshold 0; one that rounding leaves over at the end
 * aliases the row's first entry. */
int orc_alias_stride(int n, const int64_t* d_rp);
int64_t orc_build_alias_strided(int n, const int64_t* d_rp, const double* d_h,
                                const double* d_W, int stride, double* thr, int* alias);

/* Importance-reshaped transition tables (build extension, the learned
 * proposal): for domain row i with entries (j, h_ij) and importance g_j > 0,
 * P_ij = |h_ij| g_j / S_i with S_i = sum_k |h_ik| g_k, sampled by CDF
 * inversion of lcdf, and the weight factor fac_ij = S_i / g_j (times the
 * sign of h_ij), so E[w'] = sum_j P_ij fac_ij sign(h_ij) = sum_j h_ij:
 * unbiased for any g.  g is FP32 (the network's output). */
int orc_build_proposal(int n, const int64_t* d_rp, const int* d_col,
                       const double* d_h, const float* g, double* lcdf,
                       double* lfac);
/* The same proposal as per-row alias tables (Vose over |h_ij| g_j / S_i, S_i
 * summed in row order) with the factors lfac = S_i / g_j; rows of at most 64
 * entries.  Returns 0, or -(row+1) for a longer row or a bad g. */
int64_t orc_build_proposal_alias(int n, const int64_t* d_rp, const int* d_col,
                                 const double* d_h, const float* g, double* thr,
                                 int* alias, double* lfac);

/* ------------------------------------------------------------- MC solve -- */
/* ORC_EST_FORWARD: the forward method (ForwardTag) on the forward domain:
 * each state starts orc_forward_per_state(n, num_histories) histories of
 * weight 1, a history adds w b[state] at every collision to its own sum,
 * and x[start] is the average of its histories' sums; the source type is
 * ignored. */
enum { ORC_EST_COLLISION = 0, ORC_EST_EXPECTED_VALUE = 1, ORC_EST_FORWARD = 2 };
enum { ORC_SRC_RANDOM = 0, ORC_SRC_STRATIFIED = 1 };
/* ORC_SAMPLER_LEARNED_ALIAS: the learned proposal drawn with alias tables
 * (orc_build_proposal_alias), weight factor S_i / g_j as for LEARNED.
 * ORC_SAMPLER_ALIAS_STRIDED: orc_build_alias_strided tables, slot
 * floor(u stride) of the row (same draw as ORC_SAMPLER_ALIAS with len =
 * stride). */
enum { ORC_SAMPLER_CDF = 0, ORC_SAMPLER_ALIAS = 1, ORC_SAMPLER_LEARNED = 2, ORC_SAMPLER_LEARNED_ALIAS = 3,
       ORC_SAMPLER_ALIAS_STRIDED = 4 };

typedef struct orc_mc_params {
  int64_t num_histories;   /* histories per MC solve                       */
  int64_t history_length;  /* max transitions per history ("History Length") */
  double weight_cutoff;    /* relative cutoff, 0 = off                      */
  int64_t precision;       /* 64 or 32 (walk weights / CDF / W)            */
  int64_t estimator;       /* ORC_EST_*                                    */
  int64_t source;          /* ORC_SRC_*                                    */
  uint64_t seed;
  int64_t num_threads;     /* worker threads (tallies reduced in order)    */
  int64_t sampler;         /* ORC_SAMPLER_* (transition sampling)          */
  const float* importance; /* g (n) for ORC_SAMPLER_LEARNED, else NULL      */
} orc_mc_params;

/* One adjoint MC solve over the domain: x = (1/nh) * sum of tallies (plus
 * the deterministic birth term b for the expected-value estimator).
 * Returns 0 on success; *steps_out = total transitions. */
int orc_mc_solve(int n, const int64_t* d_rp, const int* d_col,
                 const double* d_h, const double* d_cdf, const double* d_W,
                 const double* b, const orc_mc_params* p,
                 uint64_t solve_index, double* x, int64_t* steps_out);

/* The same solve restricted to the histories [first, first+count) of the
 * full source (ids beyond it dropped), still normalised by the full count,
 * without the expected-value estimator's birth term, so that the windows
 * of a partition sum to orc_mc_solve's x minus that term.  Births are drawn
 * by |mag| when mag is given (weights b_k Q / q_k stay unbiased for b), as
 * a batch of systems reallocates them. */
int orc_mc_solve_window(int n, const int64_t* d_rp, const int* d_col,
                        const double* d_h, const double* d_cdf, const double* d_W,
                        const double* b, const double* mag, const orc_mc_params* p,
                        uint64_t solve_index, int64_t first, int64_t count,
                        double* x, int64_t* steps_out);

int64_t orc_forward_per_state(int n, int64_t num_histories);

/* Record the state-index path of histories [first, first+count): paths is
 * count x (max_len+1) ints, padded with -1; weights the final weight. */
int orc_record_paths(int n, const int64_t* d_rp, const int* d_col,
                     const double* d_h, const double* d_cdf, const double* d_W,
                     const double* b, const orc_mc_params* p,
                     uint64_t solve_index, int64_t first, int64_t count,
                     int max_len, int* paths, double* weights);

/* ---------------------------------------------------- composite operator -- */
/* Transpose of an n x n CSR matrix (MT::copyTranspose,
 * tpetra/src/MCLS_TpetraCrsMatrixAdapter.hpp:416-422): each transposed row
 * lists its entries in ascending source row.  t_rp holds n+1 entries, t_col /
 * t_val rp[n]. */
void orc_transpose(int n, const int64_t* rp, const int* col, const double* val,
                   int64_t* t_rp, int* t_col, double* t_val);

/* C = filter(X Y, threshold) (MT::multiply,
 * tpetra/src/MCLS_TpetraCrsMatrixAdapter.hpp:388-399, and
 * TpetraMatrixHelpers::filter, tpetra/src/MCLS_TpetraHelpers.hpp:196-263).
 * Row-by-row Gustavson product: for row i, the entries k of X's row in
 * ascending column order, then the entries j of Y's row k; C_ij starts at the
 * first product and adds the later ones in that order.  Rows of C are in
 * ascending column order.  The filter drops, per row, every entry with
 * |c| <= (signed row max) * threshold, and every exact zero (the reference
 * removes zero values with std::remove).  X and Y rows must be ascending.
 * Returns nnz(C), or -1 when it would exceed cap. */
int64_t orc_spgemm(int n, const int64_t* x_rp, const int* x_col, const double* x_val,
                   const int64_t* y_rp, const int* y_col, const double* y_val,
                   double threshold, int64_t cap, int64_t* c_rp, int* c_col,
                   double* c_val);

/* --------------------------------------------------------------- linalg -- */
void orc_spmv(int n, const int64_t* rp, const int* col, const double* val,
              const double* x, double* y);
/* rp = PL (b - A x) (PL = diag(dinv) or identity) */
void orc_prec_residual(int n, const int64_t* rp, const int* col,
                       const double* val, const double* dinv, const double* b,
                       const double* x, double* r);
double orc_norm2(int n, const double* v);
double orc_norm1(int n, const double* v);
double orc_norminf(int n, const double* v);

/* --------------------------------------------------------------- solvers -- */
typedef struct orc_solver_params {
  double tol;               /* "Convergence Tolerance"       */
  int64_t max_iters;        /* "Maximum Iterations"          */
  int64_t check_freq;       /* "Iteration Check Frequency"   */
  int64_t smoother_steps;   /* "Smoother Steps"              */
  double richardson_omega;  /* "Richardson Relaxation"       */
  double neumann_omega;     /* "Neumann Relaxation"          */
  double threshold;         /* "Composite Operator Threshold" */
  int64_t num_sets;         /* multi-set replication (>= 1)  */
  int64_t reference_mode;   /* 1 = MCLS reference MC semantics, 0 = build */
  orc_mc_params mc;
  int64_t ref_quirk;        /* reference mode: 1 = the signed b[0] CDF seed */
} orc_solver_params;

typedef struct orc_solver_result {
  int64_t iters;
  int64_t converged;
  double achieved_tol;
  int64_t histories;
  int64_t steps;
  double seconds_mc;
  double seconds_total;
} orc_solver_result;

/* MCSASolverManager<...,AdjointTag>::solve.  dinv may be NULL (no left
 * preconditioner).  res_hist (nullable) receives max_iters+1 relative
 * residuals. */
int orc_mcsa_solve(int n, const int64_t* rp, const int* col, const double* val,
                   const double* dinv, const double* b, double* x,
                   const orc_solver_params* sp, uint64_t solve_index0,
                   orc_solver_result* res, double* res_hist);

/* FixedPointSolverManager (Richardson) with the infinity-norm test. */
int orc_fixed_point_solve(int n, const int64_t* rp, const int* col,
                          const double* val, const double* dinv,
                          const double* b, double* x,
                          const orc_solver_params* sp, orc_solver_result* res);

/* ------------------------------------------------ reference semantics -- */
/* UniformAdjointSource::buildRandomSource / buildStratifiedSource with one
 * stream: fills counts[n] and returns the number of histories to emit (the
 * normalisation constant).  ref_quirk = 1 reproduces the signed b[0] CDF
 * seed of the reference; 0 uses |b[0]|. */
int64_t orc_ref_build_source(int n, const double* b, double sample_ratio_nh,
                             int stratified, int ref_quirk, uint64_t* rng,
                             int64_t* counts);

/* One reference-semantics MC solve with a single stream *rng: source
 * build, LIFO emission (highest state first), collision tally, normalise.
 * Returns the normalisation constant (histories emitted). */
int64_t orc_ref_mc_solve(int n, const int64_t* d_rp, const int* d_col,
                         const double* d_h, const double* d_cdf,
                         const double* d_W, const double* b,
                         const orc_mc_params* p, int stratified, int ref_quirk,
                         uint64_t* rng, double* x, int64_t* steps_out);

/* DomainTransporter::transport of one history from (state, weight):
 * returns the event (1 = CUTOFF, 2 = weight cutoff, 3 = absorbed on an
 * empty row) and updates state/weight/steps; tallies into x. */
int orc_ref_transport(const int64_t* d_rp, const int* d_col, const double* d_h,
                      const double* d_cdf, const double* d_W,
                      int64_t history_length, double weight_cutoff,
                      uint64_t* rng, int* state, double* weight,
                      int64_t* steps, double* x);

#ifdef __cplusplus
}
#endif
#endif
