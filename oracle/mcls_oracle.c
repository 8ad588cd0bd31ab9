/*
 * mcls_oracle.c -- CPU restatement of the MCLS adjoint MC / MCSA path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); see
 * mcls_oracle.h.  Compiled with -ffp-contract=off so that every FP64/FP32
 * operation rounds exactly as written, matching the HIP kernels, which are
 * built the same way.
 *
 * Reference files followed (paths relative to the MCLS source root):
 *   src/MCLS_Xorshift.hpp:103-109                 Xorshift (13,7,17) recurrence
 *   src/MCLS_SamplingTools.hpp:67-80              lower_bound CDF inversion
 *   src/MCLS_AlmostOptimalDomain_impl.hpp:258-366 H row build, eps drop, CDF, W
 *   src/MCLS_AlmostOptimalDomain.hpp:255-284      processTransition
 *   src/MCLS_AlmostOptimalDomain.hpp:135-136      terminateHistory (steps >= L)
 *   src/MCLS_DomainTransporter_impl.hpp:85-115    tally -> transition -> terminate
 *   src/MCLS_AdjointTally.hpp:115-123             collision tally x[i] += w
 *   src/MCLS_AdjointTally_impl.hpp:69-83          normalize / zeroOut
 *   src/MCLS_UniformAdjointSource_impl.hpp:59-263 source build and emission
 *   src/MCLS_MCSolver_impl.hpp:83-112             zeroOut -> transport -> normalize
 *   src/MCLS_MonteCarloSolverManager_impl.hpp:347-380  RHS copy * omega
 *   src/MCLS_MCSASolverManager_impl.hpp:279-432   MCSA outer loop
 *   src/MCLS_MultiSetLinearProblem_impl.hpp:103-120   sum over sets / numSets
 *   src/MCLS_RichardsonIteration_impl.hpp:135-147 x += omega*rp; update rp
 *   src/MCLS_LinearProblem_impl.hpp:411-442       rp = PL (b - A PR x)
 *   src/MCLS_FixedPointSolverManager_impl.hpp:208-320  Richardson solve, inf-norm
 *   tpetra/src/MCLS_TpetraPointJacobiPreconditioner_impl.hpp:105-139  1/a_ii
 */
#define _GNU_SOURCE
#include "mcls_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ================================================================== RNG == */

/* SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014): the build's stream-seeding function. */
uint64_t orc_splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* src/MCLS_Xorshift.hpp:103-109: x ^= x<<13; x ^= x>>7; x ^= x<<17. */
uint64_t orc_xorshift64(uint64_t* state) {
  uint64_t x = *state;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  *state = x;
  return x;
}

/* Per-history stream state.  The reference seeds one stream per rank from
 * std::random_device (src/MCLS_PRNG_impl.hpp:57-75), which is not
 * reproducible; the build gives every history its own Xorshift stream keyed
 * by (seed, solve index, history id). */
uint64_t orc_history_state(uint64_t seed, uint64_t solve_index,
                           uint64_t history_id) {
  uint64_t key = orc_splitmix64(seed ^ orc_splitmix64(solve_index));
  uint64_t s = orc_splitmix64(key ^ orc_splitmix64(history_id));
  return s ? s : 0x9E3779B97F4A7C15ULL;
}

/* 53 high bits -> [0,1) exactly. */
double orc_u01_f64(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }
/* 24 high bits -> [0,1) exactly. */
float orc_u01_f32(uint64_t r) { return (float)(r >> 40) * 0x1.0p-24f; }

/* ============================================================= sampling == */

/* src/MCLS_SamplingTools.hpp:67-80: std::lower_bound, first k with
 * cdf[k] >= u. */
int orc_lower_bound_f64(const double* cdf, int n, double u) {
  int lo = 0, len = n;
  while (len > 0) {
    int half = len >> 1;
    if (cdf[lo + half] < u) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

int orc_lower_bound_f32(const float* cdf, int n, float u) {
  int lo = 0, len = n;
  while (len > 0) {
    int half = len >> 1;
    if (cdf[lo + half] < u) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

/* First k with cdf[k] > t. */
int64_t orc_upper_bound_u64(const uint64_t* cdf, int64_t n, uint64_t t) {
  int64_t lo = 0, len = n;
  while (len > 0) {
    int64_t half = len >> 1;
    if (cdf[lo + half] <= t) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

/* =============================================================== domain == */

int orc_jacobi_dinv(int n, const int64_t* a_rp, const int* a_col,
                    const double* a_val, double* dinv) {
  /* tpetra/src/MCLS_TpetraPointJacobiPreconditioner_impl.hpp:105-139:
   * getLocalDiagCopy then reciprocal. */
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int64_t k = a_rp[i]; k < a_rp[i + 1]; ++k)
      if (a_col[k] == i) d = a_val[k];
    if (d == 0.0) return -(i + 1);
    dinv[i] = 1.0 / d;
  }
  return 0;
}

/* Row i of the domain holds column i of the composite PL*A (the transposed
 * composite operator, src/MCLS_LinearProblem_impl.hpp:179-221 with
 * tpetra/src/MCLS_TpetraCrsMatrixAdapter.hpp:389-422), entries in ascending
 * source-row order (the column order of the Tpetra transpose).  Then
 * src/MCLS_AlmostOptimalDomain_impl.hpp:296-364:
 *   h_ij = delta_ij - omega*m_ij, drop |h| < DBL_EPSILON,
 *   W_i = sum_k |h_ik| (sequential), cdf_k = (partial sum)_k / W_i. */
/* Domain rows of H^T (transpose = 1: the adjoint method,
 * getTransposeCompositeOperator) or of H (transpose = 0: the forward method,
 * getCompositeOperator, src/MCLS_LinearProblem_impl.hpp:141-172). */
static int64_t build_domain_impl(int n, const int64_t* a_rp, const int* a_col,
                                 const double* a_val, const double* dinv,
                                 double omega, double threshold, int transpose,
                                 int64_t* d_rp, int* d_col, double* d_h,
                                 double* d_cdf, double* d_W) {
  int64_t a_nnz = a_rp[n];
  int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t* pos = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
  int* tsrc = (int*)malloc((size_t)(a_nnz > 0 ? a_nnz : 1) * sizeof(int));
  double* tval = (double*)malloc((size_t)(a_nnz > 0 ? a_nnz : 1) * sizeof(double));
  int64_t ret = 0;
  for (int64_t k = 0; k < a_nnz; ++k) {
    int c = a_col[k];
    if (c < 0 || c >= n) { ret = -1; goto done; }
  }
  if (transpose) {
    /* transpose by counting sort over columns; rows of A are visited in
     * ascending order, so each transposed row is ascending in source row. */
    for (int64_t k = 0; k < a_nnz; ++k) cnt[a_col[k] + 1]++;
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    memcpy(pos, cnt, ((size_t)n + 1) * sizeof(int64_t));
    for (int r = 0; r < n; ++r) {
      for (int64_t k = a_rp[r]; k < a_rp[r + 1]; ++k) {
        int c = a_col[k];
        int64_t p = pos[c]++;
        tsrc[p] = r;
        /* composite entry (PL A)_{r c} = dinv_r * a_rc */
        tval[p] = dinv ? a_val[k] * dinv[r] : a_val[k];
      }
    }
  } else {
    /* rows of PL A, each ordered by column (insertion sort) */
    for (int i = 0; i <= n; ++i) cnt[i] = a_rp[i];
    for (int r = 0; r < n; ++r) {
      for (int64_t k = a_rp[r]; k < a_rp[r + 1]; ++k) {
        int c = a_col[k];
        double v = dinv ? a_val[k] * dinv[r] : a_val[k];
        int64_t q = k - 1;
        while (q >= a_rp[r] && tsrc[q] > c) {
          tsrc[q + 1] = tsrc[q];
          tval[q + 1] = tval[q];
          --q;
        }
        tsrc[q + 1] = c;
        tval[q + 1] = v;
      }
    }
  }
  {
    int64_t out = 0;
    d_rp[0] = 0;
    for (int i = 0; i < n; ++i) {
      int64_t beg = out;
      /* tpetra/src/MCLS_TpetraHelpers.hpp:196-263 (applied by multiply()
       * only, i.e. when a preconditioner forms the composite): drop
       * |m| <= (signed row max) * threshold. */
      double row_thr = 0.0;
      int filter = dinv != NULL && cnt[i + 1] > cnt[i];
      if (filter) {
        double mx = tval[cnt[i]];
        for (int64_t p = cnt[i] + 1; p < cnt[i + 1]; ++p)
          if (tval[p] > mx) mx = tval[p];
        row_thr = mx * threshold;
      }
      for (int64_t p = cnt[i]; p < cnt[i + 1]; ++p) {
        if (p > cnt[i] && tsrc[p] == tsrc[p - 1]) { ret = -2; goto done; }
        int j = tsrc[p];
        double m = tval[p];
        if (filter && fabs(m) <= row_thr) continue;
        double h = (j == i) ? 1.0 - omega * m : -omega * m;
        if (fabs(h) < DBL_EPSILON) continue;
        d_col[out] = j;
        d_h[out] = h;
        ++out;
      }
      /* non-normalised CDF, then W and normalisation */
      double acc = 0.0;
      for (int64_t p = beg; p < out; ++p) {
        acc = fabs(d_h[p]) + acc;
        d_cdf[p] = acc;
      }
      d_W[i] = acc;
      for (int64_t p = beg; p < out; ++p) d_cdf[p] = d_cdf[p] / acc;
      d_rp[i + 1] = out;
    }
    ret = out;
  }
done:
  free(cnt);
  free(pos);
  free(tsrc);
  free(tval);
  return ret;
}

int64_t orc_build_domain(int n, const int64_t* a_rp, const int* a_col,
                         const double* a_val, const double* dinv, double omega,
                         double threshold, int64_t* d_rp, int* d_col,
                         double* d_h, double* d_cdf, double* d_W) {
  return build_domain_impl(n, a_rp, a_col, a_val, dinv, omega, threshold, 1,
                           d_rp, d_col, d_h, d_cdf, d_W);
}

int64_t orc_build_domain_forward(int n, const int64_t* a_rp, const int* a_col,
                                 const double* a_val, const double* dinv,
                                 double omega, double threshold, int64_t* d_rp,
                                 int* d_col, double* d_h, double* d_cdf,
                                 double* d_W) {
  return build_domain_impl(n, a_rp, a_col, a_val, dinv, omega, threshold, 0,
                           d_rp, d_col, d_h, d_cdf, d_W);
}

/* Vose's alias method per domain row (M. D. Vose, "A linear algorithm for
 * generating random numbers with a given distribution", IEEE TSE 17(9),
 * 1991), run sequentially per row in a fixed order so that the CPU and the
 * GPU tables are bit-identical.  Build extension: the reference samples
 * rows only by CDF inversion. */
/* One row of Vose's method over weights wts[0..z) with total S (z <= 64):
 * s_k = (wts_k / S) z, small / large stacks popped from the top. */
static void vose_row(const double* wts, double S, int z, double* thr, int* alias) {
  double s[64];
  int small[64], large[64];
  int ns = 0, nl = 0;
  for (int k = 0; k < z; ++k) {
    s[k] = (wts[k] / S) * (double)z;
    if (s[k] < 1.0) small[ns++] = k; else large[nl++] = k;
  }
  while (ns > 0 && nl > 0) {
    int l = small[--ns];
    int g = large[--nl];
    thr[l] = s[l];
    alias[l] = g;
    s[g] = (s[g] + s[l]) - 1.0;
    if (s[g] < 1.0) small[ns++] = g; else large[nl++] = g;
  }
  while (nl > 0) { int g = large[--nl]; thr[g] = 1.0; alias[g] = g; }
  while (ns > 0) { int l = small[--ns]; thr[l] = 1.0; alias[l] = l; }
}

int64_t orc_build_alias(int n, const int64_t* d_rp, const double* d_h,
                        const double* d_W, double* thr, int* alias) {
  double wts[64];
  for (int i = 0; i < n; ++i) {
    int64_t beg = d_rp[i];
    int z = (int)(d_rp[i + 1] - beg);
    if (z > 64) return -(int64_t)(i + 1);
    for (int k = 0; k < z; ++k) wts[k] = fabs(d_h[beg + k]);
    vose_row(wts, d_W[i], z, thr + beg, alias + beg);
  }
  return 0;
}

int orc_alias_stride(int n, const int64_t* d_rp) {
  int64_t longest = 0;
  for (int i = 0; i < n; ++i)
    if (d_rp[i + 1] - d_rp[i] > longest) longest = d_rp[i + 1] - d_rp[i];
  if (longest > 64) return 0;
  return longest > 0 ? (int)longest : 1;
}

/* k_pack_strided (mcls_amd/csrc/kernels.hip) restated: Vose over the row
 * padded to `stride` slots. */
int64_t orc_build_alias_strided(int n, const int64_t* d_rp, const double* d_h,
                                const double* d_W, int stride, double* thr, int* alias) {
  if (stride < 1 || stride > 64) return -1;
  double s[64];
  int small[64], large[64];
  for (int i = 0; i < n; ++i) {
    int64_t beg = d_rp[i];
    int z = (int)(d_rp[i + 1] - beg);
    if (z > stride) return -(int64_t)(i + 1);
    double* t = thr + (int64_t)i * stride;
    int* al = alias + (int64_t)i * stride;
    if (z == 0) {
      for (int k = 0; k < stride; ++k) { t[k] = 0.0; al[k] = 0; }
      continue;
    }
    int ns = 0, nl = 0;
    for (int k = 0; k < stride; ++k) {
      s[k] = k < z ? (fabs(d_h[beg + k]) / d_W[i]) * (double)stride : 0.0;
      if (s[k] < 1.0) small[ns++] = k; else large[nl++] = k;
    }
    while (ns > 0 && nl > 0) {
      int l = small[--ns];
      int g = large[--nl];
      t[l] = s[l];
      al[l] = g;
      s[g] = (s[g] + s[l]) - 1.0;
      if (s[g] < 1.0) small[ns++] = g; else large[nl++] = g;
    }
    while (nl > 0) { int g = large[--nl]; t[g] = 1.0; al[g] = g; }
    while (ns > 0) {
      int l = small[--ns];
      t[l] = l < z ? 1.0 : 0.0;
      al[l] = l < z ? l : 0;
    }
  }
  return 0;
}

int64_t orc_build_proposal_alias(int n, const int64_t* d_rp, const int* d_col,
                                 const double* d_h, const float* g, double* thr,
                                 int* alias, double* lfac) {
  double wts[64];
  for (int i = 0; i < n; ++i) {
    int64_t beg = d_rp[i];
    int z = (int)(d_rp[i + 1] - beg);
    if (z > 64) return -(int64_t)(i + 1);
    double acc = 0.0;
    for (int k = 0; k < z; ++k) {
      double gj = (double)g[d_col[beg + k]];
      if (!(gj > 0.0) || !(gj <= DBL_MAX)) return -(int64_t)(i + 1);
      wts[k] = fabs(d_h[beg + k]) * gj;
      acc = wts[k] + acc;
    }
    vose_row(wts, acc, z, thr + beg, alias + beg);
    for (int k = 0; k < z; ++k) lfac[beg + k] = acc / (double)g[d_col[beg + k]];
  }
  return 0;
}

/* Alias draw: the high 32 bits pick the slot by multiply-shift, the low 32
 * bits decide between the slot and its alias. */
static inline int alias_pick64(uint64_t r, int len, const double* thr, const int* alias) {
  uint32_t slot = (uint32_t)(((r >> 32) * (uint64_t)len) >> 32);
  double u2 = (double)(uint32_t)r * 0x1.0p-32;
  return u2 < thr[slot] ? (int)slot : alias[slot];
}

static inline int alias_pick32(uint64_t r, int len, const float* thr, const int* alias) {
  uint32_t slot = (uint32_t)(((r >> 32) * (uint64_t)len) >> 32);
  float u2 = (float)((uint32_t)r >> 8) * 0x1.0p-24f;
  return u2 < thr[slot] ? (int)slot : alias[slot];
}

/* =============================================================== source == */

static int ceil_log2_u64(uint64_t v) {
  int l = 0;
  while (((uint64_t)1 << l) < v) ++l;
  return l;
}

/* Scale bits S: every q_i <= 2^S and sum_i q_i <= 2^62. */
static int source_scale_bits(int n) { return 62 - ceil_log2_u64((uint64_t)(n > 0 ? n : 1)); }

uint64_t orc_source_quantize(int n, const double* b, uint64_t* q, uint64_t* cdf) {
  double bmax = 0.0;
  for (int i = 0; i < n; ++i) {
    double a = fabs(b[i]);
    if (a > bmax) bmax = a;
  }
  if (!(bmax > 0.0)) {
    for (int i = 0; i < n; ++i) { q[i] = 0; cdf[i] = 0; }
    return 0;
  }
  int e;
  (void)frexp(bmax, &e); /* bmax < 2^e */
  int S = source_scale_bits(n);
  uint64_t acc = 0;
  for (int i = 0; i < n; ++i) {
    double s = ldexp(fabs(b[i]), S - e); /* exact power-of-two scaling */
    uint64_t qi = (uint64_t)ceil(s);
    q[i] = qi;
    acc += qi;
    cdf[i] = acc;
  }
  return acc;
}

/* 53-bit uniform times Q, floored: t in [0, Q). */
static uint64_t mul_u53(uint64_t r, uint64_t Q) {
  unsigned __int128 p = (unsigned __int128)(r >> 11) * (unsigned __int128)Q;
  return (uint64_t)(p >> 53);
}

int orc_source_birth(int n, const double* b, const uint64_t* q,
                     const uint64_t* cdf, uint64_t Q, uint64_t* rng,
                     double* w0) {
  if (Q == 0) return -1;
  uint64_t r = orc_xorshift64(rng);
  uint64_t t = mul_u53(r, Q);
  int64_t k = orc_upper_bound_u64(cdf, n, t);
  /* probability q_k/Q, weight b_k Q/q_k: unbiased for b. */
  *w0 = b[k] * ((double)Q / (double)q[k]);
  return (int)k;
}

/* Stratified (build semantics): n_i = ceil(q_i * nh / Q), M = sum n_i,
 * weight b_i * M / n_i, normalisation M.  Returns M; ncdf inclusive. */
static uint64_t stratified_counts(int n, const uint64_t* q, uint64_t Q,
                                  uint64_t nh, uint64_t* ncdf) {
  uint64_t acc = 0;
  for (int i = 0; i < n; ++i) {
    uint64_t ni = 0;
    if (Q) {
      unsigned __int128 num = (unsigned __int128)q[i] * nh + (Q - 1);
      ni = (uint64_t)(num / Q);
    }
    acc += ni;
    ncdf[i] = acc;
  }
  return acc;
}

/* ============================================================ MC solve === */

typedef struct {
  int n;
  const int64_t* rp;
  const int* col;
  const double* h;
  const double* cdf;
  const double* W;
  float* h32;
  float* cdf32;
  float* W32;
  int learned_alias; /* thr / alias / lfac hold the learned proposal's tables */
  int stride;    /* > 0: thr / thr32 / alias are strided tables (n * stride) */
  double* thr;   /* alias tables (sampler == ORC_SAMPLER_ALIAS) */
  float* thr32;
  int* alias;
  double* lcdf;  /* proposal tables (sampler == ORC_SAMPLER_LEARNED) */
  double* lfac;
  float* lcdf32;
  float* lfac32;
} domain_view;

int orc_build_proposal(int n, const int64_t* d_rp, const int* d_col,
                       const double* d_h, const float* g, double* lcdf,
                       double* lfac) {
  for (int i = 0; i < n; ++i) {
    int64_t beg = d_rp[i], end = d_rp[i + 1];
    double acc = 0.0;
    for (int64_t p = beg; p < end; ++p) {
      double gj = (double)g[d_col[p]];
      if (!(gj > 0.0)) return -(i + 1);
      acc = fabs(d_h[p]) * gj + acc;
      lcdf[p] = acc;
    }
    for (int64_t p = beg; p < end; ++p) {
      lcdf[p] = lcdf[p] / acc;
      lfac[p] = acc / (double)g[d_col[p]];
    }
  }
  return 0;
}

typedef struct {
  const domain_view* d;
  const double* b;
  const orc_mc_params* p;
  uint64_t solve_index;
  const uint64_t* q;     /* random source */
  const uint64_t* qcdf;
  uint64_t Q;
  const uint64_t* ncdf;  /* stratified source */
  uint64_t M;
  int64_t first, last;   /* history range */
  int64_t per_state;     /* forward: histories started in each state */
  double* x;             /* private tally, n entries */
  int64_t steps;
  /* optional path recording */
  int max_len;
  int* paths;
  double* weights;
} walk_job;

/* One history in the build semantics.  Returns the number of transitions. */
static int64_t walk_one(const walk_job* J, int64_t hid, double* x, int* path,
                        int max_len, double* wfinal) {
  const domain_view* d = J->d;
  const orc_mc_params* p = J->p;
  uint64_t s = orc_history_state(p->seed, J->solve_index, (uint64_t)hid);
  int state;
  double w0;
  const int forward = p->estimator == ORC_EST_FORWARD;
  double acc = 0.0; /* forward: sum of w b[state] along the history */
  if (forward) {
    /* UniformForwardSource::getHistory (src/MCLS_UniformForwardSource_impl.hpp:
     * 115-150): per_state consecutive histories per state, weight 1 */
    state = (int)(hid / J->per_state);
    w0 = 1.0;
  } else if (p->source == ORC_SRC_STRATIFIED) {
    int64_t k = orc_upper_bound_u64(J->ncdf, d->n, (uint64_t)hid);
    uint64_t nk = J->ncdf[k] - (k ? J->ncdf[k - 1] : 0);
    state = (int)k;
    w0 = J->b[k] * ((double)J->M / (double)nk);
  } else {
    state = orc_source_birth(d->n, J->b, J->q, J->qcdf, J->Q, &s, &w0);
  }
  const int start = state;
  int64_t L = p->history_length;
  int64_t steps = 0;
  int plen = 0;
  if (path && plen < max_len + 1) path[plen++] = state;
  if (p->precision == 32) {
    float w = (float)w0;
    float wc = (float)p->weight_cutoff * fabsf(w);
    for (;;) {
      int64_t beg = d->rp[state];
      int len = (int)(d->rp[state + 1] - beg);
      if (forward) {
        /* ForwardTally::tallyHistory (src/MCLS_ForwardTally.hpp:128-136) */
        acc = acc + (double)w * J->b[state];
      } else if (p->estimator == ORC_EST_EXPECTED_VALUE) {
        for (int k = 0; k < len; ++k) x[d->col[beg + k]] += (double)(w * d->h32[beg + k]);
      } else {
        x[state] += (double)w;
      }
      if (len == 0) break;
      uint64_t r = orc_xorshift64(&s);
      int k;
      if (d->stride) {  /* strided alias tables: slot floor(u stride) of the row */
        int64_t o = (int64_t)state * d->stride;
        k = alias_pick32(r, d->stride, d->thr32 + o, d->alias + o);
        w = w * d->W32[state];
      } else if (d->learned_alias) {  /* the same proposal drawn by alias */
        k = alias_pick32(r, len, d->thr32 + beg, d->alias + beg);
        w = w * d->lfac32[beg + k];
      } else if (d->lcdf32) {  /* importance-reshaped: P = |h| g / S, factor S / g */
        k = orc_lower_bound_f32(d->lcdf32 + beg, len, orc_u01_f32(r));
        w = w * d->lfac32[beg + k];
      } else {
        k = d->alias ? alias_pick32(r, len, d->thr32 + beg, d->alias + beg)
                     : orc_lower_bound_f32(d->cdf32 + beg, len, orc_u01_f32(r));
        w = w * d->W32[state];
      }
      if (d->h32[beg + k] < 0.0f) w = -w;
      state = d->col[beg + k];
      ++steps;
      if (path && plen < max_len + 1) path[plen++] = state;
      if (steps >= L) break;
      if (fabsf(w) < wc) break;
    }
    if (wfinal) *wfinal = (double)w;
  } else {
    double w = w0;
    double wc = p->weight_cutoff * fabs(w);
    for (;;) {
      int64_t beg = d->rp[state];
      int len = (int)(d->rp[state + 1] - beg);
      if (forward) {
        acc = acc + w * J->b[state];
      } else if (p->estimator == ORC_EST_EXPECTED_VALUE) {
        for (int k = 0; k < len; ++k) x[d->col[beg + k]] += w * d->h[beg + k];
      } else {
        x[state] += w;
      }
      if (len == 0) break;
      uint64_t r = orc_xorshift64(&s);
      int k;
      if (d->stride) {
        int64_t o = (int64_t)state * d->stride;
        k = alias_pick64(r, d->stride, d->thr + o, d->alias + o);
        w = w * d->W[state];
      } else if (d->learned_alias) {
        k = alias_pick64(r, len, d->thr + beg, d->alias + beg);
        w = w * d->lfac[beg + k];
      } else if (d->lcdf) {
        k = orc_lower_bound_f64(d->lcdf + beg, len, orc_u01_f64(r));
        w = w * d->lfac[beg + k];
      } else {
        k = d->alias ? alias_pick64(r, len, d->thr + beg, d->alias + beg)
                     : orc_lower_bound_f64(d->cdf + beg, len, orc_u01_f64(r));
        w = w * d->W[state];
      }
      if (d->h[beg + k] < 0.0) w = -w;
      state = d->col[beg + k];
      ++steps;
      if (path && plen < max_len + 1) path[plen++] = state;
      if (steps >= L) break;
      if (fabs(w) < wc) break;
    }
    if (wfinal) *wfinal = w;
  }
  /* ForwardTally::postProcessHistory: the history's sum goes to its start */
  if (forward) x[start] += acc;
  if (path)
    for (int i = plen; i < max_len + 1; ++i) path[i] = -1;
  return steps;
}

static void* walk_worker(void* arg) {
  walk_job* J = (walk_job*)arg;
  int64_t steps = 0;
  for (int64_t h = J->first; h < J->last; ++h) {
    int* path = J->paths ? J->paths + (size_t)(h - J->first) * (J->max_len + 1) : NULL;
    double* wf = J->weights ? J->weights + (h - J->first) : NULL;
    steps += walk_one(J, h, J->x, path, J->max_len, wf);
  }
  J->steps = steps;
  return NULL;
}

static void domain_view_init(domain_view* d, int n, const int64_t* rp,
                             const int* col, const double* h,
                             const double* cdf, const double* W, int need32,
                             int need_alias, const float* importance, int strided) {
  d->n = n;
  d->stride = 0;
  d->rp = rp;
  d->col = col;
  d->h = h;
  d->cdf = cdf;
  d->W = W;
  d->h32 = d->cdf32 = d->W32 = NULL;
  d->thr = NULL;
  d->thr32 = NULL;
  d->alias = NULL;
  d->lcdf = d->lfac = NULL;
  d->lcdf32 = d->lfac32 = NULL;
  d->learned_alias = 0;
  if (importance && need_alias) {
    /* the learned proposal by alias (ORC_SAMPLER_LEARNED_ALIAS) */
    int64_t nnz = rp[n];
    d->thr = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    d->thr32 = (float*)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    d->alias = (int*)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
    d->lfac = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    d->lfac32 = (float*)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    orc_build_proposal_alias(n, rp, col, h, importance, d->thr, d->alias, d->lfac);
    for (int64_t k = 0; k < nnz; ++k) {
      d->thr32[k] = (float)d->thr[k];
      d->lfac32[k] = (float)d->lfac[k];
    }
    d->learned_alias = 1;
    need_alias = 0;
    importance = NULL;
  }
  if (importance) {
    int64_t nnz = rp[n];
    d->lcdf = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    d->lfac = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    d->lcdf32 = (float*)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    d->lfac32 = (float*)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    orc_build_proposal(n, rp, col, h, importance, d->lcdf, d->lfac);
    for (int64_t k = 0; k < nnz; ++k) {
      d->lcdf32[k] = (float)d->lcdf[k];
      d->lfac32[k] = (float)d->lfac[k];
    }
  }
  if (strided) {
    int stride = orc_alias_stride(n, rp);
    int64_t m = (int64_t)n * (stride > 0 ? stride : 1);
    d->thr = (double*)malloc((size_t)(m ? m : 1) * sizeof(double));
    d->thr32 = (float*)malloc((size_t)(m ? m : 1) * sizeof(float));
    d->alias = (int*)calloc((size_t)(m ? m : 1), sizeof(int));
    if (stride > 0) orc_build_alias_strided(n, rp, h, W, stride, d->thr, d->alias);
    for (int64_t k = 0; k < m; ++k) d->thr32[k] = stride > 0 ? (float)d->thr[k] : 0.0f;
    d->stride = stride;
    need_alias = 0;
  }
  if (need_alias) {
    int64_t nnz = rp[n];
    d->thr = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    d->thr32 = (float*)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    d->alias = (int*)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
    orc_build_alias(n, rp, h, W, d->thr, d->alias);
    for (int64_t k = 0; k < nnz; ++k) d->thr32[k] = (float)d->thr[k];
  }
  if (need32) {
    int64_t nnz = rp[n];
    d->h32 = (float*)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    d->cdf32 = (float*)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    d->W32 = (float*)malloc((size_t)n * sizeof(float));
    /* FP32 tables are the FP64 tables rounded to nearest. */
    for (int64_t k = 0; k < nnz; ++k) {
      d->h32[k] = (float)h[k];
      d->cdf32[k] = (float)cdf[k];
    }
    for (int i = 0; i < n; ++i) d->W32[i] = (float)W[i];
  }
}

static void domain_view_free(domain_view* d) {
  free(d->lcdf);
  free(d->lfac);
  free(d->lcdf32);
  free(d->lfac32);
  free(d->thr);
  free(d->thr32);
  free(d->alias);
  free(d->h32);
  free(d->cdf32);
  free(d->W32);
}

/* Forward source: ceil(num_histories / n) histories per state, at least one
 * (UniformForwardSource: Sample Ratio rounded up per state). */
int64_t orc_forward_per_state(int n, int64_t num_histories) {
  if (n <= 0 || num_histories <= 0) return 1;
  return (num_histories + n - 1) / n;
}

/* Shared driver for orc_mc_solve and orc_record_paths. */
static int run_walks(int n, const int64_t* d_rp, const int* d_col,
                     const double* d_h, const double* d_cdf, const double* d_W,
                     const double* b, const double* mag, const orc_mc_params* p,
                     uint64_t solve_index, double* x, int64_t* steps_out,
                     int64_t rec_first, int64_t rec_count, int max_len,
                     int* paths, double* weights, int64_t win_first, int64_t win_count) {
  domain_view d;
  domain_view_init(&d, n, d_rp, d_col, d_h, d_cdf, d_W, p->precision == 32,
                   p->sampler == ORC_SAMPLER_ALIAS || p->sampler == ORC_SAMPLER_LEARNED_ALIAS,
                   (p->sampler == ORC_SAMPLER_LEARNED || p->sampler == ORC_SAMPLER_LEARNED_ALIAS)
                       ? p->importance : NULL,
                   p->sampler == ORC_SAMPLER_ALIAS_STRIDED);
  uint64_t* q = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
  uint64_t* qcdf = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
  uint64_t* ncdf = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
  /* births drawn by |mag| (a batch's reallocated magnitudes), weights from b */
  uint64_t Q = orc_source_quantize(n, mag ? mag : b, q, qcdf);
  uint64_t nh = (uint64_t)p->num_histories;
  uint64_t M = 0;
  const int forward = p->estimator == ORC_EST_FORWARD;
  int64_t per_state = 1;
  if (forward) {
    per_state = orc_forward_per_state(n, p->num_histories);
    nh = (uint64_t)n * (uint64_t)per_state;
    Q = 1; /* every state starts histories, whatever b is */
  } else if (p->source == ORC_SRC_STRATIFIED) {
    M = stratified_counts(n, q, Q, nh, ncdf);
    nh = M;
  }
  int64_t total_steps = 0;
  if (paths) {
    walk_job J = {&d, b, p, solve_index, q, qcdf, Q, ncdf, M,
                  rec_first, rec_first + rec_count, per_state, NULL, 0, max_len, paths, weights};
    double* scratch = (double*)calloc((size_t)n, sizeof(double));
    J.x = scratch;
    if (Q) walk_worker(&J);
    free(scratch);
    total_steps = J.steps;
  } else {
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    /* the histories walked: all, or the window [win_first, win_first + win_count) */
    uint64_t lo = 0, hi = nh;
    if (win_count >= 0) {
      lo = win_first > 0 ? (uint64_t)win_first : 0;
      hi = lo + (uint64_t)win_count < nh ? lo + (uint64_t)win_count : nh;
      if (lo > hi) lo = hi;
    }
    const uint64_t span = hi - lo;
    int nt = (int)(p->num_threads > 0 ? p->num_threads : 1);
    if ((uint64_t)nt > span) nt = span > 0 ? (int)span : 1;
    walk_job* jobs = (walk_job*)calloc((size_t)nt, sizeof(walk_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nt, sizeof(pthread_t));
    for (int t = 0; t < nt; ++t) {
      walk_job J = {&d, b, p, solve_index, q, qcdf, Q, ncdf, M,
                    (int64_t)(lo + span * (uint64_t)t / (uint64_t)nt),
                    (int64_t)(lo + span * (uint64_t)(t + 1) / (uint64_t)nt),
                    per_state, NULL, 0, 0, NULL, NULL};
      J.x = (double*)calloc((size_t)n, sizeof(double));
      jobs[t] = J;
    }
    if (Q) {
      for (int t = 1; t < nt; ++t) pthread_create(&th[t], NULL, walk_worker, &jobs[t]);
      walk_worker(&jobs[0]);
      for (int t = 1; t < nt; ++t) pthread_join(th[t], NULL);
    }
    for (int t = 0; t < nt; ++t) {
      for (int i = 0; i < n; ++i) x[i] += jobs[t].x[i];
      total_steps += jobs[t].steps;
      free(jobs[t].x);
    }
    free(jobs);
    free(th);
    /* src/MCLS_AdjointTally_impl.hpp:69-72: scale by 1/nh; forward:
     * average over the histories of each start state
     * (src/MCLS_ForwardTally_impl.hpp:174-222) */
    if (nh > 0) {
      double inv = 1.0 / (double)(forward ? (uint64_t)per_state : nh);
      for (int i = 0; i < n; ++i) x[i] = x[i] * inv;
    }
    /* the deterministic birth term belongs to the whole solve, not a window */
    if (p->estimator == ORC_EST_EXPECTED_VALUE && win_count < 0)
      for (int i = 0; i < n; ++i) x[i] = x[i] + b[i];
  }
  if (steps_out) *steps_out = total_steps;
  free(q);
  free(qcdf);
  free(ncdf);
  domain_view_free(&d);
  return 0;
}

int orc_mc_solve(int n, const int64_t* d_rp, const int* d_col,
                 const double* d_h, const double* d_cdf, const double* d_W,
                 const double* b, const orc_mc_params* p,
                 uint64_t solve_index, double* x, int64_t* steps_out) {
  return run_walks(n, d_rp, d_col, d_h, d_cdf, d_W, b, NULL, p, solve_index, x,
                   steps_out, 0, 0, 0, NULL, NULL, 0, -1);
}

int orc_mc_solve_window(int n, const int64_t* d_rp, const int* d_col,
                        const double* d_h, const double* d_cdf, const double* d_W,
                        const double* b, const double* mag, const orc_mc_params* p,
                        uint64_t solve_index, int64_t first, int64_t count,
                        double* x, int64_t* steps_out) {
  return run_walks(n, d_rp, d_col, d_h, d_cdf, d_W, b, mag, p, solve_index, x,
                   steps_out, 0, 0, 0, NULL, NULL, first, count);
}

int orc_record_paths(int n, const int64_t* d_rp, const int* d_col,
                     const double* d_h, const double* d_cdf, const double* d_W,
                     const double* b, const orc_mc_params* p,
                     uint64_t solve_index, int64_t first, int64_t count,
                     int max_len, int* paths, double* weights) {
  return run_walks(n, d_rp, d_col, d_h, d_cdf, d_W, b, NULL, p, solve_index, NULL,
                   NULL, first, count, max_len, paths, weights, 0, -1);
}

/* =================================================== composite operator == */

void orc_transpose(int n, const int64_t* rp, const int* col, const double* val,
                   int64_t* t_rp, int* t_col, double* t_val) {
  /* counting sort over columns, rows visited in ascending order */
  for (int i = 0; i <= n; ++i) t_rp[i] = 0;
  for (int64_t k = 0; k < rp[n]; ++k) t_rp[col[k] + 1]++;
  for (int i = 0; i < n; ++i) t_rp[i + 1] += t_rp[i];
  int64_t* pos = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
  memcpy(pos, t_rp, ((size_t)n + 1) * sizeof(int64_t));
  for (int r = 0; r < n; ++r) {
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
      int64_t p = pos[col[k]]++;
      t_col[p] = r;
      t_val[p] = val[k];
    }
  }
  free(pos);
}

static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

int64_t orc_spgemm(int n, const int64_t* x_rp, const int* x_col, const double* x_val,
                   const int64_t* y_rp, const int* y_col, const double* y_val,
                   double threshold, int64_t cap, int64_t* c_rp, int* c_col,
                   double* c_val) {
  /* dense accumulator per row: mark[j] == i once row i touched column j */
  int* mark = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  double* acc = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  int* list = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  for (int j = 0; j < n; ++j) mark[j] = -1;
  int64_t out = 0;
  c_rp[0] = 0;
  for (int i = 0; i < n; ++i) {
    int cnt = 0;
    for (int64_t p = x_rp[i]; p < x_rp[i + 1]; ++p) {
      const int k = x_col[p];
      const double a = x_val[p];
      for (int64_t q = y_rp[k]; q < y_rp[k + 1]; ++q) {
        const int j = y_col[q];
        const double prod = a * y_val[q];
        if (mark[j] != i) {
          mark[j] = i;
          acc[j] = prod;
          list[cnt++] = j;
        } else {
          acc[j] = acc[j] + prod;
        }
      }
    }
    qsort(list, (size_t)cnt, sizeof(int), cmp_int);
    double thr = 0.0;
    if (cnt > 0) {
      double mx = acc[list[0]];
      for (int t = 1; t < cnt; ++t)
        if (acc[list[t]] > mx) mx = acc[list[t]];
      thr = mx * threshold;
    }
    for (int t = 0; t < cnt; ++t) {
      const double v = acc[list[t]];
      if (fabs(v) <= thr || v == 0.0) continue;
      if (out >= cap) { out = -1; goto done; }
      c_col[out] = list[t];
      c_val[out] = v;
      ++out;
    }
    c_rp[i + 1] = out;
  }
done:
  free(mark);
  free(acc);
  free(list);
  return out;
}

/* =============================================================== linalg == */

void orc_spmv(int n, const int64_t* rp, const int* col, const double* val,
              const double* x, double* y) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s = s + val[k] * x[col[k]];
    y[i] = s;
  }
}

void orc_prec_residual(int n, const int64_t* rp, const int* col,
                       const double* val, const double* dinv, const double* b,
                       const double* x, double* r) {
  /* src/MCLS_LinearProblem_impl.hpp:411-442 (no right preconditioner) */
  orc_spmv(n, rp, col, val, x, r);
  for (int i = 0; i < n; ++i) {
    double t = b[i] - r[i];
    r[i] = dinv ? dinv[i] * t : t;
  }
}

double orc_norm2(int n, const double* v) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + v[i] * v[i];
  return sqrt(s);
}

double orc_norm1(int n, const double* v) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + fabs(v[i]);
  return s;
}

double orc_norminf(int n, const double* v) {
  double s = 0.0;
  for (int i = 0; i < n; ++i)
    if (fabs(v[i]) > s) s = fabs(v[i]);
  return s;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ============================================= reference MC semantics ==== */

int64_t orc_ref_build_source(int n, const double* b, double nh_total_d,
                             int stratified, int ref_quirk, uint64_t* rng,
                             int64_t* counts) {
  /* src/MCLS_UniformAdjointSource_impl.hpp:59-109: d_weight = norm1(b);
   * nh_total = (int) (N * ratio) computed by the caller. */
  int64_t nh_total = (int64_t)nh_total_d;
  double weight = orc_norm1(n, b);
  for (int i = 0; i < n; ++i) counts[i] = 0;
  if (!stratified) {
    /* buildRandomSource :182-224 */
    double* cdf = (double*)malloc((size_t)n * sizeof(double));
    cdf[0] = ref_quirk ? b[0] : fabs(b[0]);
    for (int i = 1; i < n; ++i) cdf[i] = cdf[i - 1] + fabs(b[i]);
    double back = cdf[n - 1];
    int64_t nh_domain = (int64_t)((double)nh_total * back / weight);
    for (int i = 0; i < n; ++i) cdf[i] = cdf[i] / back;
    for (int64_t h = 0; h < nh_domain; ++h) {
      double u = orc_u01_f64(orc_xorshift64(rng));
      counts[orc_lower_bound_f64(cdf, n, u)]++;
    }
    free(cdf);
    return nh_domain;
  }
  /* buildStratifiedSource :231-263 */
  double local_sum = 0.0;
  for (int i = 0; i < n; ++i) local_sum = local_sum + fabs(b[i]);
  int64_t nh_local = (int64_t)ceil((double)nh_total * local_sum / weight);
  double num_over_sum = (double)nh_local / local_sum;
  int64_t total = 0;
  for (int i = 0; i < n; ++i) {
    int64_t c = (int64_t)ceil(fabs(b[i]) * num_over_sum);
    counts[i] = c;
    total += c;
  }
  return total;
}

int orc_ref_transport(const int64_t* d_rp, const int* d_col, const double* d_h,
                      const double* d_cdf, const double* d_W,
                      int64_t history_length, double weight_cutoff,
                      uint64_t* rng, int* state, double* weight,
                      int64_t* steps, double* x) {
  /* src/MCLS_DomainTransporter_impl.hpp:85-115 */
  int st = *state;
  double w = *weight;
  double wc = weight_cutoff * fabs(w);
  int64_t s = *steps;
  int event = 1;
  for (;;) {
    x[st] += w; /* AdjointTally::tallyHistory */
    int64_t beg = d_rp[st];
    int len = (int)(d_rp[st + 1] - beg);
    if (len == 0) { event = 3; break; }
    /* AlmostOptimalDomain::processTransition */
    double u = orc_u01_f64(orc_xorshift64(rng));
    int k = orc_lower_bound_f64(d_cdf + beg, len, u);
    int sign = d_h[beg + k] > 0.0 ? 1 : -1;
    w = w * (d_W[st] * (double)sign);
    st = d_col[beg + k];
    ++s;
    if (s >= history_length) { event = 1; break; } /* terminateHistory */
    if (fabs(w) < wc) { event = 2; break; }      /* build extension */
  }
  *state = st;
  *weight = w;
  *steps = s;
  return event;
}

int64_t orc_ref_mc_solve(int n, const int64_t* d_rp, const int* d_col,
                         const double* d_h, const double* d_cdf,
                         const double* d_W, const double* b,
                         const orc_mc_params* p, int stratified, int ref_quirk,
                         uint64_t* rng, double* x, int64_t* steps_out) {
  /* src/MCLS_MCSolver_impl.hpp:83-112 */
  int64_t* counts = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  for (int i = 0; i < n; ++i) x[i] = 0.0; /* zeroOut */
  int64_t nh = 0;
  int64_t steps = 0;
  double weight = orc_norm1(n, b);
  if (weight > 0.0) {
    nh = orc_ref_build_source(n, b, (double)p->num_histories, stratified,
                              ref_quirk, rng, counts);
    /* the LIFO stack holds states pushed in ascending order, so histories
     * are emitted from the highest state down (getHistory :149-175). */
    for (int i = n - 1; i >= 0; --i) {
      for (int64_t c = 0; c < counts[i]; ++c) {
        int st = i;
        double w = weight * (b[i] > 0.0 ? 1.0 : -1.0);
        int64_t s = 0;
        orc_ref_transport(d_rp, d_col, d_h, d_cdf, d_W, p->history_length,
                          p->weight_cutoff, rng, &st, &w, &s, x);
        steps += s;
      }
    }
    if (nh > 0) {
      double inv = 1.0 / (double)nh;
      for (int i = 0; i < n; ++i) x[i] = x[i] * inv; /* normalize */
    }
  }
  free(counts);
  if (steps_out) *steps_out = steps;
  return nh;
}

/* ============================================================ solvers ==== */

typedef struct {
  int n;
  const int64_t* rp;
  const int* col;
  const double* h;
  const double* cdf;
  const double* W;
  const double* rhs;
  orc_mc_params mc;
  int quirk;
  uint64_t rng;
  double* x;
  int64_t steps;
  int64_t nh;
} set_job;

static void* ref_set_worker(void* arg) {
  set_job* S = (set_job*)arg;
  S->nh = orc_ref_mc_solve(S->n, S->rp, S->col, S->h, S->cdf, S->W, S->rhs,
                           &S->mc, S->mc.source == ORC_SRC_STRATIFIED, S->quirk,
                           &S->rng, S->x, &S->steps);
  return NULL;
}

int orc_mcsa_solve(int n, const int64_t* rp, const int* col, const double* val,
                   const double* dinv, const double* b, double* x,
                   const orc_solver_params* sp, uint64_t solve_index0,
                   orc_solver_result* res, double* res_hist) {
  double t0 = now_s();
  double mc_secs = 0.0;
  int64_t a_nnz = rp[n];
  int64_t* d_rp = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
  int* d_col = (int*)malloc((size_t)(a_nnz ? a_nnz : 1) * sizeof(int));
  double* d_h = (double*)malloc((size_t)(a_nnz ? a_nnz : 1) * sizeof(double));
  double* d_cdf = (double*)malloc((size_t)(a_nnz ? a_nnz : 1) * sizeof(double));
  double* d_W = (double*)malloc((size_t)n * sizeof(double));
  double* r = (double*)malloc((size_t)n * sizeof(double));
  double* rhs = (double*)malloc((size_t)n * sizeof(double));
  double* dx = (double*)malloc((size_t)n * sizeof(double));
  double* tmp = (double*)malloc((size_t)n * sizeof(double));
  int nsets = (int)(sp->num_sets > 0 ? sp->num_sets : 1);
  set_job* sets = (set_job*)calloc((size_t)nsets, sizeof(set_job));
  pthread_t* th = (pthread_t*)calloc((size_t)nsets, sizeof(pthread_t));
  int rc = 0;
  const int forward = sp->mc.estimator == ORC_EST_FORWARD;
  if (forward && sp->reference_mode) { rc = -3; goto out; }
  int64_t dnnz = forward
                     ? orc_build_domain_forward(n, rp, col, val, dinv, sp->neumann_omega,
                                                sp->threshold, d_rp, d_col, d_h, d_cdf, d_W)
                     : orc_build_domain(n, rp, col, val, dinv, sp->neumann_omega,
                                        sp->threshold, d_rp, d_col, d_h, d_cdf, d_W);
  if (dnnz < 0) { rc = (int)dnnz; goto out; }
  for (int s = 0; s < nsets; ++s) {
    sets[s].n = n;
    sets[s].rp = d_rp;
    sets[s].col = d_col;
    sets[s].h = d_h;
    sets[s].cdf = d_cdf;
    sets[s].W = d_W;
    sets[s].rhs = rhs;
    sets[s].mc = sp->mc;
    sets[s].quirk = (int)sp->ref_quirk;
    /* one persistent stream per set, as the reference's per-rank PRNG */
    sets[s].rng = orc_splitmix64(sp->mc.seed ^ orc_splitmix64((uint64_t)s + 1));
    if (!sets[s].rng) sets[s].rng = 0x9E3779B97F4A7C15ULL;
    sets[s].x = (double*)malloc((size_t)n * sizeof(double));
  }
  /* src/MCLS_MCSASolverManager_impl.hpp:300-318 */
  double source_norm;
  if (dinv) {
    for (int i = 0; i < n; ++i) tmp[i] = dinv[i] * b[i];
    source_norm = orc_norm2(n, tmp);
  } else {
    source_norm = orc_norm2(n, b);
  }
  if (fabs(source_norm) < 10.0 * DBL_EPSILON) source_norm = 1.0;
  double crit = sp->tol * source_norm;
  orc_prec_residual(n, rp, col, val, dinv, b, x, r);
  double rnorm = orc_norm2(n, r);
  if (res_hist) res_hist[0] = rnorm / source_norm;
  int64_t iters = 0, histories = 0, steps = 0;
  int go = 1;
  int64_t check = sp->check_freq > 0 ? sp->check_freq : 1;
  while (go) {
    ++iters;
    for (int64_t l = 0; l < sp->smoother_steps; ++l) {
      /* RichardsonIteration::doOneIteration */
      for (int i = 0; i < n; ++i) x[i] = x[i] + sp->richardson_omega * r[i];
      orc_prec_residual(n, rp, col, val, dinv, b, x, r);
    }
    /* MonteCarloSolverManager::buildMonteCarloSource: rhs = omega * rp */
    for (int i = 0; i < n; ++i) rhs[i] = r[i] * sp->neumann_omega;
    double tm = now_s();
    if (sp->reference_mode) {
      if (nsets > 1) {
        for (int s = 1; s < nsets; ++s) pthread_create(&th[s], NULL, ref_set_worker, &sets[s]);
        ref_set_worker(&sets[0]);
        for (int s = 1; s < nsets; ++s) pthread_join(th[s], NULL);
      } else {
        ref_set_worker(&sets[0]);
      }
      /* blockConstantVectorSum then scale by 1/numSets */
      for (int i = 0; i < n; ++i) dx[i] = 0.0;
      for (int s = 0; s < nsets; ++s) {
        for (int i = 0; i < n; ++i) dx[i] += sets[s].x[i];
        histories += sets[s].nh;
        steps += sets[s].steps;
      }
      for (int i = 0; i < n; ++i) dx[i] = dx[i] * (1.0 / (double)nsets);
    } else {
      int64_t st = 0;
      orc_mc_solve(n, d_rp, d_col, d_h, d_cdf, d_W, rhs, &sp->mc,
                   solve_index0 + (uint64_t)(iters - 1), dx, &st);
      histories += sp->mc.num_histories;
      steps += st;
    }
    mc_secs += now_s() - tm;
    for (int i = 0; i < n; ++i) x[i] = x[i] + dx[i];
    orc_prec_residual(n, rp, col, val, dinv, b, x, r);
    rnorm = orc_norm2(n, r);
    if (res_hist && iters <= sp->max_iters) res_hist[iters] = rnorm / source_norm;
    if (iters % check == 0) go = (rnorm > crit) && (iters < sp->max_iters);
  }
  res->iters = iters;
  res->converged = orc_norm2(n, r) <= crit;
  res->achieved_tol = rnorm / source_norm;
  res->histories = histories;
  res->steps = steps;
  res->seconds_mc = mc_secs;
out:
  for (int s = 0; s < nsets; ++s) free(sets[s].x);
  free(sets);
  free(th);
  free(d_rp);
  free(d_col);
  free(d_h);
  free(d_cdf);
  free(d_W);
  free(r);
  free(rhs);
  free(dx);
  free(tmp);
  res->seconds_total = now_s() - t0;
  return rc;
}

int orc_fixed_point_solve(int n, const int64_t* rp, const int* col,
                          const double* val, const double* dinv,
                          const double* b, double* x,
                          const orc_solver_params* sp, orc_solver_result* res) {
  /* src/MCLS_FixedPointSolverManager_impl.hpp:208-320 */
  double t0 = now_s();
  double* r = (double*)malloc((size_t)n * sizeof(double));
  double* tmp = (double*)malloc((size_t)n * sizeof(double));
  double source_norm;
  if (dinv) {
    for (int i = 0; i < n; ++i) tmp[i] = dinv[i] * b[i];
    source_norm = orc_norminf(n, tmp);
  } else {
    source_norm = orc_norminf(n, b);
  }
  if (fabs(source_norm) < 10.0 * DBL_EPSILON) source_norm = 1.0;
  double crit = sp->tol * source_norm;
  orc_prec_residual(n, rp, col, val, dinv, b, x, r);
  double rnorm = orc_norminf(n, r);
  int64_t iters = 0;
  int go = 1;
  int64_t check = sp->check_freq > 0 ? sp->check_freq : 1;
  while (go) {
    ++iters;
    for (int i = 0; i < n; ++i) x[i] = x[i] + sp->richardson_omega * r[i];
    orc_prec_residual(n, rp, col, val, dinv, b, x, r);
    rnorm = orc_norminf(n, r);
    if (iters % check == 0) go = (rnorm > crit) && (iters < sp->max_iters);
  }
  res->iters = iters;
  res->converged = orc_norminf(n, r) <= crit;
  res->achieved_tol = rnorm / source_norm;
  res->histories = 0;
  res->steps = 0;
  res->seconds_mc = 0.0;
  res->seconds_total = now_s() - t0;
  free(r);
  free(tmp);
  return 0;
}
