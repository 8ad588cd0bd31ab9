// C ABI of the gfx950 adjoint MC / MCSA solver (include/mcls_gpu.h).
//
// Host orchestration only: every arithmetic step runs in the kernels of
// kernels.hip on the context's stream.  The MCSA outer loop follows
// src/MCLS_MCSASolverManager_impl.hpp:279-432 (paths relative to the MCLS
// source root); the MC solve follows src/MCLS_MCSolver_impl.hpp:83-112 and
// src/MCLS_MonteCarloSolverManager_impl.hpp:200-235,347-380.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "mcls_gpu.h"
#include "mcls_mlp.h"
#include "mlp.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(MCLS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

#define MCLS_TRY(expr)          \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != MCLS_OK) return rc_; \
  } while (0)

// Device allocation owned by the context.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int ensure(size_t count, const char* what) {
    if (count <= n && p) return MCLS_OK;
    release();
    size_t c = count ? count : 1;
    hipError_t e = hipMalloc((void**)&p, c * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return fail(MCLS_ERR_HIP, std::string("hipMalloc(") + what + "): " + hipGetErrorString(e));
    }
    n = c;
    return MCLS_OK;
  }
};

struct Params {
  double tol = 1.0;
  int64_t max_iters = 1000;
  int64_t print_freq = 10;
  int64_t check_freq = 1;
  int64_t smoother_steps = 1;
  double richardson_omega = 1.0;
  double sample_ratio = 1.0;
  int64_t num_histories = 0;
  int source = mcls::SRC_RANDOM;
  int64_t history_length = 10;
  double neumann_omega = 1.0;
  double threshold = 0.0;
  double weight_cutoff = 0.0;
  int estimator = mcls::EST_COLLISION;
  int precision = 64;
  uint64_t seed = 0;
  int64_t use_graph = 0;
  int64_t verbosity = 0;
  int64_t walk_block = 256;
  int64_t walk_blocks_per_cu = 0;
  int64_t walk_batch = 64;
  int64_t tally_tile = 0;  // LDS tally tile entries per wave (0 = off)
  int tile_kind = mcls::kTileF64;  // tile partial sums: FP64, the walk's precision, or FP32 compensated
  bool residual_csr = false;        // "Residual Format" = "CSR": vector-CSR residual kernel
  int64_t walks_per_lane = 1;  // 2: two histories per lane (strided alias walks)
  int64_t strided_bytes = 16;  // 8: strided alias walks read the 8-byte records
  int birth_order = 0;           // "Birth Order" = "State": random births walked in state order
  int64_t compute_criteria = 0;  // print the convergence criteria after each domain build
  int forward = 0;               // "Monte Carlo Type" = "Forward"
  int sampler = mcls::SAMPLER_CDF;
};

}  // namespace

namespace mcls {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace mcls

struct mcls_gpu_context {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  Params prm;

  // matrix
  int n = 0;
  int64_t nnz = 0;
  int res_lanes = 1;  // lanes per row of the residual kernel
  bool f32 = false;
  bool has_matrix = false;
  DBuf<int> a_rp, a_col;
  DBuf<char> a_val;
  // sliced-ELL copy of A for the residual (slice widths, entry offsets,
  // columns, values in A's precision); sell_on when built
  DBuf<int> sell_w, sell_col;
  DBuf<int64_t> sell_off;
  DBuf<char> sell_val;
  bool sell_on = false;
  bool has_prec = false;
  DBuf<double> dinv;
  DBuf<uint32_t> guide;  // births' guide table, 2^guide_bits + 1 entries
  int guide_bits = 0;
  // general (sparse) left / right preconditioners: FP64 CSR, ascending rows
  struct PrecCsr {
    DBuf<int> rp, col;
    DBuf<double> val;
    int64_t nnz = 0;
    int dim = 0;
    bool on = false;
  } pl_csr, pr_csr;
  DBuf<double> pt1, pt2;  // work vectors of PL (b - A PR x)
  // state order (mcls_gpu_set_state_order): internal state k is the caller's
  // unknown ord_h[k]; vectors cross the C ABI through a gather or a scatter
  bool ordered = false;
  std::vector<int> ord_h, inv_h;
  DBuf<int> ord, ord_inv;
  DBuf<char> ustage;  // host vectors on their way through the permutation

  // domain
  bool domain_valid = false;
  double domain_omega = 0.0, domain_threshold = 0.0;
  int domain_forward = 0;  // the built domain holds rows of H (forward) or of H^T
  int64_t d_nnz = 0;
  DBuf<int> d_rp;
  DBuf<uint32_t> d_col;
  DBuf<double> d_cdf64, d_h64, d_W64;
  DBuf<float> d_cdf32, d_h32, d_W32;
  DBuf<uint4> d_hdr32, d_hdr64, d_slot64;
  DBuf<mcls::Rec32> d_rec32;
  DBuf<uint4> d_tgt32;
  DBuf<mcls::Rec32> d_tgt64;
  bool alias_valid = false;
  DBuf<uint4> d_srec;        // strided alias records, n * d_stride
  DBuf<uint2> d_srec8;       // their 8-byte form (when asked for)
  int d_stride = 0;
  bool strided_valid = false;
  bool strided8 = false;     // d_srec8 holds the current tables
  bool strided16 = false;    // d_srec holds them (8-byte walks with one history per lane skip it)
  DBuf<int> maxlen;          // longest domain row (device scalar)
  bool tgt32_valid = false, tgt64_valid = false;
  // learned proposal: importance g and its transition tables
  DBuf<float> g_imp;
  bool has_importance = false, proposal_valid = false;
  DBuf<double> d_lcdf64, d_lfac64;
  DBuf<float> d_lcdf32, d_lfac32;
  DBuf<mcls::Rec32> d_lrec32;  // the proposal as alias records (FP32 walks)
  int proposal_kind = -1;      // the sampler the proposal tables were built for
  mcls_mlp* model = nullptr;       // recomputes g on each MC source when set
  DBuf<float> feat, mlp_out;       // n x 8 features, n outputs
  DBuf<signed char> imp_mask;      // 0 = importance switched off at that state
  bool has_mask = false;

  // work vectors
  DBuf<double> x, b, r, rhs, part, scal;
  DBuf<double> t64;  // the tally (FP64 for every walk precision)
  DBuf<uint64_t> q, qcdf, ncnt, ncdf, scan_tmp;
  // "Birth Order" = "State": the random births drawn ahead, before and after
  // sorting by state, and the sort's scratch
  DBuf<uint64_t> births, births_sorted;
  DBuf<unsigned char> births_tmp;
  DBuf<unsigned long long> bmax, steps, queue;
  DBuf<int> flags;
  double* host_scal = nullptr;  // pinned, 16 doubles (see kHostM / kHostUnconv)

  uint64_t solve_index = 0;
  mcls_gpu_stats stats{};
  void* scratch = nullptr;  // this device's `scratch` (allocation-free solve)
  std::vector<double> res_hist;  // ||rp||_2 / ||PL b||_2 after each MCSA iteration (0: initial)

  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_sig = 0;
  DBuf<uint64_t> key_dev;               // stream key of the current MC solve
  DBuf<unsigned long long> sidx_dev;    // device copy of solve_index (graph mode)

  // batched systems: block-diagonal operator with rows [sys_off[s], sys_off[s+1])
  int nsys = 0;
  DBuf<int64_t> sys_off;
  DBuf<int> sys_of, sys_conv, sys_unconv;
  DBuf<double> sys_b2, sys_r2, sys_l1, mag;
  // convergence-criteria work (power iteration vectors, scalars, column maxima)
  DBuf<double> pw_x, pw_y, pw_scal;
  DBuf<unsigned long long> pw_cmax, pw_mm;
  std::vector<int64_t> sys_off_h;
  std::vector<double> sys_rel;  // per-system relative residuals of the last solve
};

namespace {

using mcls::WalkArgs;
using mcls::WalkConfig;

uint64_t splitmix64_h(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t solve_key(uint64_t seed, uint64_t solve_index) {
  return splitmix64_h(seed ^ splitmix64_h(solve_index));
}

bool general_prec(const mcls_gpu_context* c) { return c->pl_csr.on || c->pr_csr.on; }

// the learned proposal, sampled by CDF inversion or by alias records
bool learned_sampler(const mcls_gpu_context* c) {
  return c->prm.sampler == mcls::SAMPLER_LEARNED || c->prm.sampler == mcls::SAMPLER_LEARNED_ALIAS;
}

int check_ctx(mcls_gpu_context* c) {
  if (!c) return fail(MCLS_ERR_INVALID_ARG, "null context");
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return fail(MCLS_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  return MCLS_OK;
}

int check_matrix(mcls_gpu_context* c) {
  if (!c->has_matrix) return fail(MCLS_ERR_STATE, "no matrix set (mcls_gpu_set_matrix_csr)");
  return MCLS_OK;
}

hipMemcpyKind kind_to_dev(int ptr_kind) {
  return ptr_kind == MCLS_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
}
hipMemcpyKind kind_from_dev(int ptr_kind) {
  return ptr_kind == MCLS_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
}

int check_kind(int ptr_kind) {
  if (ptr_kind != MCLS_HOST && ptr_kind != MCLS_DEVICE)
    return fail(MCLS_ERR_INVALID_ARG, "ptr_kind must be MCLS_HOST or MCLS_DEVICE");
  return MCLS_OK;
}

// A caller's vector (n rows of w elements of elem bytes) into the internal
// numbering on the context's stream: dst[k] = src[ord[k]].
int vec_in(mcls_gpu_context* c, const void* src, int ptr_kind, void* dst, int w, size_t elem) {
  const size_t bytes = (size_t)c->n * (size_t)w * elem;
  if (!c->ordered) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind_to_dev(ptr_kind), c->stream));
    return MCLS_OK;
  }
  const void* from = src;
  if (ptr_kind != MCLS_DEVICE) {
    if (c->ustage.n < bytes) HIP_TRY(hipStreamSynchronize(c->stream));
    MCLS_TRY(c->ustage.ensure(bytes, "order staging"));
    HIP_TRY(hipMemcpyAsync(c->ustage.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    from = c->ustage.p;
  }
  HIP_TRY(mcls::launch_gather_rows(c->n, w, elem, c->ord.p, from, dst, c->stream));
  return MCLS_OK;
}

// An internal vector out to the caller's numbering: dst[ord[k]] = src[k].
int vec_out(mcls_gpu_context* c, const void* src, void* dst, int ptr_kind, int w, size_t elem) {
  const size_t bytes = (size_t)c->n * (size_t)w * elem;
  if (!c->ordered) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind_from_dev(ptr_kind), c->stream));
    return MCLS_OK;
  }
  if (ptr_kind == MCLS_DEVICE) {
    HIP_TRY(mcls::launch_scatter_rows(c->n, w, elem, c->ord.p, src, dst, c->stream));
    return MCLS_OK;
  }
  if (c->ustage.n < bytes) HIP_TRY(hipStreamSynchronize(c->stream));
  MCLS_TRY(c->ustage.ensure(bytes, "order staging"));
  HIP_TRY(mcls::launch_scatter_rows(c->n, w, elem, c->ord.p, src, c->ustage.p, c->stream));
  HIP_TRY(hipMemcpyAsync(dst, c->ustage.p, bytes, hipMemcpyDeviceToHost, c->stream));
  return MCLS_OK;
}

int read_flags(mcls_gpu_context* c, int out[mcls::FLAG_COUNT]) {
  HIP_TRY(hipMemcpyAsync(out, c->flags.p, sizeof(int) * mcls::FLAG_COUNT, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCLS_OK;
}

int reset_flags(mcls_gpu_context* c) {
  MCLS_TRY(c->flags.ensure(mcls::FLAG_COUNT, "flags"));
  HIP_TRY(mcls::launch_fill_i32(c->flags.p, mcls::FLAG_COUNT, INT_MAX, c->stream));
  return MCLS_OK;
}

int ensure_vectors(mcls_gpu_context* c) {
  size_t n = (size_t)c->n;
  MCLS_TRY(c->x.ensure(n, "x"));
  MCLS_TRY(c->b.ensure(n, "b"));
  MCLS_TRY(c->r.ensure(n, "r"));
  MCLS_TRY(c->rhs.ensure(n, "rhs"));
  MCLS_TRY(c->part.ensure(2 * (size_t)mcls::kResidualPartials + 1, "partials"));
  MCLS_TRY(c->scal.ensure(8, "scalars"));
  MCLS_TRY(c->q.ensure(n, "q"));
  MCLS_TRY(c->qcdf.ensure(n, "qcdf"));
  MCLS_TRY(c->ncnt.ensure(n, "ncnt"));
  MCLS_TRY(c->ncdf.ensure(n, "ncdf"));
  MCLS_TRY(c->scan_tmp.ensure(mcls::scan_tmp_elems((int64_t)n) + 1, "scan tmp"));
  c->guide_bits = mcls::guide_bits_for((int64_t)n);
  MCLS_TRY(c->guide.ensure(((size_t)1 << c->guide_bits) + 1, "birth guide"));
  MCLS_TRY(c->bmax.ensure(1, "bmax"));
  MCLS_TRY(c->steps.ensure(1, "steps"));
  MCLS_TRY(c->queue.ensure(mcls::kQueueWords, "queue"));
  MCLS_TRY(c->key_dev.ensure(1, "key"));
  MCLS_TRY(c->sidx_dev.ensure(1, "solve index"));
  MCLS_TRY(c->mag.ensure(n, "source magnitudes"));
  MCLS_TRY(c->pt1.ensure(n, "prec work 1"));
  MCLS_TRY(c->pt2.ensure(n, "prec work 2"));
  // the tally: FP64 for every walk precision
  if (!c->t64.p || c->t64.n < n) {
    MCLS_TRY(c->t64.ensure(n, "tally64"));
    HIP_TRY(hipMemsetAsync(c->t64.p, 0, n * sizeof(double), c->stream));
  }
  return MCLS_OK;
}

bool tile_on(const mcls_gpu_context* c, bool record) {
  return !record && c->prm.tally_tile > 0 && !c->prm.forward &&
         (c->prm.estimator == mcls::EST_COLLISION || c->prm.estimator == mcls::EST_EXPECTED_VALUE);
}

// The LDS tile kind of the walk (FP64 walks keep FP64 partial sums).
int tile_kind_of(const Params& p) {
  if (p.precision == 64) return mcls::kTileF64;
  return p.tile_kind;
}

// The per-solve buffers of an MC solve: the context's own, or those the
// allocation-free entry point carves out of the static scratch array.
struct SolveBufs {
  uint64_t *q, *qcdf, *ncnt, *ncdf, *scan_tmp;
  unsigned long long *bmax, *steps, *queue;
  double* t64;  // the tally (FP64 for every walk precision)
  uint32_t* guide = nullptr;  // births' guide table (nullable)
  int guide_bits = 0;
  bool may_allocate = true;   // false: the allocation-free entry point's buffers
};

SolveBufs ctx_bufs(mcls_gpu_context* c) {
  return SolveBufs{c->q.p,    c->qcdf.p,  c->ncnt.p,  c->ncdf.p, c->scan_tmp.p,
                   c->bmax.p, c->steps.p, c->queue.p, c->t64.p, c->guide.p, c->guide_bits};
}

void invalidate_graph(mcls_gpu_context* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_sig = 0;
}

// --------------------------------------------------------------- domain
int convergence_criteria(mcls_gpu_context* c, int64_t max_iters, double tol, mcls_gpu_criteria* out);

// ------------------------------------------------------ composite operator
// Read-only view of an FP64 CSR matrix, and an owned one.
struct CsrRef {
  const int* rp;
  const int* col;
  const double* val;
  int64_t nnz;
};
struct DevCsr {
  DBuf<int> rp, col;
  DBuf<double> val;
  int64_t nnz = 0;
  CsrRef ref() const { return CsrRef{rp.p, col.p, val.p, nnz}; }
};

int scan_total(mcls_gpu_context* c, const int* cnt, int* out_rp, int64_t* total) {
  // the row counts are summed on the host in 64 bits first, so that an
  // operator past 2^31 entries is refused instead of wrapping the scan
  const int n = c->n;
  std::vector<int> h((size_t)n);
  HIP_TRY(hipMemcpyAsync(h.data(), cnt, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  int64_t t = 0;
  for (int v : h) t += v;
  if (t > INT32_MAX) return fail(MCLS_ERR_UNSUPPORTED, "composite operator has more than 2^31-1 entries");
  DBuf<int> tmp;
  MCLS_TRY(tmp.ensure(mcls::scan_tmp_elems(n) + 1, "scan tmp"));
  HIP_TRY(mcls::launch_scan_i32_excl(cnt, out_rp, n, tmp.p, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *total = t;
  return MCLS_OK;
}

int check_composite_flags(mcls_gpu_context* c) {
  int flags[mcls::FLAG_COUNT];
  MCLS_TRY(read_flags(c, flags));
  if (flags[mcls::FLAG_BAD_COLUMN] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "column index out of range in row " + std::to_string(flags[mcls::FLAG_BAD_COLUMN] - 1));
  if (flags[mcls::FLAG_DUPLICATE] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "duplicate column in row " + std::to_string(flags[mcls::FLAG_DUPLICATE] - 1));
  if (flags[mcls::FLAG_NONFINITE] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "non-finite matrix value in row " + std::to_string(flags[mcls::FLAG_NONFINITE] - 1));
  return MCLS_OK;
}

// MT::copyTranspose: rows of T ascending in source row.
int csr_transpose(mcls_gpu_context* c, const int* rp, const int* col, const void* val, bool f32, int64_t nnz,
                  DevCsr& T) {
  const int n = c->n;
  hipStream_t s = c->stream;
  DBuf<int> cnt, fill;
  MCLS_TRY(cnt.ensure((size_t)n, "transpose counts"));
  MCLS_TRY(fill.ensure((size_t)n, "transpose fill"));
  MCLS_TRY(T.rp.ensure((size_t)n + 1, "transpose rp"));
  MCLS_TRY(T.col.ensure((size_t)(nnz > 0 ? nnz : 1), "transpose col"));
  MCLS_TRY(T.val.ensure((size_t)(nnz > 0 ? nnz : 1), "transpose val"));
  HIP_TRY(mcls::launch_fill_i32(cnt.p, n, 0, s));
  HIP_TRY(mcls::launch_fill_i32(fill.p, n, 0, s));
  HIP_TRY(mcls::launch_count_cols(n, rp, col, val, f32, cnt.p, c->flags.p, s));
  MCLS_TRY(scan_total(c, cnt.p, T.rp.p, &T.nnz));
  HIP_TRY(mcls::launch_scatter_transpose(n, rp, col, val, f32, nullptr, T.rp.p, fill.p, T.col.p, T.val.p, s));
  HIP_TRY(mcls::launch_sort_rows(n, T.rp.p, T.col.p, T.val.p, c->flags.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  return MCLS_OK;
}

int csr_transpose(mcls_gpu_context* c, const CsrRef& M, DevCsr& T) {
  return csr_transpose(c, M.rp, M.col, M.val, false, M.nnz, T);
}

// C = filter(X Y, threshold) (MT::multiply + TpetraMatrixHelpers::filter).
int csr_multiply(mcls_gpu_context* c, const CsrRef& X, const CsrRef& Y, double threshold, DevCsr& C) {
  const int n = c->n;
  hipStream_t s = c->stream;
  DBuf<int> head, cnt, u_rp, u_col;
  DBuf<double> u_val;
  MCLS_TRY(head.ensure((size_t)(X.nnz > 0 ? X.nnz : 1), "product cursors"));
  MCLS_TRY(cnt.ensure((size_t)n, "product counts"));
  MCLS_TRY(u_rp.ensure((size_t)n + 1, "product rp"));
  HIP_TRY(mcls::launch_spgemm_count(n, X.rp, X.col, Y.rp, Y.col, head.p, cnt.p, s));
  int64_t unnz = 0;
  MCLS_TRY(scan_total(c, cnt.p, u_rp.p, &unnz));
  MCLS_TRY(u_col.ensure((size_t)(unnz > 0 ? unnz : 1), "product col"));
  MCLS_TRY(u_val.ensure((size_t)(unnz > 0 ? unnz : 1), "product val"));
  HIP_TRY(mcls::launch_spgemm_fill(n, X.rp, X.col, X.val, Y.rp, Y.col, Y.val, head.p, u_rp.p, u_col.p, u_val.p, s));
  HIP_TRY(mcls::launch_filter_count(n, u_rp.p, u_val.p, threshold, cnt.p, s));
  MCLS_TRY(C.rp.ensure((size_t)n + 1, "composite rp"));
  MCLS_TRY(scan_total(c, cnt.p, C.rp.p, &C.nnz));
  MCLS_TRY(C.col.ensure((size_t)(C.nnz > 0 ? C.nnz : 1), "composite col"));
  MCLS_TRY(C.val.ensure((size_t)(C.nnz > 0 ? C.nnz : 1), "composite val"));
  HIP_TRY(mcls::launch_filter_fill(n, u_rp.p, u_col.p, u_val.p, threshold, C.rp.p, C.col.p, C.val.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  return MCLS_OK;
}

// The diagonal left preconditioner as a CSR matrix, for the general
// composite that a sparse right preconditioner requires.
int diag_as_csr(mcls_gpu_context* c, DevCsr& D) {
  const int n = c->n;
  std::vector<int> idx((size_t)n + 1);
  for (int i = 0; i <= n; ++i) idx[(size_t)i] = i;
  MCLS_TRY(D.rp.ensure((size_t)n + 1, "diag rp"));
  MCLS_TRY(D.col.ensure((size_t)n, "diag col"));
  MCLS_TRY(D.val.ensure((size_t)n, "diag val"));
  HIP_TRY(hipMemcpy(D.rp.p, idx.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(D.col.p, idx.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(D.val.p, c->dinv.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice));
  D.nnz = n;
  return MCLS_OK;
}

// Rows the domain is built from when a sparse preconditioner is present:
// the composite operator of src/MCLS_LinearProblem_impl.hpp:139-221, the
// products taken in the reference's order and each filtered:
//   forward:  PL (A PR), PL A, A PR
//   adjoint:  PR^T (A^T PL^T), PR^T A^T, A^T PL^T  (rows of the transpose)
int build_composite(mcls_gpu_context* c, bool forward, DevCsr& out) {
  const int n = c->n;
  hipStream_t s = c->stream;
  const double thr = c->prm.threshold;
  const bool has_l = c->pl_csr.on || c->has_prec;
  const bool has_r = c->pr_csr.on;
  DevCsr Ld;  // a diagonal left preconditioner as CSR
  CsrRef L{}, R{};
  if (c->pl_csr.on) {
    L = CsrRef{c->pl_csr.rp.p, c->pl_csr.col.p, c->pl_csr.val.p, c->pl_csr.nnz};
  } else if (c->has_prec) {
    MCLS_TRY(diag_as_csr(c, Ld));
    L = Ld.ref();
  }
  if (has_r) R = CsrRef{c->pr_csr.rp.p, c->pr_csr.col.p, c->pr_csr.val.p, c->pr_csr.nnz};
  MCLS_TRY(reset_flags(c));
  if (forward) {
    // A as FP64 with ascending rows
    DevCsr A;
    MCLS_TRY(A.rp.ensure((size_t)n + 1, "A64 rp"));
    MCLS_TRY(A.col.ensure((size_t)(c->nnz > 0 ? c->nnz : 1), "A64 col"));
    MCLS_TRY(A.val.ensure((size_t)(c->nnz > 0 ? c->nnz : 1), "A64 val"));
    HIP_TRY(mcls::launch_copy_rows(n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, nullptr, A.rp.p, A.col.p, A.val.p,
                                   c->flags.p, s));
    HIP_TRY(mcls::launch_sort_rows(n, A.rp.p, A.col.p, A.val.p, c->flags.p, s));
    A.nnz = c->nnz;
    MCLS_TRY(check_composite_flags(c));
    if (has_l && has_r) {
      DevCsr T;
      MCLS_TRY(csr_multiply(c, A.ref(), R, thr, T));
      MCLS_TRY(csr_multiply(c, L, T.ref(), thr, out));
    } else if (has_l) {
      MCLS_TRY(csr_multiply(c, L, A.ref(), thr, out));
    } else {
      MCLS_TRY(csr_multiply(c, A.ref(), R, thr, out));
    }
  } else {
    DevCsr AT;
    MCLS_TRY(csr_transpose(c, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, c->nnz, AT));
    MCLS_TRY(check_composite_flags(c));
    if (has_l && has_r) {
      DevCsr LT, T, RT;
      MCLS_TRY(csr_transpose(c, L, LT));
      MCLS_TRY(csr_multiply(c, AT.ref(), LT.ref(), thr, T));
      MCLS_TRY(csr_transpose(c, R, RT));
      MCLS_TRY(csr_multiply(c, RT.ref(), T.ref(), thr, out));
    } else if (has_r) {
      DevCsr RT;
      MCLS_TRY(csr_transpose(c, R, RT));
      MCLS_TRY(csr_multiply(c, RT.ref(), AT.ref(), thr, out));
    } else {
      DevCsr LT;
      MCLS_TRY(csr_transpose(c, L, LT));
      MCLS_TRY(csr_multiply(c, AT.ref(), LT.ref(), thr, out));
    }
  }
  return check_composite_flags(c);
}

int build_domain(mcls_gpu_context* c) {
  MCLS_TRY(check_matrix(c));
  const int n = c->n;
  hipStream_t s = c->stream;
  invalidate_graph(c);
  HIP_TRY(hipEventRecord(c->ev[0], s));
  MCLS_TRY(reset_flags(c));
  DBuf<int> t_rp, t_src, keep, itmp;
  DBuf<double> t_val;
  MCLS_TRY(keep.ensure((size_t)n, "keep"));
  MCLS_TRY(itmp.ensure(mcls::scan_tmp_elems(n) + 1, "scan tmp"));
  const double* dinv = c->has_prec ? c->dinv.p : nullptr;
  const bool forward = c->prm.forward != 0;
  // Source rows of the domain: the composite operator (rows of PL A, or of
  // its transpose for the adjoint method), as (src_rp, src_col, src_val).
  const bool general = general_prec(c);
  DevCsr comp;
  const int* src_rp;
  int* src_col;
  double* src_val;
  if (general) {
    MCLS_TRY(build_composite(c, forward, comp));
    MCLS_TRY(reset_flags(c));
    src_rp = comp.rp.p;
    src_col = comp.col.p;
    src_val = comp.val.p;
  } else {
    DBuf<int> cnt, fill;
    MCLS_TRY(cnt.ensure((size_t)n, "col counts"));
    MCLS_TRY(fill.ensure((size_t)n, "fill"));
    MCLS_TRY(t_rp.ensure((size_t)n + 1, "transpose rp"));
    MCLS_TRY(t_src.ensure((size_t)c->nnz, "transpose src"));
    MCLS_TRY(t_val.ensure((size_t)c->nnz, "transpose val"));
    HIP_TRY(mcls::launch_fill_i32(cnt.p, n, 0, s));
    HIP_TRY(mcls::launch_fill_i32(fill.p, n, 0, s));
    if (forward) {
      // src/MCLS_MonteCarloSolverManager_impl.hpp:243-248: rows of PL A
      HIP_TRY(mcls::launch_copy_rows(n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, dinv, t_rp.p, t_src.p, t_val.p,
                                     c->flags.p, s));
    } else {
      HIP_TRY(mcls::launch_count_cols(n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, cnt.p, c->flags.p, s));
      HIP_TRY(mcls::launch_scan_i32_excl(cnt.p, t_rp.p, n, itmp.p, s));
      HIP_TRY(mcls::launch_scatter_transpose(n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, dinv, t_rp.p, fill.p,
                                             t_src.p, t_val.p, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    src_rp = t_rp.p;
    src_col = t_src.p;
    src_val = t_val.p;
  }
  // with a diagonal PL the multiply's row filter runs here; the general
  // composite has filtered inside its products
  const bool filt = c->has_prec && !general;
  const double omega = c->prm.neumann_omega, thr = c->prm.threshold;
  HIP_TRY(mcls::launch_domain_rows(n, src_rp, src_col, src_val, omega, thr, filt, keep.p, c->flags.p, s));
  MCLS_TRY(c->d_rp.ensure((size_t)n + 1, "domain rp"));
  HIP_TRY(mcls::launch_scan_i32_excl(keep.p, c->d_rp.p, n, itmp.p, s));
  int flags[mcls::FLAG_COUNT];
  MCLS_TRY(read_flags(c, flags));
  if (flags[mcls::FLAG_BAD_COLUMN] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "column index out of range in row " + std::to_string(flags[mcls::FLAG_BAD_COLUMN] - 1));
  if (flags[mcls::FLAG_DUPLICATE] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "duplicate column in domain row " + std::to_string(flags[mcls::FLAG_DUPLICATE] - 1));
  if (flags[mcls::FLAG_NONFINITE] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "non-finite matrix value in row " + std::to_string(flags[mcls::FLAG_NONFINITE] - 1));
  int dnnz = 0;
  HIP_TRY(hipMemcpy(&dnnz, c->d_rp.p + n, sizeof(int), hipMemcpyDeviceToHost));
  c->d_nnz = dnnz;
  size_t m = (size_t)(dnnz > 0 ? dnnz : 1);
  MCLS_TRY(c->d_col.ensure(m, "domain col"));
  MCLS_TRY(c->d_h64.ensure(m, "domain h64"));
  MCLS_TRY(c->d_h32.ensure(m, "domain h32"));
  MCLS_TRY(c->d_cdf64.ensure(m, "domain cdf64"));
  MCLS_TRY(c->d_cdf32.ensure(m, "domain cdf32"));
  MCLS_TRY(c->d_W64.ensure((size_t)n, "domain W64"));
  MCLS_TRY(c->d_W32.ensure((size_t)n, "domain W32"));
  HIP_TRY(mcls::launch_domain_fill(n, src_rp, src_col, src_val, omega, thr, filt, c->d_rp.p, c->d_col.p,
                                   c->d_h64.p, c->d_h32.p, c->d_cdf64.p, c->d_cdf32.p, c->d_W64.p, c->d_W32.p, s));
  HIP_TRY(hipEventRecord(c->ev[1], s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->stats.ms_domain = ms;
  c->stats.domain_nnz = dnnz;
  MCLS_TRY(c->d_hdr32.ensure((size_t)n, "row headers 32"));
  MCLS_TRY(c->d_hdr64.ensure((size_t)n, "row headers 64"));
  HIP_TRY(mcls::launch_pack_rows(n, c->d_rp.p, c->d_col.p, c->d_h64.p, c->d_W64.p, c->d_W32.p, c->d_hdr32.p,
                                 c->d_hdr64.p, nullptr, nullptr, false, c->flags.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  c->alias_valid = false;
  c->strided_valid = false;
  c->tgt32_valid = false;
  c->tgt64_valid = false;
  c->proposal_valid = false;
  c->domain_valid = true;
  c->domain_omega = omega;
  c->domain_threshold = thr;
  c->domain_forward = forward ? 1 : 0;
  if (c->prm.compute_criteria) {
    mcls_gpu_criteria cr{};
    MCLS_TRY(convergence_criteria(c, 5000, 1e-6, &cr));
    // src/MCLS_AlmostOptimalDomain_impl.hpp:95-112 and :519-531
    std::printf("\n||H||_inf = %.10g\nsum|H|_min = %.10g\n\nH Balance Parameters\nmin: %.10g\nmax: %.10g\nave: %.10g\n",
                cr.inf_norm, cr.min_row_sum, cr.balance_min, cr.balance_max, cr.balance_ave);
    std::printf("\nNeumann/Ulam Convergence Criteria\n---------------------------------\n"
                "rho(H)  = %.10g\nrho(H+) = %.10g\nrho(H*) = %.10g\n---------------------------------\n\n",
                cr.rho_h, cr.rho_habs, cr.rho_hstar);
    std::fflush(stdout);
  }
  return MCLS_OK;
}

// Alias slot records, built on first use of the alias sampler.
int ensure_alias(mcls_gpu_context* c) {
  if (c->prm.sampler != mcls::SAMPLER_ALIAS || c->alias_valid) return MCLS_OK;
  const size_t m = (size_t)(c->d_nnz > 0 ? c->d_nnz : 1);
  MCLS_TRY(c->d_rec32.ensure(m, "alias records 32"));
  MCLS_TRY(c->d_slot64.ensure(m, "alias slots 64"));
  MCLS_TRY(reset_flags(c));
  HIP_TRY(mcls::launch_pack_rows(c->n, c->d_rp.p, c->d_col.p, c->d_h64.p, c->d_W64.p, c->d_W32.p, c->d_hdr32.p,
                                 c->d_hdr64.p, c->d_rec32.p, c->d_slot64.p, true, c->flags.p, c->stream));
  int flags[mcls::FLAG_COUNT];
  MCLS_TRY(read_flags(c, flags));
  if (flags[mcls::FLAG_LONG_ROW] != INT_MAX)
    return fail(MCLS_ERR_UNSUPPORTED, "alias sampler: domain row " + std::to_string(flags[mcls::FLAG_LONG_ROW] - 1) +
                                          " has more than 64 entries (use the CDF sampler)");
  c->alias_valid = true;
  return MCLS_OK;
}

// Strided alias records (FP32 walks, collision or forward estimator), built
// on first use: the stride is the longest row's length.
int ensure_strided(mcls_gpu_context* c) {
  if (c->prm.sampler != mcls::SAMPLER_ALIAS_STRIDED || c->strided_valid) return MCLS_OK;
  if (c->prm.precision != 32)
    return fail(MCLS_ERR_UNSUPPORTED, "Transition Sampler 'Alias Strided' needs FP32 walks (Walk Precision 'FP32')");
  if (!c->prm.forward && c->prm.estimator == mcls::EST_EXPECTED_VALUE)
    return fail(MCLS_ERR_UNSUPPORTED, "Transition Sampler 'Alias Strided' tallies collisions only");
  MCLS_TRY(c->maxlen.ensure(1, "row length"));
  HIP_TRY(mcls::launch_fill_i32(c->maxlen.p, 1, 0, c->stream));
  HIP_TRY(mcls::launch_max_row_len(c->n, c->d_rp.p, c->maxlen.p, c->stream));
  int longest = 0;
  HIP_TRY(hipMemcpyAsync(&longest, c->maxlen.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (longest > 64)
    return fail(MCLS_ERR_UNSUPPORTED, "Transition Sampler 'Alias Strided': a domain row has " + std::to_string(longest) +
                                          " entries (at most 64)");
  const int stride = longest > 0 ? longest : 1;
  const bool want8 = c->prm.strided_bytes == 8;
  if (want8 && (stride > 32 || c->n >= (1 << 26) - 1))
    return fail(MCLS_ERR_UNSUPPORTED, "Strided Record Bytes 8 needs rows of at most 32 entries and fewer than 2^26 - 1 "
                                      "unknowns");
  // the 16-byte records only where a walk reads them: the default, and two
  // histories per lane (k_walk2 has no 8-byte form)
  const bool want16 = !want8 || c->prm.walks_per_lane == 2;
  if (want16)
    MCLS_TRY(c->d_srec.ensure((size_t)c->n * (size_t)stride, "strided alias records"));
  else
    c->d_srec.release();
  if (want8) MCLS_TRY(c->d_srec8.ensure((size_t)c->n * (size_t)stride, "strided alias records (8-byte)"));
  else c->d_srec8.release();
  HIP_TRY(mcls::launch_pack_strided(c->n, c->d_rp.p, c->d_col.p, c->d_h64.p, c->d_W64.p, c->d_W32.p, stride,
                                    want16 ? c->d_srec.p : nullptr, want8 ? c->d_srec8.p : nullptr, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->d_stride = stride;
  c->strided8 = want8;
  c->strided16 = want16;
  c->strided_valid = true;
  return MCLS_OK;
}

int ensure_proposal_buffers(mcls_gpu_context* c) {
  const size_t m = (size_t)(c->d_nnz > 0 ? c->d_nnz : 1);
  if (c->prm.sampler == mcls::SAMPLER_LEARNED_ALIAS) {
    if (c->prm.precision != 32)
      return fail(MCLS_ERR_UNSUPPORTED, "Transition Sampler 'Learned Alias' needs FP32 walks (Walk Precision 'FP32')");
    MCLS_TRY(c->d_lrec32.ensure(m, "proposal alias records"));
  } else {
    MCLS_TRY(c->d_lcdf64.ensure(m, "proposal cdf64"));
    MCLS_TRY(c->d_lcdf32.ensure(m, "proposal cdf32"));
    MCLS_TRY(c->d_lfac64.ensure(m, "proposal factors64"));
    MCLS_TRY(c->d_lfac32.ensure(m, "proposal factors32"));
  }
  MCLS_TRY(c->g_imp.ensure((size_t)c->n, "importance"));
  return MCLS_OK;
}

// The proposal tables of the sampler in use from c->g_imp (on the stream).
int build_proposal_tables(mcls_gpu_context* c) {
  if (c->prm.sampler == mcls::SAMPLER_LEARNED_ALIAS) {
    HIP_TRY(mcls::launch_pack_proposal_alias(c->n, c->d_rp.p, c->d_col.p, c->d_h64.p, c->g_imp.p, c->d_lrec32.p,
                                             c->flags.p, c->stream));
  } else {
    HIP_TRY(mcls::launch_build_proposal(c->n, c->d_rp.p, c->d_col.p, c->d_h64.p, c->g_imp.p, c->d_lcdf64.p,
                                        c->d_lcdf32.p, c->d_lfac64.p, c->d_lfac32.p, c->flags.p, c->stream));
  }
  return MCLS_OK;
}

// Importance features of source v into c->feat (device).
int compute_features(mcls_gpu_context* c, const double* v) {
  MCLS_TRY(c->feat.ensure((size_t)c->n * 8, "features"));
  HIP_TRY(mcls::launch_feature_sums(c->n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, v, c->part.p, c->scal.p + 4,
                                    c->stream));
  HIP_TRY(mcls::launch_features(c->n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, v, c->scal.p + 4, c->feat.p,
                                c->stream));
  return MCLS_OK;
}

// g from the attached network on source v, then the proposal tables; all on
// the stream (capturable).
int model_importance(mcls_gpu_context* c, const double* v) {
  MCLS_TRY(ensure_proposal_buffers(c));
  MCLS_TRY(c->mlp_out.ensure((size_t)c->n, "network output"));
  MCLS_TRY(compute_features(c, v));
  HIP_TRY(mcls::launch_mlp_forward(mcls::mlp_params_device(c->model), c->feat.p, c->n, c->mlp_out.p, c->stream));
  HIP_TRY(mcls::launch_importance(c->n, c->mlp_out.p, c->has_mask ? c->imp_mask.p : nullptr, c->g_imp.p,
                                  c->stream));
  return build_proposal_tables(c);
}

// Learned-proposal tables from the stored importance (rebuilt after a
// domain rebuild or a new g).  With a network attached the tables are
// rebuilt for every MC source instead.
int ensure_proposal(mcls_gpu_context* c) {
  if (!learned_sampler(c) || (c->proposal_valid && c->proposal_kind == c->prm.sampler)) return MCLS_OK;
  if (c->prm.sampler == mcls::SAMPLER_LEARNED_ALIAS) {
    // alias records hold rows of at most 64 entries
    MCLS_TRY(reset_flags(c));
    HIP_TRY(mcls::launch_flag_long_rows(c->n, c->d_rp.p, 64, c->flags.p, c->stream));
    int lf[mcls::FLAG_COUNT];
    MCLS_TRY(read_flags(c, lf));
    if (lf[mcls::FLAG_LONG_ROW] != INT_MAX)
      return fail(MCLS_ERR_UNSUPPORTED, "domain row " + std::to_string(lf[mcls::FLAG_LONG_ROW] - 1) +
                                            " has more than 64 entries (use the 'Learned' sampler)");
  }
  if (c->model) return ensure_proposal_buffers(c);
  if (!c->has_importance)
    return fail(MCLS_ERR_STATE, "Transition Sampler 'Learned' needs an importance (mcls_gpu_set_importance)");
  MCLS_TRY(ensure_proposal_buffers(c));
  MCLS_TRY(reset_flags(c));
  MCLS_TRY(build_proposal_tables(c));
  int flags[mcls::FLAG_COUNT];
  MCLS_TRY(read_flags(c, flags));
  if (flags[mcls::FLAG_BAD_IMPORTANCE] != INT_MAX)
    return fail(MCLS_ERR_NUMERIC, "importance g must be finite and > 0 (domain row " +
                                      std::to_string(flags[mcls::FLAG_BAD_IMPORTANCE] - 1) + ")");
  if (flags[mcls::FLAG_LONG_ROW] != INT_MAX)
    return fail(MCLS_ERR_UNSUPPORTED, "domain row " + std::to_string(flags[mcls::FLAG_LONG_ROW] - 1) +
                                          " has more than 64 entries (use the 'Learned' sampler)");
  c->proposal_valid = true;
  c->proposal_kind = c->prm.sampler;
  return MCLS_OK;
}

// Target records of the walk precision in use, for the CDF and learned
// samplers (alias walks read their own records).
int ensure_targets(mcls_gpu_context* c) {
  if (c->prm.sampler == mcls::SAMPLER_ALIAS || c->prm.sampler == mcls::SAMPLER_LEARNED_ALIAS ||
      c->prm.sampler == mcls::SAMPLER_ALIAS_STRIDED)
    return MCLS_OK;
  const size_t m = (size_t)(c->d_nnz > 0 ? c->d_nnz : 1);
  // CDF walks of both precisions read the FP64 records (which hold the CDF
  // value); the FP32 learned walk reads the 16-byte FP32 records
  const bool rec32 = c->prm.precision == 32 && c->prm.sampler == mcls::SAMPLER_LEARNED;
  if (rec32 && !c->tgt32_valid) {
    MCLS_TRY(c->d_tgt32.ensure(m, "target records 32"));
    HIP_TRY(mcls::launch_pack_targets(c->d_nnz, c->d_rp.p, c->d_col.p, c->d_W64.p, c->d_W32.p, c->d_cdf64.p,
                                      c->d_tgt32.p, nullptr, c->stream));
    c->tgt32_valid = true;
  } else if (!rec32 && !c->tgt64_valid) {
    MCLS_TRY(c->d_tgt64.ensure(m, "target records 64"));
    HIP_TRY(mcls::launch_pack_targets(c->d_nnz, c->d_rp.p, c->d_col.p, c->d_W64.p, c->d_W32.p, c->d_cdf64.p,
                                      nullptr, c->d_tgt64.p, c->stream));
    c->tgt64_valid = true;
  }
  return MCLS_OK;
}

int ensure_domain_impl(mcls_gpu_context* c);
int ensure_domain(mcls_gpu_context* c) {
  MCLS_TRY(ensure_domain_impl(c));
  MCLS_TRY(ensure_alias(c));
  MCLS_TRY(ensure_strided(c));
  MCLS_TRY(ensure_targets(c));
  return ensure_proposal(c);
}

int ensure_domain_impl(mcls_gpu_context* c) {
  if (c->domain_valid && c->domain_omega == c->prm.neumann_omega && c->domain_threshold == c->prm.threshold &&
      c->domain_forward == c->prm.forward)
    return MCLS_OK;
  return build_domain(c);
}

// ------------------------------------------------------------- MC solve
// Forward source (src/MCLS_UniformForwardSource_impl.hpp:78-91): every
// state starts ceil(Sample Ratio) histories, or ceil(histories / N) when the
// count is given.
int64_t forward_per_state(const mcls_gpu_context* c) {
  int64_t k;
  if (c->prm.num_histories > 0)
    k = (c->prm.num_histories + c->n - 1) / (c->n > 0 ? c->n : 1);
  else
    k = (int64_t)std::ceil(c->prm.sample_ratio);
  return k > 0 ? k : 1;
}

int64_t num_histories(const mcls_gpu_context* c) {
  if (c->prm.forward) return (int64_t)c->n * forward_per_state(c);
  if (c->prm.num_histories > 0) return c->prm.num_histories;
  // src/MCLS_UniformAdjointSource_impl.hpp:82-86: N * Sample Ratio (int)
  return (int64_t)((double)c->n * c->prm.sample_ratio);
}

WalkConfig walk_config(const mcls_gpu_context* c, bool record, int64_t nh) {
  WalkConfig wc;
  wc.precision = c->prm.precision;
  wc.estimator = c->prm.forward ? mcls::EST_FORWARD : c->prm.estimator;
  wc.source = c->prm.source;
  wc.sampler = c->prm.sampler;
  wc.record = record;
  wc.tile = !tile_on(c, record) ? mcls::kTileOff : tile_kind_of(c->prm);
  wc.walks_per_lane = (int)c->prm.walks_per_lane;
  wc.block = (int)c->prm.walk_block;
  int64_t per_cu = c->prm.walk_blocks_per_cu > 0 ? c->prm.walk_blocks_per_cu : 8;
  int64_t cap = (int64_t)c->num_cus * per_cu;
  int64_t need = (nh + wc.block - 1) / wc.block;
  if (need < 1) need = 1;
  int64_t g = need < cap ? need : cap;
  g = (g + mcls::kQueueSlots - 1) / mcls::kQueueSlots * mcls::kQueueSlots;
  wc.grid = (int)g;
  return wc;
}

WalkArgs walk_args(mcls_gpu_context* c, const SolveBufs& sb, const double* rhs, uint64_t key, int64_t first,
                   int64_t last) {
  WalkArgs a{};
  a.d.n = c->n;
  a.d.rp = c->d_rp.p;
  a.d.col = c->d_col.p;
  a.d.cdf64 = c->d_cdf64.p;
  a.d.cdf32 = c->d_cdf32.p;
  a.d.W64 = c->d_W64.p;
  a.d.W32 = c->d_W32.p;
  a.d.h64 = c->d_h64.p;
  a.d.h32 = c->d_h32.p;
  a.d.hdr32 = c->d_hdr32.p;
  a.d.hdr64 = c->d_hdr64.p;
  a.d.rec32 = c->d_rec32.p;
  a.d.srec = c->strided16 ? c->d_srec.p : nullptr;
  a.d.srec8 = c->strided8 && c->prm.strided_bytes == 8 ? c->d_srec8.p : nullptr;
  a.d.stride = (uint32_t)c->d_stride;
  a.d.slot64 = c->d_slot64.p;
  a.d.tgt32 = c->d_tgt32.p;
  a.d.tgt64 = c->d_tgt64.p;
  a.d.lcdf64 = c->d_lcdf64.p;
  a.d.lcdf32 = c->d_lcdf32.p;
  a.d.lfac64 = c->d_lfac64.p;
  a.d.lfac32 = c->d_lfac32.p;
  a.d.lrec32 = c->d_lrec32.p;
  a.queue = sb.queue;
  a.rhs = rhs;
  a.q = sb.q;
  a.qcdf = sb.qcdf;
  a.ncdf = sb.ncdf;
  a.key = key;
  a.first = first;
  a.last = last;
  a.history_length = c->prm.history_length;
  a.cutoff = c->prm.weight_cutoff;
  a.tally64 = sb.t64;
  a.tile_slots = (int)c->prm.tally_tile;
  a.tile_shift = 32;
  for (int64_t t = c->prm.tally_tile; t > 1; t >>= 1) --a.tile_shift;
  // with a state order, states close in the grid are close in number, so
  // consecutive states in distinct sets keep a walk cluster's states apart
  // (2.3% faster than the hash on the 256^3 27-point chosen solve)
  a.tile_direct = c->ordered ? 1 : 0;
  a.steps = sb.steps;
  a.batch = (int)c->prm.walk_batch;
  a.per_state = forward_per_state(c);
  a.guide = sb.guide;  // over qcdf (random) or ncdf (stratified), as the source built it
  a.guide_bits = sb.guide_bits;
  return a;
}

// Source tables for rhs (device): q, qcdf (and stratified counts).  Births
// are allocated by |mag| (mag = rhs unless a batch reallocates them).
int build_source(mcls_gpu_context* c, const SolveBufs& sb, const double* rhs, int64_t nh,
                 const double* mag = nullptr) {
  hipStream_t s = c->stream;
  const int n = c->n;
  if (!mag) mag = rhs;
  HIP_TRY(mcls::launch_fill_u64(sb.bmax, 1, 0ULL, s));
  HIP_TRY(mcls::launch_absmax(mag, n, sb.bmax, s));
  HIP_TRY(mcls::launch_quantize(mag, n, sb.bmax, sb.q, s));
  HIP_TRY(mcls::launch_scan_u64(sb.q, sb.qcdf, n, sb.scan_tmp, s));
  if (c->prm.source == mcls::SRC_STRATIFIED) {
    HIP_TRY(mcls::launch_strat_counts(sb.q, n, sb.qcdf + (n - 1), (uint64_t)nh, sb.ncnt, s));
    HIP_TRY(mcls::launch_scan_u64(sb.ncnt, sb.ncdf, n, sb.scan_tmp, s));
    if (sb.guide) HIP_TRY(mcls::launch_build_guide(n, sb.ncdf, sb.guide_bits, sb.guide, s));
  } else if (sb.guide) {
    HIP_TRY(mcls::launch_build_guide(n, sb.qcdf, sb.guide_bits, sb.guide, s));
  }
  return MCLS_OK;
}
int build_source(mcls_gpu_context* c, const double* rhs, int64_t nh, const double* mag = nullptr) {
  return build_source(c, ctx_bufs(c), rhs, nh, mag);
}

// "Birth Order" = "State" (random source, adjoint walks): draw the births of
// history ids [s0, s1) ahead of the walk and sort them by state, so that the
// walk takes its histories in birth-state order - the locality a stratified
// source has by construction - with every history's id, stream and path
// unchanged.  Sets a.order / a.order_base; leaves them null otherwise.
int prepare_birth_order(mcls_gpu_context* c, const SolveBufs& sb, int64_t s0, int64_t s1, WalkArgs& a) {
  a.order = nullptr;
  a.order_base = 0;
  if (!c->prm.birth_order || c->prm.source != mcls::SRC_RANDOM || c->prm.forward || s1 <= s0) return MCLS_OK;
  if (!sb.may_allocate) return MCLS_OK;  // the allocation-free solve walks in id order
  if (c->prm.use_graph)
    return fail(MCLS_ERR_UNSUPPORTED, "Birth Order 'State' is not available with Use Graph");
  const int64_t count = s1 - s0;
  if (count >= ((int64_t)1 << 32)) return fail(MCLS_ERR_UNSUPPORTED, "Birth Order 'State' needs fewer than 2^32 histories");
  // with a guide table the sort key is the draw's guide bucket (about four
  // states each, buckets in state order) and the states are resolved after
  // the sort; else the key is the state itself
  int bits = 1;
  if (sb.guide) bits = sb.guide_bits;
  else
    while (bits < 31 && ((int64_t)1 << bits) < (int64_t)c->n) ++bits;
  hipStream_t s = c->stream;
  MCLS_TRY(c->births.ensure((size_t)count, "birth order"));
  MCLS_TRY(c->births_sorted.ensure((size_t)count, "birth order"));
  size_t tb = 0;
  HIP_TRY(mcls::launch_sort_births(c->births.p, c->births_sorted.p, count, bits, nullptr, &tb, s));
  MCLS_TRY(c->births_tmp.ensure(tb > 0 ? tb : 1, "birth order sort"));
  HIP_TRY(mcls::launch_random_births(s0, count, a.key, a.key_dev, sb.qcdf, sb.guide, sb.guide_bits, c->n,
                                     c->births.p, s));
  HIP_TRY(mcls::launch_sort_births(c->births.p, c->births_sorted.p, count, bits, c->births_tmp.p, &tb, s));
  if (sb.guide)
    HIP_TRY(mcls::launch_resolve_births(s0, count, a.key, a.key_dev, sb.qcdf, sb.guide, sb.guide_bits, c->n,
                                        c->births_sorted.p, s));
  a.order = c->births_sorted.p;
  a.order_base = s0;
  return MCLS_OK;
}

// One MC solve of source rhs (device) into the tally; then x (+)= tally/nh.
// accumulate: x += ..., else x = ...  `timed` brackets the source build and
// the walk with events (not inside a graph capture); `key_dev` (nullable)
// makes the walk read its stream key from device memory.
// window < 1: walk only that fraction of the source's histories, a run of
// consecutive history ids from the middle of the queue (the autotuner's
// pilots: same births per state as the full solve, fewer states).
// wcount >= 0: walk only the history ids [wfirst, wfirst + wcount) of the
// full source, still normalised by its full count, without the expected-value
// birth term (a window of the solve; windows of a partition sum to it).
// mag (nullable): the magnitudes births are drawn by, when they are not |rhs|
// (a batch's reallocation to its unconverged systems, set by the MCSA
// iteration that computed them).
int mc_solve_device(mcls_gpu_context* c, const SolveBufs& sb, const double* rhs, double* x, bool accumulate,
                    uint64_t sidx, bool timed = true, const uint64_t* key_dev = nullptr, double window = 1.0,
                    int64_t wfirst = 0, int64_t wcount = -1, const double* mag = nullptr) {
  hipStream_t s = c->stream;
  const int64_t nh = num_histories(c);
  if (timed) HIP_TRY(hipEventRecord(c->ev[2], s));
  if (c->prm.sampler == mcls::SAMPLER_LEARNED_ALIAS && c->prm.precision != 32)
    return fail(MCLS_ERR_UNSUPPORTED, "Transition Sampler 'Learned Alias' needs FP32 walks (Walk Precision 'FP32')");
  if (c->prm.sampler == mcls::SAMPLER_ALIAS_STRIDED && !c->strided_valid)
    return fail(MCLS_ERR_STATE, "strided alias records not built");
  if (learned_sampler(c) && c->model) MCLS_TRY(model_importance(c, rhs));
  const bool forward = c->prm.forward != 0;
  if (forward && c->prm.estimator != mcls::EST_COLLISION)
    return fail(MCLS_ERR_UNSUPPORTED, "the forward method tallies collisions only (Estimator Type 'Collision')");
  if (!forward) MCLS_TRY(build_source(c, sb, rhs, nh, mag));
  if (timed) HIP_TRY(hipEventRecord(c->ev[3], s));
  const bool strat = !forward && c->prm.source == mcls::SRC_STRATIFIED;
  const int64_t last = strat ? nh + c->n : nh;
  int64_t first = 0, end = last;
  if (window < 1.0 && last > 0) {
    const int64_t span = std::max<int64_t>(1, (int64_t)((double)last * window));
    first = (last - span) / 2;
    end = first + span;
  }
  if (wcount >= 0) {
    first = std::min(std::max<int64_t>(wfirst, 0), last);
    end = std::min(first + wcount, last);
  }
  WalkArgs a = walk_args(c, sb, rhs, solve_key(c->prm.seed, sidx), first, end);
  a.key_dev = key_dev;
  // ordered births: the whole source's, walked from position first (a pilot
  // window then sees all of its states' births), or only the ids of a
  // history-id window
  if (wcount >= 0)
    MCLS_TRY(prepare_birth_order(c, sb, first, end, a));
  else
    MCLS_TRY(prepare_birth_order(c, sb, 0, last, a));
  WalkConfig wc = walk_config(c, false, end - first);
  if (wc.tile != mcls::kTileOff) {
    const size_t elem = mcls::tile_entry_bytes(wc.tile, c->prm.precision == 32 ? 4 : 8);  // key + partial sum(s)
    if ((size_t)(wc.block / 64) * (size_t)c->prm.tally_tile * elem > 65536)
      return fail(MCLS_ERR_INVALID_ARG, "Tally Tile x waves per block exceeds 64 KiB of LDS per workgroup");
  }
  if (nh > 0 && end > first) HIP_TRY(mcls::launch_walk(a, wc, s));
  if (timed) HIP_TRY(hipEventRecord(c->ev[4], s));
  const bool ev = !forward && c->prm.estimator == mcls::EST_EXPECTED_VALUE && wcount < 0;
  // forward: each state's sum over its own histories (ForwardTally::finalize)
  double inv = nh > 0 ? 1.0 / (double)(forward ? forward_per_state(c) : nh) : 0.0;
  const uint64_t* mtot = strat ? sb.ncdf + (c->n - 1) : nullptr;
  if (accumulate)
    HIP_TRY(mcls::launch_apply_tally(c->n, x, sb.t64, nullptr, rhs, ev, inv, mtot, s));
  else
    HIP_TRY(mcls::launch_store_tally(c->n, x, sb.t64, nullptr, rhs, ev, inv, mtot, s));
  c->stats.walk_launches += nh > 0 ? 1 : 0;
  return MCLS_OK;
}

int mc_solve_device(mcls_gpu_context* c, const double* rhs, double* x, bool accumulate, uint64_t sidx,
                    bool timed = true, const uint64_t* key_dev = nullptr, double window = 1.0) {
  return mc_solve_device(c, ctx_bufs(c), rhs, x, accumulate, sidx, timed, key_dev, window);
}

int accumulate_mc_timing(mcls_gpu_context* c) {
  float a = 0.f, b = 0.f;
  HIP_TRY(hipEventElapsedTime(&a, c->ev[2], c->ev[3]));
  HIP_TRY(hipEventElapsedTime(&b, c->ev[3], c->ev[4]));
  c->stats.ms_source += a;
  c->stats.ms_walk += b;
  c->stats.last_walk_ms = b;
  return MCLS_OK;
}

int histories_done(mcls_gpu_context* c, const SolveBufs& sb, int64_t* out) {
  if (c->prm.forward) {
    *out = num_histories(c);
    return MCLS_OK;
  }
  if (c->prm.source == mcls::SRC_STRATIFIED) {
    uint64_t M = 0;
    HIP_TRY(hipMemcpy(&M, sb.ncdf + (c->n - 1), sizeof(uint64_t), hipMemcpyDeviceToHost));
    *out = (int64_t)M;
  } else {
    uint64_t Q = 0;
    HIP_TRY(hipMemcpy(&Q, sb.qcdf + (c->n - 1), sizeof(uint64_t), hipMemcpyDeviceToHost));
    *out = Q ? num_histories(c) : 0;
  }
  return MCLS_OK;
}
int histories_done(mcls_gpu_context* c, int64_t* out) { return histories_done(c, ctx_bufs(c), out); }

// r = D (b - A x), D diagonal or none, by the sliced-ELL kernel when its copy
// of A is built and selected, else by the vector-CSR one; partial sums of r^2
// to part, *nparts of them.
int unprec_residual(mcls_gpu_context* c, const double* dinv, const double* b, const double* x, double* r,
                    int* nparts) {
  const int n = c->n;
  if (c->sell_on && !c->prm.residual_csr) {
    HIP_TRY(mcls::launch_sell_residual(n, c->sell_w.p, c->sell_off.p, c->sell_col.p, c->sell_val.p, c->f32, dinv, b,
                                       x, r, c->part.p, c->stream));
    *nparts = mcls::sell_blocks((n + 63) / 64);
  } else {
    HIP_TRY(mcls::launch_prec_residual(n, c->res_lanes, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, dinv, b, x, r,
                                       c->part.p, c->stream));
    *nparts = mcls::residual_blocks(n, c->res_lanes);
  }
  return MCLS_OK;
}

// Sliced-ELL copy of A (n x n, rows as stored on the device) for the
// residual.  Skipped, leaving the CSR kernel in use, when padding the slices
// to their longest rows would add more than a quarter to the entries or the
// copy does not fit in device memory.
int build_sell(mcls_gpu_context* c) {
  c->sell_on = false;
  const int n = c->n;
  const int ns = (n + 63) / 64;
  if (ns == 0) return MCLS_OK;
  MCLS_TRY(c->sell_w.ensure((size_t)ns, "sell widths"));
  HIP_TRY(mcls::launch_sell_widths(n, c->a_rp.p, c->sell_w.p, c->stream));
  std::vector<int> w((size_t)ns);
  HIP_TRY(hipMemcpyAsync(w.data(), c->sell_w.p, sizeof(int) * (size_t)ns, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<int64_t> off((size_t)ns);
  int64_t tot = 0;
  for (int t = 0; t < ns; ++t) {
    off[(size_t)t] = tot;
    tot += 64 * (int64_t)w[(size_t)t];
  }
  if (tot > c->nnz + c->nnz / 4 + 64) return MCLS_OK;
  const size_t vsz = c->f32 ? sizeof(float) : sizeof(double);
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  const size_t need = (size_t)tot * (sizeof(int) + vsz) + sizeof(int64_t) * (size_t)ns;
  if (need + (size_t)(1u << 30) > free_b) return MCLS_OK;
  MCLS_TRY(c->sell_off.ensure((size_t)ns, "sell offsets"));
  MCLS_TRY(c->sell_col.ensure((size_t)tot, "sell columns"));
  MCLS_TRY(c->sell_val.ensure((size_t)tot * vsz, "sell values"));
  HIP_TRY(hipMemcpyAsync(c->sell_off.p, off.data(), sizeof(int64_t) * (size_t)ns, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(mcls::launch_sell_fill(n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, c->sell_w.p, c->sell_off.p,
                                 c->sell_col.p, c->sell_val.p, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->sell_on = true;
  return MCLS_OK;
}

// r = PL (b - A PR x) (src/MCLS_LinearProblem_impl.hpp:411-442) with a
// diagonal or sparse PL and an optional sparse PR; partial sums of r^2 go to
// part, *nparts of them.
int prec_residual(mcls_gpu_context* c, const double* b, const double* x, double* r, int* nparts) {
  hipStream_t s = c->stream;
  const int n = c->n;
  const double* xin = x;
  if (c->pr_csr.on) {
    HIP_TRY(mcls::launch_csr_spmv(n, c->pr_csr.rp.p, c->pr_csr.col.p, c->pr_csr.val.p, x, c->pt1.p, s));
    xin = c->pt1.p;
  }
  if (c->pl_csr.on) {
    int np = 0;
    MCLS_TRY(unprec_residual(c, nullptr, b, xin, c->pt2.p, &np));
    HIP_TRY(mcls::launch_csr_spmv(n, c->pl_csr.rp.p, c->pl_csr.col.p, c->pl_csr.val.p, c->pt2.p, r, s));
    HIP_TRY(mcls::launch_scaled_sq(n, nullptr, r, c->part.p, s));
    *nparts = mcls::residual_blocks(n, 1);
  } else {
    MCLS_TRY(unprec_residual(c, c->has_prec ? c->dinv.p : nullptr, b, xin, r, nparts));
  }
  return MCLS_OK;
}

// y = PL v (diagonal, sparse or none)
int apply_left_prec(mcls_gpu_context* c, const double* v, double* y) {
  hipStream_t s = c->stream;
  if (c->pl_csr.on) {
    HIP_TRY(mcls::launch_csr_spmv(c->n, c->pl_csr.rp.p, c->pl_csr.col.p, c->pl_csr.val.p, v, y, s));
  } else if (c->has_prec) {
    HIP_TRY(mcls::launch_diag_scale(c->n, c->dinv.p, v, y, s));
  } else {
    HIP_TRY(mcls::launch_scale_copy(c->n, 1.0, v, y, s));
  }
  return MCLS_OK;
}

// x = PR x (the recovery of the original solution, right preconditioning)
int recover_right_prec(mcls_gpu_context* c, double* x) {
  if (!c->pr_csr.on) return MCLS_OK;
  hipStream_t s = c->stream;
  HIP_TRY(mcls::launch_csr_spmv(c->n, c->pr_csr.rp.p, c->pr_csr.col.p, c->pr_csr.val.p, x, c->pt1.p, s));
  HIP_TRY(mcls::launch_scale_copy(c->n, 1.0, c->pt1.p, x, s));
  return MCLS_OK;
}

// r = PL (b - A PR x), norm^2 into scal[slot] and host_scal[slot]
int residual(mcls_gpu_context* c, const double* b, const double* x, double* r, int slot) {
  hipStream_t s = c->stream;
  int nb = 0;
  MCLS_TRY(prec_residual(c, b, x, r, &nb));
  HIP_TRY(mcls::launch_sum_partials(c->part.p, nb, c->scal.p + slot, s));
  HIP_TRY(hipMemcpyAsync(c->host_scal + slot, c->scal.p + slot, sizeof(double), hipMemcpyDeviceToHost, s));
  return MCLS_OK;
}

// Refuse a non-finite MC source (a NaN or Inf in b or PL): its births and
// weights would be meaningless.  ||v||^2 by the fixed-order reduction into
// scal[1]; synchronises the stream.
int check_finite_source(mcls_gpu_context* c, const double* v) {
  hipStream_t s = c->stream;
  HIP_TRY(mcls::launch_scaled_sq(c->n, nullptr, v, c->part.p, s));
  HIP_TRY(mcls::launch_sum_partials(c->part.p, mcls::residual_blocks(c->n, 1), c->scal.p + 1, s));
  HIP_TRY(hipMemcpyAsync(c->host_scal + 1, c->scal.p + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (!std::isfinite(c->host_scal[1])) return fail(MCLS_ERR_NUMERIC, "non-finite Monte Carlo source");
  return MCLS_OK;
}

int set_param_common(mcls_gpu_context* c) {
  invalidate_graph(c);
  return MCLS_OK;
}

bool key_is(const char* a, const char* b) { return std::strcmp(a, b) == 0; }

// Host-visible scalars in the pinned block host_scal (16 doubles):
//   [0] ||rp||^2   [2] ||PL b||^2   [8] M (stratified histories, bits)
//   [9] unconverged systems (int in the low word)
constexpr int kHostM = 8, kHostUnconv = 9;

// One MCSA iteration (src/MCLS_MCSASolverManager_impl.hpp:356-406): smoother
// steps, MC correction of the preconditioned residual, update, residual and
// its norm.  Everything stays on the stream; with timed = false it contains
// no event records and reads its stream key from key_dev, so it can be
// captured once into a hipGraph and replayed.
int mcsa_iteration(mcls_gpu_context* c, bool timed, const uint64_t* key_dev, uint64_t sidx) {
  hipStream_t s = c->stream;
  const Params& p = c->prm;
  const int n = c->n;
  if (timed) HIP_TRY(hipEventRecord(c->ev[5], s));
  for (int64_t l = 0; l < p.smoother_steps; ++l) {
    // RichardsonIteration::doOneIteration: x += omega rp; rp = PL(b - A x)
    HIP_TRY(mcls::launch_axpy(n, p.richardson_omega, c->r.p, c->x.p, s));
    int np = 0;
    MCLS_TRY(prec_residual(c, c->b.p, c->x.p, c->r.p, &np));
  }
  // buildMonteCarloSource with the internal solver: rhs = omega rp
  if (c->nsys > 0) {
    HIP_TRY(mcls::launch_segment_norm(c->nsys, c->sys_off.p, c->r.p, nullptr, false, c->sys_l1.p, s));
    HIP_TRY(mcls::launch_batch_source(n, c->sys_of.p, c->sys_off.p, c->r.p, p.neumann_omega, c->sys_conv.p,
                                      c->sys_l1.p, c->rhs.p, c->mag.p, s));
  } else {
    HIP_TRY(mcls::launch_scale_copy(n, p.neumann_omega, c->r.p, c->rhs.p, s));
  }
  if (timed) HIP_TRY(hipEventRecord(c->ev[6], s));
  if (key_dev) HIP_TRY(mcls::launch_next_key(p.seed, c->sidx_dev.p, c->key_dev.p, s));
  // x += delta x; a batch draws its births by the reallocated magnitudes
  MCLS_TRY(mc_solve_device(c, ctx_bufs(c), c->rhs.p, c->x.p, true, sidx, timed, key_dev, 1.0, 0, -1,
                           c->nsys > 0 ? c->mag.p : nullptr));
  MCLS_TRY(residual(c, c->b.p, c->x.p, c->r.p, 0));
  if (c->nsys > 0) {
    HIP_TRY(mcls::launch_segment_norm(c->nsys, c->sys_off.p, c->r.p, nullptr, true, c->sys_r2.p, s));
    HIP_TRY(mcls::launch_batch_check(c->nsys, c->sys_r2.p, c->sys_b2.p, p.tol, c->sys_conv.p, c->sys_unconv.p, s));
    HIP_TRY(hipMemcpyAsync(c->host_scal + kHostUnconv, c->sys_unconv.p, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  if (p.source == mcls::SRC_STRATIFIED && !p.forward)
    HIP_TRY(hipMemcpyAsync(c->host_scal + kHostM, c->ncdf.p + (n - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  if (timed) HIP_TRY(hipEventRecord(c->ev[7], s));
  return MCLS_OK;
}

// Spectral radii by power iteration on the domain operator (the growth
// factor averaged over windows of 32 steps in log space, which also holds
// when the dominant eigenvalues are a +-rho or complex pair) and the row /
// column statistics of |H|; src/MCLS_AlmostOptimalDomain_impl.hpp:439-576
// (Anasazi Generalized Davidson there).
int convergence_criteria(mcls_gpu_context* c, int64_t max_iters, double tol, mcls_gpu_criteria* out) {
  hipStream_t s = c->stream;
  const int n = c->n;
  MCLS_TRY(c->pw_x.ensure((size_t)n, "power vector"));
  MCLS_TRY(c->pw_y.ensure((size_t)n, "power vector"));
  MCLS_TRY(c->pw_scal.ensure(8, "power scalars"));
  MCLS_TRY(c->pw_cmax.ensure((size_t)n, "column maxima"));
  MCLS_TRY(c->pw_mm.ensure(4, "row statistics"));
  MCLS_TRY(ensure_vectors(c));
  constexpr int kWindow = 32;
  double* nrm2 = c->pw_scal.p;
  double* logsum = c->pw_scal.p + 1;
  double rho[3] = {0, 0, 0};
  for (int mode = 0; mode < 3; ++mode) {
    double* x = c->pw_x.p;
    double* y = c->pw_y.p;
    HIP_TRY(mcls::launch_power_start(n, 0x5EEDULL + (uint64_t)mode, x, s));
    HIP_TRY(mcls::launch_fill_f64(logsum, 1, 0.0, s));
    double prev_log = 0.0, prev_est = -1.0, est = 0.0;
    int64_t it = 0;
    bool conv = false;
    while (it < max_iters) {
      HIP_TRY(mcls::launch_domain_power(n, c->d_rp.p, c->d_col.p, c->d_h64.p, c->d_W64.p, mode, x,
                                        it == 0 ? nullptr : nrm2, y, c->part.p, nrm2, logsum, s));
      std::swap(x, y);
      ++it;
      if (it % kWindow == 0) {
        double L = 0.0;
        HIP_TRY(hipMemcpyAsync(&L, logsum, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (!std::isfinite(L)) {  // the iterate vanished: nilpotent on the start vector
          est = 0.0;
          conv = true;
          break;
        }
        if (it > kWindow) {
          est = std::exp((L - prev_log) / kWindow);
          if (prev_est >= 0.0 && std::fabs(est - prev_est) <= tol * est) {
            conv = true;
            prev_log = L;
            break;
          }
          prev_est = est;
        }
        prev_log = L;
      }
    }
    rho[mode] = est;
    out->iterations[mode] = it;
    out->converged[mode] = conv ? 1 : 0;
  }
  out->rho_h = rho[0];
  out->rho_habs = rho[1];
  out->rho_hstar = rho[2];
  double* sums = c->pw_scal.p + 2;
  HIP_TRY(mcls::launch_balance(n, c->d_rp.p, c->d_col.p, c->d_h64.p, c->d_W64.p, c->pw_cmax.p, c->pw_mm.p,
                               c->part.p, sums, s));
  unsigned long long mm[4];
  double sm[2];
  HIP_TRY(hipMemcpyAsync(mm, c->pw_mm.p, sizeof(mm), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(sm, sums, sizeof(sm), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  auto as_double = [](unsigned long long b) {
    double d;
    std::memcpy(&d, &b, sizeof(d));
    return d;
  };
  out->inf_norm = as_double(mm[0]);
  out->min_row_sum = as_double(mm[1]);
  out->balance_min = sm[1] > 0 ? as_double(mm[2]) : 0.0;
  out->balance_max = sm[1] > 0 ? as_double(mm[3]) : 0.0;
  out->balance_ave = sm[1] > 0 ? sm[0] / sm[1] : 0.0;
  return MCLS_OK;
}

}  // namespace

// ------------------------------------------ scratch of the allocation-free solve
// mcls_gpu_mc_solve_noalloc keeps every per-call buffer of an MC solve (the
// source tables, the tally, the work-queue and step counters, the scan
// partials and, with a state order, b and x in the internal numbering) in
// this array: 1 MiB for the counters and scan partials, then 64 bytes per
// unknown for at most MCLS_NOALLOC_MAX_N unknowns (65 MiB in all).  Each GPU
// the library runs on holds its own copy until the process ends.  In host
// code `scratch` names a host-side stand-in: its device address on a GPU is
// the one hipGetSymbolAddress returns there, kept per device below.
constexpr size_t kNoAllocFixedBytes = size_t(1) << 20;
constexpr size_t kNoAllocPerState = 64;
constexpr size_t kNoAllocBytes = kNoAllocFixedBytes + kNoAllocPerState * (size_t)MCLS_NOALLOC_MAX_N;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_dev;  // device address of `scratch`, per device

// Once per device, with that device current (mcls_gpu_create).
int resolve_scratch(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize((size_t)device + 1, nullptr);
  if (!g_scratch_dev[(size_t)device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MCLS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[(size_t)device] = p;
  }
  *out = g_scratch_dev[(size_t)device];
  return MCLS_OK;
}

// The domain and the configured sampler's tables exist for the current
// operator and parameters (what ensure_domain would otherwise build).
bool tables_ready(const mcls_gpu_context* c) {
  if (!c->domain_valid || c->domain_omega != c->prm.neumann_omega || c->domain_threshold != c->prm.threshold ||
      c->domain_forward != c->prm.forward)
    return false;
  switch (c->prm.sampler) {
    case mcls::SAMPLER_ALIAS:
      return c->alias_valid;
    case mcls::SAMPLER_ALIAS_STRIDED:
      return c->strided_valid;
    case mcls::SAMPLER_LEARNED_ALIAS:
      return c->proposal_valid && c->proposal_kind == c->prm.sampler;
    case mcls::SAMPLER_LEARNED:
      return c->proposal_valid && c->proposal_kind == c->prm.sampler &&
             (c->prm.precision == 32 ? c->tgt32_valid : c->tgt64_valid);
    default:
      return c->tgt64_valid;
  }
}

}  // namespace

// ================================================================ C ABI
extern "C" {

const char* mcls_gpu_last_error(void) { return g_last_error.c_str(); }

int mcls_gpu_create(mcls_gpu_context** out, int device, unsigned flags) {
  (void)flags;
  if (!out) return fail(MCLS_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(MCLS_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(MCLS_ERR_NO_DEVICE, "device index out of range");
  HIP_TRY(hipSetDevice(device));
  void* scr = nullptr;
  MCLS_TRY(resolve_scratch(device, &scr));
  auto* c = new mcls_gpu_context();
  c->device = device;
  c->scratch = scr;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  hipError_t se = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (se != hipSuccess) {
    delete c;
    return fail(MCLS_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(se));
  }
  for (auto& ev : c->ev) {
    if (hipEventCreate(&ev) != hipSuccess) {
      mcls_gpu_destroy(c);
      return fail(MCLS_ERR_HIP, "hipEventCreate failed");
    }
  }
  if (hipHostMalloc((void**)&c->host_scal, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    c->host_scal = nullptr;
    mcls_gpu_destroy(c);
    return fail(MCLS_ERR_HIP, "hipHostMalloc failed");
  }
  *out = c;
  return MCLS_OK;
}

int mcls_gpu_destroy(mcls_gpu_context* c) {
  if (!c) return MCLS_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  invalidate_graph(c);
  for (auto& ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->host_scal) (void)hipHostFree(c->host_scal);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MCLS_OK;
}

int mcls_gpu_synchronize(mcls_gpu_context* c) {
  MCLS_TRY(check_ctx(c));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCLS_OK;
}

int mcls_gpu_set_state_order(mcls_gpu_context* c, int64_t n, const int32_t* order, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  if (order && (n <= 0 || n > INT32_MAX - 1)) return fail(MCLS_ERR_INVALID_ARG, "n must be in [1, 2^31-2]");
  HIP_TRY(hipStreamSynchronize(c->stream));
  // the operator and everything set for it were numbered the old way
  c->has_matrix = false;
  c->domain_valid = false;
  c->has_prec = false;
  c->pl_csr.on = false;
  c->pr_csr.on = false;
  c->has_importance = false;
  c->proposal_valid = false;
  c->has_mask = false;
  c->nsys = 0;
  invalidate_graph(c);
  c->ordered = false;
  c->ord_h.clear();
  c->inv_h.clear();
  if (!order) return MCLS_OK;
  std::vector<int> o((size_t)n), inv((size_t)n, -1);
  HIP_TRY(hipMemcpy(o.data(), order, sizeof(int) * (size_t)n,
                    ptr_kind == MCLS_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
  for (int64_t k = 0; k < n; ++k) {
    const int v = o[(size_t)k];
    if (v < 0 || v >= n) return fail(MCLS_ERR_INVALID_ARG, "state order entry " + std::to_string(k) + " out of range");
    if (inv[(size_t)v] >= 0)
      return fail(MCLS_ERR_INVALID_ARG, "state order names unknown " + std::to_string(v) + " twice");
    inv[(size_t)v] = (int)k;
  }
  MCLS_TRY(c->ord.ensure((size_t)n, "state order"));
  MCLS_TRY(c->ord_inv.ensure((size_t)n, "state order inverse"));
  MCLS_TRY(c->ustage.ensure((size_t)n * sizeof(double), "order staging"));
  HIP_TRY(hipMemcpy(c->ord.p, o.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->ord_inv.p, inv.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
  c->ord_h.swap(o);
  c->inv_h.swap(inv);
  c->ordered = true;
  return MCLS_OK;
}

int mcls_gpu_set_matrix_csr(mcls_gpu_context* c, int64_t n, int64_t nnz, const int64_t* row_ptr,
                            const int32_t* col, const void* val, int val_dtype, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  if (n <= 0 || n > INT32_MAX - 1) return fail(MCLS_ERR_INVALID_ARG, "n must be in [1, 2^31-2]");
  if (nnz < 0 || nnz > INT32_MAX) return fail(MCLS_ERR_INVALID_ARG, "nnz must be in [0, 2^31-1]");
  if (!row_ptr || (nnz > 0 && (!col || !val))) return fail(MCLS_ERR_INVALID_ARG, "null CSR array");
  if (val_dtype != MCLS_F64 && val_dtype != MCLS_F32) return fail(MCLS_ERR_INVALID_ARG, "val_dtype");
  if (c->ordered && n != (int64_t)c->ord_h.size())
    return fail(MCLS_ERR_INVALID_ARG, "n differs from the length of the state order");
  // Row pointers are validated on the host (monotone, in range) and narrowed
  // to int32.
  std::vector<int64_t> rp64((size_t)n + 1);
  HIP_TRY(hipMemcpy(rp64.data(), row_ptr, sizeof(int64_t) * (size_t)(n + 1),
                    ptr_kind == MCLS_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
  if (rp64[0] != 0 || rp64[(size_t)n] != nnz) return fail(MCLS_ERR_MATRIX, "row_ptr[0] != 0 or row_ptr[n] != nnz");
  std::vector<int> rp32((size_t)n + 1);
  for (int64_t i = 0; i <= n; ++i) {
    if (i > 0 && rp64[(size_t)i] < rp64[(size_t)i - 1]) return fail(MCLS_ERR_MATRIX, "row_ptr not monotone");
    rp32[(size_t)i] = (int)rp64[(size_t)i];
  }
  const size_t vsz = val_dtype == MCLS_F32 ? sizeof(float) : sizeof(double);
  c->has_matrix = false;
  c->domain_valid = false;
  invalidate_graph(c);
  MCLS_TRY(c->a_rp.ensure((size_t)n + 1, "A rp"));
  MCLS_TRY(c->a_col.ensure((size_t)nnz, "A col"));
  MCLS_TRY(c->a_val.ensure((size_t)nnz * vsz, "A val"));
  // With a state order the caller's arrays are validated where they land,
  // then permuted into the context's numbering (P A P^T).
  DBuf<int> src_rp, src_col;
  DBuf<char> src_val;
  int* up_rp = c->a_rp.p;
  int* up_col = c->a_col.p;
  char* up_val = c->a_val.p;
  if (c->ordered) {
    MCLS_TRY(src_rp.ensure((size_t)n + 1, "A rp (caller order)"));
    MCLS_TRY(src_col.ensure((size_t)nnz, "A col (caller order)"));
    MCLS_TRY(src_val.ensure((size_t)nnz * vsz, "A val (caller order)"));
    up_rp = src_rp.p;
    up_col = src_col.p;
    up_val = src_val.p;
  }
  HIP_TRY(hipMemcpy(up_rp, rp32.data(), sizeof(int) * (size_t)(n + 1), hipMemcpyHostToDevice));
  if (nnz > 0) {
    HIP_TRY(hipMemcpy(up_col, col, sizeof(int) * (size_t)nnz, kind_to_dev(ptr_kind)));
    HIP_TRY(hipMemcpy(up_val, val, vsz * (size_t)nnz, kind_to_dev(ptr_kind)));
  }
  MCLS_TRY(reset_flags(c));
  HIP_TRY(mcls::launch_validate_csr((int)n, up_rp, up_col, up_val, val_dtype == MCLS_F32, c->flags.p, c->stream));
  int flags[mcls::FLAG_COUNT];
  MCLS_TRY(read_flags(c, flags));
  if (flags[mcls::FLAG_BAD_COLUMN] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "column index out of range in row " + std::to_string(flags[mcls::FLAG_BAD_COLUMN] - 1));
  if (flags[mcls::FLAG_NONFINITE] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "non-finite matrix value in row " + std::to_string(flags[mcls::FLAG_NONFINITE] - 1));
  if (c->ordered) {
    std::vector<int> rpp((size_t)n + 1);
    rpp[0] = 0;
    for (int64_t k = 0; k < n; ++k) {
      const int r = c->ord_h[(size_t)k];
      rpp[(size_t)k + 1] = rpp[(size_t)k] + (rp32[(size_t)r + 1] - rp32[(size_t)r]);
    }
    HIP_TRY(hipMemcpy(c->a_rp.p, rpp.data(), sizeof(int) * (size_t)(n + 1), hipMemcpyHostToDevice));
    HIP_TRY(mcls::launch_permute_csr((int)n, c->ord.p, c->ord_inv.p, src_rp.p, src_col.p, src_val.p,
                                     val_dtype == MCLS_F32, c->a_rp.p, c->a_col.p, c->a_val.p, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  c->n = (int)n;
  c->nnz = nnz;
  c->f32 = val_dtype == MCLS_F32;
  c->has_matrix = true;
  c->res_lanes = mcls::residual_lanes(n, nnz);
  MCLS_TRY(build_sell(c));
  c->nsys = 0;  // system offsets describe the previous operator
  c->has_importance = false;
  c->proposal_valid = false;
  if (c->has_prec && (int64_t)c->dinv.n < n) c->has_prec = false;
  // sparse preconditioners of another size no longer apply
  if (c->pl_csr.on && c->pl_csr.dim != n) c->pl_csr.on = false;
  if (c->pr_csr.on && c->pr_csr.dim != n) c->pr_csr.on = false;
  MCLS_TRY(ensure_vectors(c));
  return MCLS_OK;
}

int mcls_gpu_set_left_prec_diag(mcls_gpu_context* c, const double* dinv, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (!dinv) return fail(MCLS_ERR_INVALID_ARG, "null dinv");
  MCLS_TRY(c->dinv.ensure((size_t)c->n, "dinv"));
  MCLS_TRY(vec_in(c, dinv, ptr_kind, c->dinv.p, 1, sizeof(double)));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->has_prec = true;
  c->pl_csr.on = false;
  c->domain_valid = false;
  invalidate_graph(c);
  return MCLS_OK;
}

int mcls_gpu_set_left_prec_point_jacobi(mcls_gpu_context* c) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_matrix(c));
  MCLS_TRY(c->dinv.ensure((size_t)c->n, "dinv"));
  MCLS_TRY(reset_flags(c));
  HIP_TRY(mcls::launch_jacobi(c->n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, c->dinv.p, c->flags.p, c->stream));
  int flags[mcls::FLAG_COUNT];
  MCLS_TRY(read_flags(c, flags));
  if (flags[mcls::FLAG_ZERO_DIAG] != INT_MAX) {
    c->has_prec = false;
    return fail(MCLS_ERR_MATRIX, "zero diagonal in row " + std::to_string(flags[mcls::FLAG_ZERO_DIAG] - 1));
  }
  c->has_prec = true;
  c->pl_csr.on = false;
  c->domain_valid = false;
  invalidate_graph(c);
  return MCLS_OK;
}

int mcls_gpu_clear_left_prec(mcls_gpu_context* c) {
  MCLS_TRY(check_ctx(c));
  c->has_prec = false;
  c->pl_csr.on = false;
  c->domain_valid = false;
  invalidate_graph(c);
  return MCLS_OK;
}

namespace {
// A sparse preconditioner: n x n (the operator's n), FP64, each row's
// columns strictly ascending.  Validated on the host and on the device.
int set_prec_csr(mcls_gpu_context* c, mcls_gpu_context::PrecCsr& P, const int64_t* row_ptr, const int32_t* col,
                 const double* val, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (!row_ptr) return fail(MCLS_ERR_INVALID_ARG, "null CSR array");
  const int n = c->n;
  const hipMemcpyKind to_host = ptr_kind == MCLS_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
  std::vector<int64_t> rp64((size_t)n + 1);
  HIP_TRY(hipMemcpy(rp64.data(), row_ptr, sizeof(int64_t) * ((size_t)n + 1), to_host));
  const int64_t nnz = rp64[(size_t)n];
  if (rp64[0] != 0 || nnz < 0 || nnz > INT32_MAX) return fail(MCLS_ERR_MATRIX, "row_ptr[0] != 0 or row_ptr[n] out of range");
  if (nnz > 0 && (!col || !val)) return fail(MCLS_ERR_INVALID_ARG, "null CSR array");
  std::vector<int> rp32((size_t)n + 1), col_h((size_t)(nnz > 0 ? nnz : 1));
  for (int i = 0; i <= n; ++i) {
    if (i > 0 && rp64[(size_t)i] < rp64[(size_t)i - 1]) return fail(MCLS_ERR_MATRIX, "row_ptr not monotone");
    rp32[(size_t)i] = (int)rp64[(size_t)i];
  }
  if (nnz > 0) HIP_TRY(hipMemcpy(col_h.data(), col, sizeof(int) * (size_t)nnz, to_host));
  for (int i = 0; i < n; ++i)
    for (int k = rp32[(size_t)i] + 1; k < rp32[(size_t)i + 1]; ++k)
      if (col_h[(size_t)k] <= col_h[(size_t)k - 1])
        return fail(MCLS_ERR_MATRIX, "preconditioner columns not strictly ascending in row " + std::to_string(i));
  // With a state order: P M P^T on the host, each row re-sorted by column.
  std::vector<double> val_h;
  const double* val_src = val;
  int val_kind = ptr_kind;
  if (c->ordered && nnz > 0) {
    for (int i = 0; i < n; ++i)
      for (int k = rp32[(size_t)i]; k < rp32[(size_t)i + 1]; ++k)
        if (col_h[(size_t)k] < 0 || col_h[(size_t)k] >= n)
          return fail(MCLS_ERR_MATRIX, "column index out of range in row " + std::to_string(i));
    std::vector<double> v0((size_t)nnz);
    HIP_TRY(hipMemcpy(v0.data(), val, sizeof(double) * (size_t)nnz, to_host));
    std::vector<int> rpp((size_t)n + 1), colp((size_t)nnz);
    val_h.resize((size_t)nnz);
    std::vector<std::pair<int, double>> row;
    rpp[0] = 0;
    for (int k = 0; k < n; ++k) {
      const int r = c->ord_h[(size_t)k];
      row.clear();
      for (int q = rp32[(size_t)r]; q < rp32[(size_t)r + 1]; ++q)
        row.emplace_back(c->inv_h[(size_t)col_h[(size_t)q]], v0[(size_t)q]);
      std::sort(row.begin(), row.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
        return a.first < b.first;
      });
      int o = rpp[(size_t)k];
      for (const auto& e : row) {
        colp[(size_t)o] = e.first;
        val_h[(size_t)o] = e.second;
        ++o;
      }
      rpp[(size_t)k + 1] = o;
    }
    rp32.swap(rpp);
    col_h.swap(colp);
    val_src = val_h.data();
    val_kind = MCLS_HOST;
  }
  P.on = false;
  MCLS_TRY(P.rp.ensure((size_t)n + 1, "prec rp"));
  MCLS_TRY(P.col.ensure((size_t)(nnz > 0 ? nnz : 1), "prec col"));
  MCLS_TRY(P.val.ensure((size_t)(nnz > 0 ? nnz : 1), "prec val"));
  HIP_TRY(hipMemcpy(P.rp.p, rp32.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice));
  if (nnz > 0) {
    HIP_TRY(hipMemcpy(P.col.p, col_h.data(), sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P.val.p, val_src, sizeof(double) * (size_t)nnz, kind_to_dev(val_kind)));
  }
  MCLS_TRY(reset_flags(c));
  HIP_TRY(mcls::launch_validate_csr(n, P.rp.p, P.col.p, P.val.p, false, c->flags.p, c->stream));
  int flags[mcls::FLAG_COUNT];
  MCLS_TRY(read_flags(c, flags));
  if (flags[mcls::FLAG_BAD_COLUMN] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "column index out of range in row " + std::to_string(flags[mcls::FLAG_BAD_COLUMN] - 1));
  if (flags[mcls::FLAG_NONFINITE] != INT_MAX)
    return fail(MCLS_ERR_MATRIX, "non-finite value in row " + std::to_string(flags[mcls::FLAG_NONFINITE] - 1));
  P.nnz = nnz;
  P.dim = n;
  P.on = true;
  MCLS_TRY(ensure_vectors(c));
  c->domain_valid = false;
  invalidate_graph(c);
  return MCLS_OK;
}
}  // namespace

int mcls_gpu_set_left_prec_matrix(mcls_gpu_context* c, const int64_t* row_ptr, const int32_t* col, const double* val,
                                  int ptr_kind) {
  MCLS_TRY(set_prec_csr(c, c->pl_csr, row_ptr, col, val, ptr_kind));
  c->has_prec = false;  // replaces a diagonal left preconditioner
  return MCLS_OK;
}

int mcls_gpu_set_right_prec_matrix(mcls_gpu_context* c, const int64_t* row_ptr, const int32_t* col, const double* val,
                                   int ptr_kind) {
  return set_prec_csr(c, c->pr_csr, row_ptr, col, val, ptr_kind);
}

int mcls_gpu_clear_right_prec(mcls_gpu_context* c) {
  MCLS_TRY(check_ctx(c));
  c->pr_csr.on = false;
  c->domain_valid = false;
  invalidate_graph(c);
  return MCLS_OK;
}

int mcls_gpu_get_left_prec_diag(mcls_gpu_context* c, double* dinv, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  if (!c->has_prec) return fail(MCLS_ERR_STATE, c->pl_csr.on ? "the left preconditioner is not diagonal"
                                                           : "no left preconditioner");
  if (!dinv) return fail(MCLS_ERR_INVALID_ARG, "null dinv");
  MCLS_TRY(vec_out(c, c->dinv.p, dinv, ptr_kind, 1, sizeof(double)));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCLS_OK;
}

int mcls_gpu_set_param_i64(mcls_gpu_context* c, const char* key, int64_t v) {
  if (!c || !key) return fail(MCLS_ERR_INVALID_ARG, "null context or key");
  Params& p = c->prm;
  if (key_is(key, "Maximum Iterations")) {
    if (v < 0) return fail(MCLS_ERR_INVALID_ARG, "Maximum Iterations < 0");
    p.max_iters = v;
  } else if (key_is(key, "Iteration Print Frequency")) {
    p.print_freq = v;
  } else if (key_is(key, "Iteration Check Frequency")) {
    if (v < 1) return fail(MCLS_ERR_INVALID_ARG, "Iteration Check Frequency < 1");
    p.check_freq = v;
  } else if (key_is(key, "Smoother Steps")) {
    if (v < 0) return fail(MCLS_ERR_INVALID_ARG, "Smoother Steps < 0");
    p.smoother_steps = v;
  } else if (key_is(key, "History Length")) {
    if (v < 1) return fail(MCLS_ERR_INVALID_ARG, "History Length < 1");
    p.history_length = v;
  } else if (key_is(key, "Number of Histories")) {
    if (v < 0) return fail(MCLS_ERR_INVALID_ARG, "Number of Histories < 0");
    p.num_histories = v;
  } else if (key_is(key, "Random Seed")) {
    p.seed = (uint64_t)v;
  } else if (key_is(key, "Use Graph")) {
    p.use_graph = v;
  } else if (key_is(key, "Verbosity")) {
    p.verbosity = v;
  } else if (key_is(key, "Walk Block Size")) {
    if (v != 64 && v != 128 && v != 256) return fail(MCLS_ERR_INVALID_ARG, "Walk Block Size must be 64, 128 or 256");
    p.walk_block = v;
  } else if (key_is(key, "Walk Blocks Per CU")) {
    if (v < 0 || v > 64) return fail(MCLS_ERR_INVALID_ARG, "Walk Blocks Per CU must be in [0, 64]");
    p.walk_blocks_per_cu = v;
  } else if (key_is(key, "Compute Convergence Criteria")) {
    p.compute_criteria = v ? 1 : 0;
  } else if (key_is(key, "Tally Tile")) {
    if (v != 0 && (v < 64 || v > 2048 || (v & (v - 1)) != 0))
      return fail(MCLS_ERR_INVALID_ARG, "Tally Tile must be 0 or a power of two in [64, 2048]");
    p.tally_tile = v;
    if (c->has_matrix) MCLS_TRY(ensure_vectors(c));
  } else if (key_is(key, "Walks Per Lane")) {
    if (v != 1 && v != 2) return fail(MCLS_ERR_INVALID_ARG, "Walks Per Lane must be 1 or 2");
    if (v == 2 && !c->strided16) c->strided_valid = false;  // its 16-byte records, built on next use
    p.walks_per_lane = v;
  } else if (key_is(key, "Strided Record Bytes")) {
    if (v != 16 && v != 8) return fail(MCLS_ERR_INVALID_ARG, "Strided Record Bytes must be 16 or 8");
    if ((v == 8 && !c->strided8) || (v == 16 && !c->strided16))
      c->strided_valid = false;  // built with the tables on next use
    p.strided_bytes = v;
  } else if (key_is(key, "Walk Batch")) {
    if (v < 64 || v > 4096 || v % 64 != 0)
      return fail(MCLS_ERR_INVALID_ARG, "Walk Batch must be a multiple of 64 in [64, 4096]");
    p.walk_batch = v;
  } else if (key_is(key, "MC Check Frequency") || key_is(key, "MC Buffer Size") ||
             key_is(key, "Overlap Size") || key_is(key, "Number of Sets")) {
    // accepted; single domain, single set
  } else {
    return fail(MCLS_ERR_INVALID_ARG, std::string("unknown integer parameter '") + key + "'");
  }
  return set_param_common(c);
}

int mcls_gpu_set_param_f64(mcls_gpu_context* c, const char* key, double v) {
  if (!c || !key) return fail(MCLS_ERR_INVALID_ARG, "null context or key");
  Params& p = c->prm;
  if (!std::isfinite(v)) return fail(MCLS_ERR_INVALID_ARG, std::string("non-finite value for '") + key + "'");
  if (key_is(key, "Convergence Tolerance")) {
    p.tol = v;
  } else if (key_is(key, "Richardson Relaxation")) {
    p.richardson_omega = v;
  } else if (key_is(key, "Sample Ratio")) {
    if (v < 0) return fail(MCLS_ERR_INVALID_ARG, "Sample Ratio < 0");
    p.sample_ratio = v;
  } else if (key_is(key, "Neumann Relaxation")) {
    if (!(v > 0)) return fail(MCLS_ERR_INVALID_ARG, "Neumann Relaxation must be > 0");
    p.neumann_omega = v;
  } else if (key_is(key, "Composite Operator Threshold")) {
    p.threshold = v;
  } else if (key_is(key, "Weight Cutoff")) {
    if (v < 0) return fail(MCLS_ERR_INVALID_ARG, "Weight Cutoff < 0");
    p.weight_cutoff = v;
  } else if (key_is(key, "History Length")) {
    // getValidParameters declares it a double (MonteCarloSolverManager_impl:105)
    if (v < 1) return fail(MCLS_ERR_INVALID_ARG, "History Length < 1");
    p.history_length = (int64_t)v;
  } else {
    return fail(MCLS_ERR_INVALID_ARG, std::string("unknown double parameter '") + key + "'");
  }
  return set_param_common(c);
}

int mcls_gpu_set_param_str(mcls_gpu_context* c, const char* key, const char* v) {
  if (!c || !key || !v) return fail(MCLS_ERR_INVALID_ARG, "null context, key or value");
  Params& p = c->prm;
  if (key_is(key, "Source Sampling Type")) {
    if (key_is(v, "Random")) p.source = mcls::SRC_RANDOM;
    else if (key_is(v, "Stratified")) p.source = mcls::SRC_STRATIFIED;
    else return fail(MCLS_ERR_INVALID_ARG, std::string("Source Sampling Type '") + v + "'");
  } else if (key_is(key, "Birth Order")) {
    if (key_is(v, "Id")) p.birth_order = 0;
    else if (key_is(v, "State")) p.birth_order = 1;
    else return fail(MCLS_ERR_INVALID_ARG, std::string("Birth Order '") + v + "'");
  } else if (key_is(key, "Estimator Type")) {
    if (key_is(v, "Collision")) p.estimator = mcls::EST_COLLISION;
    else if (key_is(v, "Expected Value")) p.estimator = mcls::EST_EXPECTED_VALUE;
    else return fail(MCLS_ERR_INVALID_ARG, std::string("Estimator Type '") + v + "'");
  } else if (key_is(key, "Monte Carlo Type")) {
    // ForwardTag / AdjointTag of the reference's managers
    if (key_is(v, "Adjoint")) p.forward = 0;
    else if (key_is(v, "Forward")) p.forward = 1;
    else return fail(MCLS_ERR_INVALID_ARG, std::string("Monte Carlo Type '") + v + "'");
  } else if (key_is(key, "Transition Sampler")) {
    if (key_is(v, "CDF")) p.sampler = mcls::SAMPLER_CDF;
    else if (key_is(v, "Alias")) p.sampler = mcls::SAMPLER_ALIAS;
    else if (key_is(v, "Learned")) p.sampler = mcls::SAMPLER_LEARNED;
    else if (key_is(v, "Learned Alias")) p.sampler = mcls::SAMPLER_LEARNED_ALIAS;
    else if (key_is(v, "Alias Strided")) p.sampler = mcls::SAMPLER_ALIAS_STRIDED;
    else return fail(MCLS_ERR_INVALID_ARG, std::string("Transition Sampler '") + v + "'");
  } else if (key_is(key, "Walk Precision")) {
    if (key_is(v, "FP64")) p.precision = 64;
    else if (key_is(v, "FP32")) p.precision = 32;
    else return fail(MCLS_ERR_INVALID_ARG, std::string("Walk Precision '") + v + "'");
    if (c->has_matrix) MCLS_TRY(ensure_vectors(c));
  } else if (key_is(key, "Tally Tile Precision")) {
    if (key_is(v, "FP64")) p.tile_kind = mcls::kTileF64;
    else if (key_is(v, "Walk")) p.tile_kind = mcls::kTileWalk;
    else if (key_is(v, "Compensated")) p.tile_kind = mcls::kTileComp;
    else return fail(MCLS_ERR_INVALID_ARG, std::string("Tally Tile Precision '") + v + "'");
  } else if (key_is(key, "Residual Format")) {
    if (key_is(v, "Sliced ELL")) p.residual_csr = false;
    else if (key_is(v, "CSR")) p.residual_csr = true;
    else return fail(MCLS_ERR_INVALID_ARG, std::string("Residual Format '") + v + "'");
  } else if (key_is(key, "Fixed Point Type")) {
    if (!key_is(v, "Richardson"))
      return fail(MCLS_ERR_UNSUPPORTED, std::string("Fixed Point Type '") + v + "' (only Richardson)");
  } else if (key_is(key, "Transport Type")) {
    if (!key_is(v, "Global") && !key_is(v, "Subdomain"))
      return fail(MCLS_ERR_INVALID_ARG, std::string("Transport Type '") + v + "'");
  } else {
    return fail(MCLS_ERR_INVALID_ARG, std::string("unknown string parameter '") + key + "'");
  }
  return set_param_common(c);
}

int mcls_gpu_build_domain(mcls_gpu_context* c) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(build_domain(c));
  // the transition tables of the configured sampler, so that a solve that
  // follows builds nothing (the learned proposal once it has an importance)
  MCLS_TRY(ensure_alias(c));
  MCLS_TRY(ensure_strided(c));
  MCLS_TRY(ensure_targets(c));
  if (learned_sampler(c) && (c->has_importance || c->model)) MCLS_TRY(ensure_proposal(c));
  return MCLS_OK;
}

int mcls_gpu_domain_nnz(mcls_gpu_context* c, int64_t* nnz) {
  MCLS_TRY(check_ctx(c));
  if (!nnz) return fail(MCLS_ERR_INVALID_ARG, "null nnz");
  if (!c->domain_valid) return fail(MCLS_ERR_STATE, "domain not built");
  *nnz = c->d_nnz;
  return MCLS_OK;
}

int mcls_gpu_get_domain(mcls_gpu_context* c, int64_t* rp, int32_t* col, double* h, double* cdf, double* W) {
  MCLS_TRY(check_ctx(c));
  if (!c->domain_valid) return fail(MCLS_ERR_STATE, "domain not built");
  const size_t n = (size_t)c->n, m = (size_t)c->d_nnz;
  if (rp) {
    std::vector<int> tmp(n + 1);
    HIP_TRY(hipMemcpy(tmp.data(), c->d_rp.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    for (size_t i = 0; i <= n; ++i) rp[i] = tmp[i];
  }
  if (col && m) {
    std::vector<uint32_t> tmp(m);
    HIP_TRY(hipMemcpy(tmp.data(), c->d_col.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < m; ++k) col[k] = (int32_t)(tmp[k] & 0x7FFFFFFFu);
  }
  if (h && m) HIP_TRY(hipMemcpy(h, c->d_h64.p, sizeof(double) * m, hipMemcpyDeviceToHost));
  if (cdf && m) HIP_TRY(hipMemcpy(cdf, c->d_cdf64.p, sizeof(double) * m, hipMemcpyDeviceToHost));
  if (W) HIP_TRY(hipMemcpy(W, c->d_W64.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return MCLS_OK;
}

int mcls_gpu_prec_residual(mcls_gpu_context* c, const double* b, const double* x, double* r, double* norm2,
                           int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (!b || !x) return fail(MCLS_ERR_INVALID_ARG, "null vector");
  MCLS_TRY(vec_in(c, b, ptr_kind, c->b.p, 1, sizeof(double)));
  MCLS_TRY(vec_in(c, x, ptr_kind, c->x.p, 1, sizeof(double)));
  MCLS_TRY(residual(c, c->b.p, c->x.p, c->r.p, 0));
  if (r) MCLS_TRY(vec_out(c, c->r.p, r, ptr_kind, 1, sizeof(double)));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (norm2) *norm2 = std::sqrt(c->host_scal[0]);
  return MCLS_OK;
}

int mcls_gpu_mc_solve_raw(mcls_gpu_context* c, const double* rhs, uint64_t solve_index, double* x) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_matrix(c));
  if (!rhs || !x) return fail(MCLS_ERR_INVALID_ARG, "null vector");
  MCLS_TRY(ensure_vectors(c));
  MCLS_TRY(ensure_domain(c));
  c->stats.walk_launches = 0;
  c->stats.ms_walk = c->stats.ms_source = 0.0;
  HIP_TRY(mcls::launch_fill_u64(c->steps.p, 1, 0ULL, c->stream));
  MCLS_TRY(vec_in(c, rhs, MCLS_HOST, c->rhs.p, 1, sizeof(double)));
  MCLS_TRY(check_finite_source(c, c->rhs.p));
  MCLS_TRY(mc_solve_device(c, c->rhs.p, c->x.p, false, solve_index));
  MCLS_TRY(vec_out(c, c->x.p, x, MCLS_HOST, 1, sizeof(double)));
  unsigned long long steps = 0;
  HIP_TRY(hipMemcpyAsync(&steps, c->steps.p, sizeof(steps), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  MCLS_TRY(accumulate_mc_timing(c));
  c->stats.steps = (int64_t)steps;
  MCLS_TRY(histories_done(c, &c->stats.histories));
  return MCLS_OK;
}

int mcls_gpu_mc_solve_window(mcls_gpu_context* c, const double* rhs, uint64_t solve_index, int64_t first,
                             int64_t count, double* x) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_matrix(c));
  if (!rhs || !x) return fail(MCLS_ERR_INVALID_ARG, "null vector");
  if (first < 0 || count < 0) return fail(MCLS_ERR_INVALID_ARG, "negative history window");
  MCLS_TRY(ensure_vectors(c));
  MCLS_TRY(ensure_domain(c));
  c->stats.walk_launches = 0;
  c->stats.ms_walk = c->stats.ms_source = 0.0;
  HIP_TRY(mcls::launch_fill_u64(c->steps.p, 1, 0ULL, c->stream));
  MCLS_TRY(vec_in(c, rhs, MCLS_HOST, c->rhs.p, 1, sizeof(double)));
  MCLS_TRY(check_finite_source(c, c->rhs.p));
  MCLS_TRY(mc_solve_device(c, ctx_bufs(c), c->rhs.p, c->x.p, false, solve_index, true, nullptr, 1.0, first, count));
  MCLS_TRY(vec_out(c, c->x.p, x, MCLS_HOST, 1, sizeof(double)));
  unsigned long long steps = 0;
  HIP_TRY(hipMemcpyAsync(&steps, c->steps.p, sizeof(steps), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  MCLS_TRY(accumulate_mc_timing(c));
  c->stats.steps = (int64_t)steps;
  return MCLS_OK;
}

int mcls_gpu_record_paths(mcls_gpu_context* c, const double* rhs, uint64_t solve_index, int64_t first,
                          int64_t count, int32_t max_len, int32_t* paths, double* weights) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_matrix(c));
  if (!rhs || !paths || !weights || count <= 0 || first < 0 || max_len < 0)
    return fail(MCLS_ERR_INVALID_ARG, "bad record_paths arguments");
  if (count > (int64_t)1 << 24 || (int64_t)count * (max_len + 1) > ((int64_t)1 << 28))
    return fail(MCLS_ERR_INVALID_ARG, "record_paths request too large");
  MCLS_TRY(ensure_vectors(c));
  MCLS_TRY(ensure_domain(c));
  hipStream_t s = c->stream;
  DBuf<int> dpaths;
  DBuf<double> dw;
  MCLS_TRY(dpaths.ensure((size_t)count * (size_t)(max_len + 1), "paths"));
  MCLS_TRY(dw.ensure((size_t)count, "weights"));
  HIP_TRY(mcls::launch_fill_i32(dpaths.p, count * (int64_t)(max_len + 1), -1, s));
  HIP_TRY(mcls::launch_fill_f64(dw.p, count, 0.0, s));
  HIP_TRY(mcls::launch_fill_u64(c->steps.p, 1, 0ULL, s));
  MCLS_TRY(vec_in(c, rhs, MCLS_HOST, c->rhs.p, 1, sizeof(double)));
  const int64_t nh = num_histories(c);
  if (!c->prm.forward) MCLS_TRY(build_source(c, c->rhs.p, nh));
  else if (first + count > nh) return fail(MCLS_ERR_INVALID_ARG, "history range beyond the forward source");
  if (c->prm.sampler == mcls::SAMPLER_ALIAS_STRIDED && !c->strided_valid)
    return fail(MCLS_ERR_STATE, "strided alias records not built");
  WalkArgs a = walk_args(c, ctx_bufs(c), c->rhs.p, solve_key(c->prm.seed, solve_index), first, first + count);
  MCLS_TRY(prepare_birth_order(c, ctx_bufs(c), first, first + count, a));
  a.paths = dpaths.p;
  a.wfinal = dw.p;
  a.max_len = max_len;
  WalkConfig wc = walk_config(c, true, count);
  HIP_TRY(mcls::launch_walk(a, wc, s));
  // clear the tally the recorded walks wrote
  HIP_TRY(mcls::launch_store_tally(c->n, c->x.p, c->t64.p, nullptr, c->rhs.p, false, 0.0, nullptr, s));
  HIP_TRY(hipMemcpyAsync(paths, dpaths.p, sizeof(int) * (size_t)count * (size_t)(max_len + 1),
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(weights, dw.p, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (c->ordered) {  // states in the caller's numbering
    const size_t m = (size_t)count * (size_t)(max_len + 1);
    for (size_t i = 0; i < m; ++i)
      if (paths[i] >= 0) paths[i] = c->ord_h[(size_t)paths[i]];
  }
  return MCLS_OK;
}

int mcls_gpu_mc_solve(mcls_gpu_context* c, const double* b, double* x, int ptr_kind) {
  auto t0 = std::chrono::steady_clock::now();
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (c->nsys > 0 && general_prec(c))
    return fail(MCLS_ERR_UNSUPPORTED, "sparse preconditioners are not supported with batched systems");
  if (!b || !x) return fail(MCLS_ERR_INVALID_ARG, "null vector");
  MCLS_TRY(ensure_vectors(c));
  MCLS_TRY(ensure_domain(c));
  hipStream_t s = c->stream;
  c->stats.walk_launches = 0;
  c->stats.ms_walk = c->stats.ms_source = c->stats.ms_residual = 0.0;
  HIP_TRY(mcls::launch_fill_u64(c->steps.p, 1, 0ULL, s));
  MCLS_TRY(vec_in(c, b, ptr_kind, c->b.p, 1, sizeof(double)));
  // buildMonteCarloSource (src/MCLS_MonteCarloSolverManager_impl.hpp:347-380):
  // rhs = omega * PL b
  MCLS_TRY(apply_left_prec(c, c->b.p, c->rhs.p));
  if (c->prm.neumann_omega != 1.0) HIP_TRY(mcls::launch_scale_copy(c->n, c->prm.neumann_omega, c->rhs.p, c->rhs.p, s));
  MCLS_TRY(check_finite_source(c, c->rhs.p));
  const uint64_t sidx = c->solve_index++;
  MCLS_TRY(mc_solve_device(c, c->rhs.p, c->x.p, false, sidx));
  // right preconditioning: x = PR x (src/MCLS_MonteCarloSolverManager_impl.hpp:220-229)
  MCLS_TRY(recover_right_prec(c, c->x.p));
  MCLS_TRY(vec_out(c, c->x.p, x, ptr_kind, 1, sizeof(double)));
  unsigned long long steps = 0;
  HIP_TRY(hipMemcpyAsync(&steps, c->steps.p, sizeof(steps), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  MCLS_TRY(accumulate_mc_timing(c));
  c->stats.steps = (int64_t)steps;
  MCLS_TRY(histories_done(c, &c->stats.histories));
  c->stats.iterations = 0;
  c->stats.converged = 1;
  c->stats.solve_index = (int64_t)c->solve_index;
  c->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return MCLS_OK;
}

int mcls_gpu_mc_solve_noalloc(mcls_gpu_context* c, const double* b, double* x) {
  auto t0 = std::chrono::steady_clock::now();
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_matrix(c));
  if (!b || !x) return fail(MCLS_ERR_INVALID_ARG, "null vector");
  if ((int64_t)c->n > (int64_t)MCLS_NOALLOC_MAX_N)
    return fail(MCLS_ERR_INVALID_ARG, "mcls_gpu_mc_solve_noalloc: n = " + std::to_string(c->n) +
                                          " exceeds its scratch limit of " + std::to_string(MCLS_NOALLOC_MAX_N) +
                                          " unknowns");
  if (c->nsys > 0) return fail(MCLS_ERR_UNSUPPORTED, "mcls_gpu_mc_solve_noalloc: not with batched systems");
  if (general_prec(c))
    return fail(MCLS_ERR_UNSUPPORTED, "mcls_gpu_mc_solve_noalloc: not with sparse preconditioners");
  if (c->model)
    return fail(MCLS_ERR_UNSUPPORTED, "mcls_gpu_mc_solve_noalloc: not with an importance network attached");
  if (!tables_ready(c))
    return fail(MCLS_ERR_STATE, "mcls_gpu_mc_solve_noalloc: the domain or the Transition Sampler's tables are not "
                                "built for the current operator and parameters (call mcls_gpu_build_domain)");
  if (!c->scratch) return fail(MCLS_ERR_STATE, "scratch array not resolved on this device");
  hipStream_t s = c->stream;
  const size_t n = (size_t)c->n;
  // carve the per-call buffers out of the scratch array (256-byte aligned)
  unsigned char* base = (unsigned char*)c->scratch;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  SolveBufs sb{};
  sb.may_allocate = false;  // nothing below may allocate (ordered births walk in id order)
  sb.queue = (unsigned long long*)take(mcls::kQueueWords * sizeof(unsigned long long));
  sb.steps = (unsigned long long*)take(sizeof(unsigned long long));
  sb.bmax = (unsigned long long*)take(sizeof(unsigned long long));
  sb.scan_tmp = (uint64_t*)take((mcls::scan_tmp_elems((int64_t)n) + 1) * sizeof(uint64_t));
  if (off > kNoAllocFixedBytes) return fail(MCLS_ERR_INVALID_ARG, "scratch layout exceeds its fixed part");
  off = kNoAllocFixedBytes;
  double* rhs = (double*)take(n * sizeof(double));
  sb.q = (uint64_t*)take(n * sizeof(uint64_t));
  sb.qcdf = (uint64_t*)take(n * sizeof(uint64_t));
  sb.ncnt = (uint64_t*)take(n * sizeof(uint64_t));
  sb.ncdf = (uint64_t*)take(n * sizeof(uint64_t));
  sb.t64 = (double*)take(n * sizeof(double));
  double* b_int = c->ordered ? (double*)take(n * sizeof(double)) : nullptr;
  double* x_int = c->ordered ? (double*)take(n * sizeof(double)) : nullptr;
  if (off > kNoAllocBytes) return fail(MCLS_ERR_INVALID_ARG, "scratch layout exceeds the scratch array");

  c->stats.walk_launches = 0;
  c->stats.ms_walk = c->stats.ms_source = c->stats.ms_residual = 0.0;
  // the array keeps the last call's contents: clear what must start at zero
  // (fill kernels, so that the call can be captured in a graph)
  HIP_TRY(mcls::launch_fill_u64(sb.steps, 1, 0ULL, s));
  HIP_TRY(mcls::launch_fill_f64(sb.t64, (int64_t)n, 0.0, s));
  const double* bb = b;
  if (c->ordered) {
    HIP_TRY(mcls::launch_gather_rows((int64_t)n, 1, sizeof(double), c->ord.p, b, b_int, s));
    bb = b_int;
  }
  // buildMonteCarloSource (src/MCLS_MonteCarloSolverManager_impl.hpp:347-380): rhs = omega PL b
  if (c->has_prec)
    HIP_TRY(mcls::launch_diag_scale(c->n, c->dinv.p, bb, rhs, s));
  else
    HIP_TRY(mcls::launch_scale_copy(c->n, 1.0, bb, rhs, s));
  if (c->prm.neumann_omega != 1.0) HIP_TRY(mcls::launch_scale_copy(c->n, c->prm.neumann_omega, rhs, rhs, s));
  const uint64_t sidx = c->solve_index++;
  MCLS_TRY(mc_solve_device(c, sb, rhs, c->ordered ? x_int : x, false, sidx));
  if (c->ordered) HIP_TRY(mcls::launch_scatter_rows((int64_t)n, 1, sizeof(double), c->ord.p, x_int, x, s));
  unsigned long long steps = 0;
  HIP_TRY(hipMemcpyAsync(&steps, sb.steps, sizeof(steps), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  MCLS_TRY(accumulate_mc_timing(c));
  c->stats.steps = (int64_t)steps;
  MCLS_TRY(histories_done(c, sb, &c->stats.histories));
  c->stats.iterations = 0;
  c->stats.converged = 1;
  c->stats.solve_index = (int64_t)c->solve_index;
  c->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return MCLS_OK;
}

int mcls_gpu_mcsa_solve(mcls_gpu_context* c, const double* b, double* x, int ptr_kind) {
  auto t0 = std::chrono::steady_clock::now();
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (c->nsys > 0 && general_prec(c))
    return fail(MCLS_ERR_UNSUPPORTED, "sparse preconditioners are not supported with batched systems");
  if (!b || !x) return fail(MCLS_ERR_INVALID_ARG, "null vector");
  MCLS_TRY(ensure_vectors(c));
  MCLS_TRY(ensure_domain(c));
  hipStream_t s = c->stream;
  const int n = c->n;
  const Params& p = c->prm;
  mcls_gpu_stats& st = c->stats;
  st.walk_launches = 0;
  st.ms_walk = st.ms_source = st.ms_residual = 0.0;
  st.histories = st.steps = 0;
  HIP_TRY(mcls::launch_fill_u64(c->steps.p, 1, 0ULL, s));
  MCLS_TRY(vec_in(c, b, ptr_kind, c->b.p, 1, sizeof(double)));
  MCLS_TRY(vec_in(c, x, ptr_kind, c->x.p, 1, sizeof(double)));

  // src/MCLS_MCSASolverManager_impl.hpp:300-318: ||PL b||_2, homogeneous case
  if (c->pl_csr.on) {
    MCLS_TRY(apply_left_prec(c, c->b.p, c->pt2.p));
    HIP_TRY(mcls::launch_scaled_sq(n, nullptr, c->pt2.p, c->part.p, s));
  } else {
    HIP_TRY(mcls::launch_scaled_sq(n, c->has_prec ? c->dinv.p : nullptr, c->b.p, c->part.p, s));
  }
  HIP_TRY(mcls::launch_sum_partials(c->part.p, mcls::residual_blocks(n, 1), c->scal.p + 2, s));
  HIP_TRY(hipMemcpyAsync(c->host_scal + 2, c->scal.p + 2, sizeof(double), hipMemcpyDeviceToHost, s));
  MCLS_TRY(residual(c, c->b.p, c->x.p, c->r.p, 0));
  HIP_TRY(hipStreamSynchronize(s));
  double source_norm = std::sqrt(c->host_scal[2]);
  if (std::fabs(source_norm) < 10.0 * DBL_EPSILON) source_norm = 1.0;
  const double crit = p.tol * source_norm;
  double rnorm = std::sqrt(c->host_scal[0]);
  c->res_hist.assign(1, rnorm / source_norm);
  if (!std::isfinite(rnorm)) return fail(MCLS_ERR_NUMERIC, "non-finite initial residual");
  if (p.verbosity > 0) {
    std::printf("MCLS MCSA (gfx950): N=%d domain nnz=%lld histories/iter=%lld\n", n, (long long)c->d_nnz,
                (long long)num_histories(c));
  }
  int64_t iters = 0;
  bool go = true;
  const int64_t nh = num_histories(c);
  const bool batch = c->nsys > 0;
  if (batch) {
    // per-system ||PL b_s||^2 and convergence flags
    HIP_TRY(mcls::launch_segment_norm(c->nsys, c->sys_off.p, c->b.p, c->has_prec ? c->dinv.p : nullptr, true,
                                      c->sys_b2.p, s));
    HIP_TRY(mcls::launch_fill_i32(c->sys_conv.p, c->nsys, 0, s));
  }
  // hipGraph of one iteration (build option "Use Graph"): captured once per
  // problem and parameter set, replayed every iteration.
  const bool graph = p.use_graph != 0;
  if (graph) {
    unsigned long long si = c->solve_index;
    HIP_TRY(hipMemcpyAsync(c->sidx_dev.p, &si, sizeof(si), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!c->graph_exec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc = mcsa_iteration(c, false, c->key_dev.p, 0);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (rc != MCLS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      HIP_TRY(e);
      c->graph = g;
      HIP_TRY(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
      st.walk_launches = 0;  // the capture itself launched nothing
    }
  }
  while (go) {
    ++iters;
    const uint64_t sidx = c->solve_index++;
    if (graph) {
      HIP_TRY(hipEventRecord(c->ev[5], s));
      HIP_TRY(hipGraphLaunch(c->graph_exec, s));
      HIP_TRY(hipEventRecord(c->ev[7], s));
      HIP_TRY(hipStreamSynchronize(s));
      float a = 0.f;
      HIP_TRY(hipEventElapsedTime(&a, c->ev[5], c->ev[7]));
      st.ms_walk += a;  // whole iteration (phases are not separated in a graph)
      st.walk_launches += nh > 0 ? 1 : 0;
    } else {
      MCLS_TRY(mcsa_iteration(c, true, nullptr, sidx));
      HIP_TRY(hipStreamSynchronize(s));
      MCLS_TRY(accumulate_mc_timing(c));
      float a = 0.f, b2 = 0.f;
      HIP_TRY(hipEventElapsedTime(&a, c->ev[5], c->ev[6]));
      HIP_TRY(hipEventElapsedTime(&b2, c->ev[4], c->ev[7]));
      st.ms_residual += a + b2;
    }
    int64_t hd = nh;  // the forward source's count is nh exactly
    if (p.source == mcls::SRC_STRATIFIED && !p.forward) {
      uint64_t M = 0;
      std::memcpy(&M, c->host_scal + kHostM, sizeof(M));
      hd = (int64_t)M;
    }
    st.histories += hd;
    rnorm = std::sqrt(c->host_scal[0]);
    c->res_hist.push_back(rnorm / source_norm);
    if (!std::isfinite(rnorm)) {
      st.iterations = iters;
      st.converged = 0;
      return fail(MCLS_ERR_NUMERIC, "non-finite residual at MCSA iteration " + std::to_string(iters));
    }
    if (iters % p.check_freq == 0) {
      if (batch) {
        int unconv = 0;
        std::memcpy(&unconv, c->host_scal + kHostUnconv, sizeof(int));
        go = unconv > 0 && iters < p.max_iters;
      } else {
        go = (rnorm > crit) && (iters < p.max_iters);
      }
    }
    if (p.verbosity > 0 && ((p.print_freq > 0 && iters % p.print_freq == 0) || !go))
      std::printf("%18lld%18.4e\n", (long long)iters, rnorm / source_norm);
  }
  // src/MCLS_MCSASolverManager_impl.hpp:410-420: x = PR x when right preconditioned
  MCLS_TRY(recover_right_prec(c, c->x.p));
  MCLS_TRY(vec_out(c, c->x.p, x, ptr_kind, 1, sizeof(double)));
  unsigned long long steps = 0;
  HIP_TRY(hipMemcpyAsync(&steps, c->steps.p, sizeof(steps), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  st.steps = (int64_t)steps;
  st.iterations = iters;
  st.converged = rnorm <= crit;
  st.achieved_tol = rnorm / source_norm;
  st.source_norm = source_norm;
  if (batch) {
    // per-system relative residuals; the batch converges when all do
    std::vector<double> r2((size_t)c->nsys), b2((size_t)c->nsys);
    HIP_TRY(hipMemcpy(r2.data(), c->sys_r2.p, sizeof(double) * r2.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(b2.data(), c->sys_b2.p, sizeof(double) * b2.size(), hipMemcpyDeviceToHost));
    c->sys_rel.assign((size_t)c->nsys, 0.0);
    double worst = 0.0;
    for (int k = 0; k < c->nsys; ++k) {
      double bn = std::sqrt(b2[(size_t)k]);
      if (std::fabs(bn) < 10.0 * DBL_EPSILON) bn = 1.0;
      c->sys_rel[(size_t)k] = std::sqrt(r2[(size_t)k]) / bn;
      worst = std::max(worst, c->sys_rel[(size_t)k]);
    }
    st.achieved_tol = worst;
    st.converged = worst <= p.tol;
  }
  st.solve_index = (int64_t)c->solve_index;
  st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return MCLS_OK;
}

int mcls_gpu_get_residual_history(mcls_gpu_context* c, double* hist, int64_t cap, int64_t* count) {
  MCLS_TRY(check_ctx(c));
  if (!count || (cap > 0 && !hist)) return fail(MCLS_ERR_INVALID_ARG, "null output");
  const int64_t m = (int64_t)c->res_hist.size();
  for (int64_t k = 0; k < m && k < cap; ++k) hist[k] = c->res_hist[(size_t)k];
  *count = m;
  return MCLS_OK;
}

int mcls_gpu_set_importance(mcls_gpu_context* c, const float* g, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (!g) return fail(MCLS_ERR_INVALID_ARG, "null importance");
  MCLS_TRY(c->g_imp.ensure((size_t)c->n, "importance"));
  MCLS_TRY(vec_in(c, g, ptr_kind, c->g_imp.p, 1, sizeof(float)));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->has_importance = true;
  c->proposal_valid = false;
  invalidate_graph(c);
  if (c->domain_valid && learned_sampler(c)) return ensure_proposal(c);
  return MCLS_OK;
}

int mcls_gpu_convergence_criteria(mcls_gpu_context* c, int64_t max_iters, double tol, mcls_gpu_criteria* out) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_matrix(c));
  if (!out) return fail(MCLS_ERR_INVALID_ARG, "null output");
  if (max_iters < 64 || !(tol > 0.0)) return fail(MCLS_ERR_INVALID_ARG, "max_iters must be >= 64 and tol > 0");
  MCLS_TRY(ensure_domain(c));
  *out = mcls_gpu_criteria{};
  return convergence_criteria(c, max_iters, tol, out);
}

int mcls_gpu_autotune(mcls_gpu_context* c, const double* b, int ptr_kind, int64_t pilot, mcls_gpu_tuning* out) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (!b || !out) return fail(MCLS_ERR_INVALID_ARG, "null vector or output");
  if (pilot < 0) return fail(MCLS_ERR_INVALID_ARG, "pilot history count must be >= 0");
  MCLS_TRY(ensure_vectors(c));
  MCLS_TRY(ensure_domain(c));
  hipStream_t s = c->stream;
  MCLS_TRY(vec_in(c, b, ptr_kind, c->b.p, 1, sizeof(double)));
  HIP_TRY(mcls::launch_fill_f64(c->x.p, c->n, 0.0, s));
  MCLS_TRY(residual(c, c->b.p, c->x.p, c->r.p, 0));  // pilot source PL b
  const Params saved = c->prm;
  const int64_t saved_nsys = c->nsys;
  const int64_t nh = num_histories(c);
  int64_t np = pilot > 0 ? pilot : std::max<int64_t>(nh / 8, (int64_t)1 << 20);
  if (nh > 0 && np > nh) np = nh;
  // a pilot walks np consecutive histories of the full source, so that the
  // births per state (what the LDS tile and the caches see) are the solve's
  const double window = nh > 0 ? (double)np / (double)nh : 1.0;
  c->nsys = 0;  // pilot births follow |PL b| over the whole operator
  auto pilot_win = [&](int64_t blk, int64_t bpc, int64_t bat, int64_t tile, double win, double* ms) -> int {
    c->prm.walk_block = blk;
    c->prm.walk_blocks_per_cu = bpc;
    c->prm.walk_batch = bat;
    c->prm.tally_tile = tile;
    MCLS_TRY(ensure_vectors(c));
    // the faster of two pilots: shapes within a few percent of each other
    // are otherwise ranked by timing noise
    double fastest = 0.0;
    for (int rep = 0; rep < 2; ++rep) {
      MCLS_TRY(mc_solve_device(c, c->r.p, c->x.p, false, 0, true, nullptr, win));
      HIP_TRY(hipStreamSynchronize(s));
      float t = 0.f;
      HIP_TRY(hipEventElapsedTime(&t, c->ev[3], c->ev[4]));
      if (rep == 0 || t < fastest) fastest = t;
    }
    *ms = fastest;
    return MCLS_OK;
  };
  auto pilot_ms = [&](int64_t blk, int64_t bpc, int64_t bat, int64_t tile, double* ms) -> int {
    return pilot_win(blk, bpc, bat, tile, window, ms);
  };
  // a tile of `tile` entries per wave fits the 64 KiB of LDS of a block
  const bool tiled_est =
      !saved.forward && (saved.estimator == mcls::EST_COLLISION || saved.estimator == mcls::EST_EXPECTED_VALUE);
  const size_t tile_elem = mcls::tile_entry_bytes(tile_kind_of(saved), saved.precision == 32 ? 4 : 8);
  auto tile_fits = [&](int64_t blk, int64_t tile) {
    return !tiled_est || tile <= 0 || (size_t)(blk / 64) * (size_t)tile * tile_elem <= 65536;
  };
  int rc = MCLS_OK;
  mcls_gpu_tuning best{};
  best.block = saved.walk_block;
  best.blocks_per_cu = saved.walk_blocks_per_cu > 0 ? saved.walk_blocks_per_cu : 8;
  best.batch = saved.walk_batch;
  best.tally_tile = saved.tally_tile;
  double ms = 0.0;
  // warm-up, then the configuration on entry; an entry tile over the LDS
  // limit is not timed (ms_default = -1) and the search starts untiled
  best.candidates = 0;
  if (tile_fits(best.block, best.tally_tile)) {
    rc = pilot_ms(best.block, best.blocks_per_cu, best.batch, best.tally_tile, &ms);
    if (rc == MCLS_OK) rc = pilot_ms(best.block, best.blocks_per_cu, best.batch, best.tally_tile, &ms);
    best.ms_default = ms;
    best.ms_best = ms;
    best.candidates = 1;
  } else {
    best.tally_tile = 0;
    best.ms_default = -1.0;
    best.ms_best = -1.0;
  }
  // launch shape: 8, 16 or 32 resident waves per CU at each block size
  const int64_t blocks[3] = {64, 128, 256};
  const int64_t waves[3] = {8, 16, 32};
  double block_ms[3] = {0.0, 0.0, 0.0};
  int64_t block_bpc[3] = {0, 0, 0};  // the best resident count of each block size
  for (int i = 0; i < 3 && rc == MCLS_OK; ++i)
    for (int j = 0; j < 3 && rc == MCLS_OK; ++j) {
      const int64_t bpc = waves[j] * 64 / blocks[i];
      if (!tile_fits(blocks[i], best.tally_tile)) continue;
      rc = pilot_ms(blocks[i], bpc, best.batch, best.tally_tile, &ms);
      ++best.candidates;
      if (rc == MCLS_OK && (block_bpc[i] == 0 || ms < block_ms[i])) {
        block_ms[i] = ms;
        block_bpc[i] = bpc;
      }
      if (rc == MCLS_OK && (best.ms_best < 0.0 || ms < best.ms_best)) {
        best.ms_best = ms;
        best.block = blocks[i];
        best.blocks_per_cu = bpc;
      }
    }
  // history ids per queue grab at the chosen shape
  const int64_t batches[4] = {64, 128, 256, 512};
  const int64_t shape_batch = best.batch;
  for (int k = 0; k < 4 && rc == MCLS_OK; ++k) {
    if (batches[k] == shape_batch) continue;
    rc = pilot_ms(best.block, best.blocks_per_cu, batches[k], best.tally_tile, &ms);
    ++best.candidates;
    if (rc == MCLS_OK && ms < best.ms_best) {
      best.ms_best = ms;
      best.batch = batches[k];
    }
  }
  // LDS tally tile (adjoint collision and expected-value estimators): off, 128, 512 or 1024
  // entries per wave, within 64 KiB of LDS per workgroup
  if (tiled_est) {
    const int64_t tiles[4] = {0, 128, 512, 1024};
    const int64_t shape_tile = best.tally_tile;
    for (int k = 0; k < 4 && rc == MCLS_OK; ++k) {
      if (tiles[k] == shape_tile || !tile_fits(best.block, tiles[k])) continue;
      rc = pilot_ms(best.block, best.blocks_per_cu, best.batch, tiles[k], &ms);
      ++best.candidates;
      if (rc == MCLS_OK && ms < best.ms_best) {
        best.ms_best = ms;
        best.tally_tile = tiles[k];
      }
    }
  }
  // The block size, confirmed over the whole source: pilots over a window
  // of it have ranked 256-thread blocks ahead of smaller ones that walk the
  // whole source faster.  The best shape of each block size, with the chosen
  // batch and tile, walks the full source; the fastest is kept.
  if (rc == MCLS_OK) {
    double full_best = -1.0;
    int64_t fb = best.block, fbpc = best.blocks_per_cu;
    for (int i = 0; i < 3 && rc == MCLS_OK; ++i) {
      if (block_bpc[i] == 0 || !tile_fits(blocks[i], best.tally_tile)) continue;
      rc = pilot_win(blocks[i], block_bpc[i], best.batch, best.tally_tile, 1.0, &ms);
      ++best.candidates;
      if (rc == MCLS_OK && (full_best < 0.0 || ms < full_best)) {
        full_best = ms;
        fb = blocks[i];
        fbpc = block_bpc[i];
      }
    }
    // The batch, confirmed the same way at that shape: the window holds an
    // eighth of the solve's grabs, and has ranked 256 ids per grab ahead of
    // 512 that walk the whole source faster.  The chosen batch's neighbours
    // walk the full source too; the fastest is kept.
    const int64_t fbat0 = best.batch;
    int64_t fbat = fbat0;
    if (full_best >= 0.0) {
      for (int k = 0; k < 4 && rc == MCLS_OK; ++k) {
        if (batches[k] != fbat0 / 2 && batches[k] != fbat0 * 2) continue;
        rc = pilot_win(fb, fbpc, batches[k], best.tally_tile, 1.0, &ms);
        ++best.candidates;
        if (rc == MCLS_OK && ms < full_best) {
          full_best = ms;
          fbat = batches[k];
        }
      }
    }
    // ms_best is the window pilot of the configuration returned
    if (rc == MCLS_OK && (fb != best.block || fbpc != best.blocks_per_cu || fbat != fbat0)) {
      rc = pilot_ms(fb, fbpc, fbat, best.tally_tile, &ms);
      ++best.candidates;
      best.ms_best = ms;
    }
    best.block = fb;
    best.blocks_per_cu = fbpc;
    best.batch = fbat;
    best.ms_confirmed = full_best;
  }
  c->prm = saved;
  c->nsys = saved_nsys;
  MCLS_TRY(rc);
  c->prm.walk_block = best.block;
  c->prm.walk_blocks_per_cu = best.blocks_per_cu;
  c->prm.walk_batch = best.batch;
  c->prm.tally_tile = best.tally_tile;
  MCLS_TRY(ensure_vectors(c));
  invalidate_graph(c);
  *out = best;
  return MCLS_OK;
}

int mcls_gpu_set_importance_diagonal(mcls_gpu_context* c) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_matrix(c));
  MCLS_TRY(c->g_imp.ensure((size_t)c->n, "importance"));
  HIP_TRY(mcls::launch_diag_abs(c->n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, c->g_imp.p, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->has_importance = true;
  c->proposal_valid = false;
  invalidate_graph(c);
  if (c->domain_valid && learned_sampler(c)) return ensure_proposal(c);
  return MCLS_OK;
}

int mcls_gpu_get_importance(mcls_gpu_context* c, float* g, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  if (!g) return fail(MCLS_ERR_INVALID_ARG, "null output");
  if (!c->g_imp.p) return fail(MCLS_ERR_STATE, "no importance has been set or computed");
  HIP_TRY(hipStreamSynchronize(c->stream));
  MCLS_TRY(vec_out(c, c->g_imp.p, g, ptr_kind, 1, sizeof(float)));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCLS_OK;
}

int mcls_gpu_set_importance_model(mcls_gpu_context* c, mcls_mlp* m) {
  MCLS_TRY(check_ctx(c));
  if (m && mcls::mlp_device(m) != c->device)
    return fail(MCLS_ERR_INVALID_ARG, "network and context are on different devices");
  c->model = m;
  c->proposal_valid = false;
  invalidate_graph(c);
  return MCLS_OK;
}

int mcls_gpu_set_importance_mask(mcls_gpu_context* c, const signed char* mask, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  invalidate_graph(c);
  if (!mask) {
    c->has_mask = false;
    return MCLS_OK;
  }
  MCLS_TRY(c->imp_mask.ensure((size_t)c->n, "importance mask"));
  MCLS_TRY(vec_in(c, mask, ptr_kind, c->imp_mask.p, 1, 1));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->has_mask = true;
  return MCLS_OK;
}

int mcls_gpu_importance_features(mcls_gpu_context* c, const double* v, float* X, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (!v || !X) return fail(MCLS_ERR_INVALID_ARG, "null vector");
  MCLS_TRY(ensure_vectors(c));
  MCLS_TRY(vec_in(c, v, ptr_kind, c->rhs.p, 1, sizeof(double)));
  MCLS_TRY(compute_features(c, c->rhs.p));
  MCLS_TRY(vec_out(c, c->feat.p, X, ptr_kind, 8, sizeof(float)));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCLS_OK;
}

int mcls_gpu_importance_targets(mcls_gpu_context* c, const double* v, uint64_t solve_index, float* T, int ptr_kind) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_kind(ptr_kind));
  MCLS_TRY(check_matrix(c));
  if (!v || !T) return fail(MCLS_ERR_INVALID_ARG, "null vector");
  MCLS_TRY(ensure_vectors(c));
  // one plain adjoint MC solve of v (almost-optimal transitions)
  const int sampler = c->prm.sampler;
  if (sampler == mcls::SAMPLER_LEARNED || sampler == mcls::SAMPLER_LEARNED_ALIAS) c->prm.sampler = mcls::SAMPLER_CDF;
  int rc = ensure_domain(c);
  if (rc == MCLS_OK) {
    rc = vec_in(c, v, ptr_kind, c->rhs.p, 1, sizeof(double));
    if (rc == MCLS_OK) rc = mc_solve_device(c, c->rhs.p, c->x.p, false, solve_index);
  }
  c->prm.sampler = sampler;
  MCLS_TRY(rc);
  MCLS_TRY(c->mlp_out.ensure((size_t)c->n, "targets"));
  // y = diag(A) x: for symmetric A under point Jacobi the adjoint tally
  // approximates the right Perron vector v of |H|, and D v is the left one,
  // the importance that keeps walk weights from growing
  HIP_TRY(mcls::launch_scale_by_diag(c->n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, c->x.p, c->r.p, c->stream));
  HIP_TRY(mcls::launch_feature_sums(c->n, c->a_rp.p, c->a_col.p, c->a_val.p, c->f32, c->r.p, c->part.p,
                                    c->scal.p + 4, c->stream));
  HIP_TRY(mcls::launch_targets(c->n, c->r.p, c->scal.p + 4, c->mlp_out.p, c->stream));
  MCLS_TRY(vec_out(c, c->mlp_out.p, T, ptr_kind, 1, sizeof(float)));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCLS_OK;
}

int mcls_gpu_set_systems(mcls_gpu_context* c, int64_t nsys, const int64_t* offsets) {
  MCLS_TRY(check_ctx(c));
  MCLS_TRY(check_matrix(c));
  invalidate_graph(c);
  if (nsys == 0) {
    c->nsys = 0;
    return MCLS_OK;
  }
  if (nsys < 0 || nsys > c->n || !offsets) return fail(MCLS_ERR_INVALID_ARG, "bad system count or offsets");
  if (c->ordered) return fail(MCLS_ERR_UNSUPPORTED, "batched systems with a state order");
  if (offsets[0] != 0 || offsets[nsys] != c->n) return fail(MCLS_ERR_INVALID_ARG, "offsets must span [0, n]");
  std::vector<int> sys_of((size_t)c->n);
  for (int64_t s = 0; s < nsys; ++s) {
    if (offsets[s + 1] <= offsets[s]) return fail(MCLS_ERR_INVALID_ARG, "system offsets must increase");
    for (int64_t i = offsets[s]; i < offsets[s + 1]; ++i) sys_of[(size_t)i] = (int)s;
  }
  MCLS_TRY(c->sys_off.ensure((size_t)nsys + 1, "system offsets"));
  MCLS_TRY(c->sys_of.ensure((size_t)c->n, "system of row"));
  MCLS_TRY(c->sys_conv.ensure((size_t)nsys, "system flags"));
  MCLS_TRY(c->sys_unconv.ensure(1, "unconverged count"));
  MCLS_TRY(c->sys_b2.ensure((size_t)nsys, "system source norms"));
  MCLS_TRY(c->sys_r2.ensure((size_t)nsys, "system residual norms"));
  MCLS_TRY(c->sys_l1.ensure((size_t)nsys, "system 1-norms"));
  HIP_TRY(hipMemcpy(c->sys_off.p, offsets, sizeof(int64_t) * (size_t)(nsys + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->sys_of.p, sys_of.data(), sizeof(int) * sys_of.size(), hipMemcpyHostToDevice));
  c->sys_off_h.assign(offsets, offsets + nsys + 1);
  c->nsys = (int)nsys;
  return MCLS_OK;
}

int mcls_gpu_get_system_residuals(mcls_gpu_context* c, double* rel) {
  MCLS_TRY(check_ctx(c));
  if (!rel) return fail(MCLS_ERR_INVALID_ARG, "null output");
  if (c->nsys == 0 || c->sys_rel.size() != (size_t)c->nsys)
    return fail(MCLS_ERR_STATE, "no batched solve has run");
  std::memcpy(rel, c->sys_rel.data(), sizeof(double) * c->sys_rel.size());
  return MCLS_OK;
}

int mcls_gpu_get_stats(mcls_gpu_context* c, mcls_gpu_stats* out) {
  if (!c || !out) return fail(MCLS_ERR_INVALID_ARG, "null argument");
  *out = c->stats;
  return MCLS_OK;
}

}  // extern "C"
