// Launchers for the gfx950 kernels of the adjoint MC / MCSA path.
// All launches go to the caller's stream and never synchronise.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcls {

// Device error flags written by kernels (first failing row + 1, 0 = ok).
enum DeviceFlag : int {
  FLAG_BAD_COLUMN = 0,   // column index out of range
  FLAG_DUPLICATE = 1,    // duplicate column within a row
  FLAG_ZERO_DIAG = 2,    // zero / missing diagonal under point Jacobi
  FLAG_NONFINITE = 3,    // NaN / Inf in a matrix value
  FLAG_LONG_ROW = 4,     // domain row longer than the alias builder handles
  FLAG_BAD_IMPORTANCE = 5,  // importance g <= 0 or non-finite
  FLAG_COUNT = 8
};

// Adjoint domain in HBM: rows of H^T (CSR).  col packs the column index in
// bits 0..30 and the sign of h in bit 31 (1 = negative).
// Packed per-row headers (one 16-byte load per step):
//   hdr32[i] = {start, len, bits(W32_i), 0}
//   hdr64[i] = {start, len, lo(W64_i), hi(W64_i)}
// Alias slot records, packed columns carry the sign of h in bit 31.
//   slot64[k] = {lo(thr_k), hi(thr_k), col_k, col_alias(k)}   (16 bytes)
// FP32 walks use 32-byte records that also carry the row header of both
// possible next states, so a step reads one record instead of a record and
// the next row's header:
//   rec32[k].a = {bits((float)thr_k), col_k, col_alias(k), len(col_k) | len(col_alias(k)) << 16}
//   rec32[k].b = {start(col_k), start(col_alias(k)), bits(W32(col_k)), bits(W32(col_alias(k)))}
// Target records for the CDF and learned samplers: entry k of a row names
// its column and that column's row header, read once the CDF search has
// chosen k (the next step then needs no header load):
//   tgt32[k]   = {col_k, start(col_k), len(col_k), bits(W32(col_k))}
//   tgt64[k].a = {col_k, start(col_k), len(col_k), 0}
//   tgt64[k].b = {lo(W64(col_k)), hi(W64(col_k)), 0, 0}
struct __attribute__((aligned(32))) Rec32 {
  uint4 a;
  uint4 b;
};

struct DomainView {
  int n;
  const int* rp;          // n+1
  const uint32_t* col;    // nnz, packed
  const double* cdf64;    // nnz
  const float* cdf32;     // nnz
  const double* W64;      // n
  const float* W32;       // n
  const double* h64;      // nnz
  const float* h32;       // nnz
  const uint4* hdr32;     // n
  const uint4* hdr64;     // n
  const Rec32* rec32;     // nnz (alias sampler, FP32 walks)
  const uint4* slot64;    // nnz (alias sampler, FP64 walks)
  const uint4* tgt32;     // nnz (CDF / learned samplers, FP32 walks)
  const Rec32* tgt64;     // nnz (CDF / learned samplers, FP64 walks; b.zw = cdf64)
  // learned proposal: P_ij = |h_ij| g_j / S_i (CDF), weight factor S_i / g_j
  const double* lcdf64;   // nnz
  const float* lcdf32;    // nnz
  const double* lfac64;   // nnz
  const float* lfac32;    // nnz
  // learned proposal sampled by alias (FP32 walks): per entry a 32-byte
  // record {thr, col, alias col, lens | start, alias start, S_i/g_j of both}
  const Rec32* lrec32;    // nnz
  // strided alias tables: n * stride records {thr, col, alias col, W_i}
  const uint4* srec;
  uint32_t stride;        // the longest row's length (>= 1)
  // the same tables in 8-byte records (nullable; n < 2^26 - 1, stride <= 32):
  // {col (26 bits) | alias slot << 26 (5 bits) | sign << 31, thr}, the row's
  // W read from W32 and the alias slot's own record read for its column
  const uint2* srec8;
};

// EST_FORWARD: the forward method on the forward domain (rows of H): history
// h starts at state h / per_state with weight 1 and adds w rhs[state] to its
// own sum at every collision; the sum goes to its start state at the end.
enum Estimator : int { EST_COLLISION = 0, EST_EXPECTED_VALUE = 1, EST_FORWARD = 2 };
enum SourceMode : int { SRC_RANDOM = 0, SRC_STRATIFIED = 1 };
// SAMPLER_ALIAS_STRIDED: alias tables padded to a fixed stride (the longest
// row) per row (FP32 walks): the slot of row i is i * stride + floor(u stride) and
// its 16-byte record carries the row's own W, so no row header is carried.
enum Sampler : int {
  SAMPLER_CDF = 0,
  SAMPLER_ALIAS = 1,
  SAMPLER_LEARNED = 2,
  SAMPLER_LEARNED_ALIAS = 3,
  SAMPLER_ALIAS_STRIDED = 4
};
// marks the records of an absorbing row (no entries): the walk ends there
constexpr uint32_t kStridedAbsorbing = 0xFFFFFFFFu;
constexpr uint32_t kStrided8Absorbing = 0x7FFFFFFFu;  // 8-byte records' first word

// Work queue of the walk kernel: one counter per XCD slot, 128 bytes apart.
constexpr int kQueueSlots = 8;
constexpr int kQueueStride = 16;  // uint64 words between counters
constexpr int kQueueWords = kQueueSlots * kQueueStride;

struct WalkArgs {
  DomainView d;
  const double* rhs;       // source vector (n)
  const uint64_t* q;       // quantised |rhs| (n)
  const uint64_t* qcdf;    // inclusive prefix of q (n); Q = qcdf[n-1]
  const uint64_t* ncdf;    // stratified: inclusive prefix of counts; M = ncdf[n-1]
  uint64_t key;            // per-solve stream key
  const uint64_t* key_dev; // when set, the key is read from device memory
  int64_t first;           // first history id
  int64_t last;            // one past the last history id (random source)
  int64_t history_length;  // max transitions
  double cutoff;           // relative weight cutoff
  double* tally64;         // the tally (FP64 for every walk precision)
  unsigned long long* steps;  // total transitions (atomic)
  unsigned long long* queue;  // kQueueWords counters, zeroed by launch_walk
  int* paths;              // recording: (last-first) x (max_len+1)
  double* wfinal;          // recording: final weight per history
  int max_len;
  int batch;               // history ids a wave takes per queue grab (multiple of 64)
  int64_t per_state;       // forward method: histories per start state
  int tile_slots;          // LDS tally tile entries per wave (power of two), 0 = none
  int tile_shift;          // 32 - log2(tile_slots)
  int tile_direct;         // 1: tile set = state mod sets (locality-ordered
                           // states); 0: Fibonacci hash of the state
  // guide table of the birth search (nullable), over the top guide_bits
  // bits of the source's range (launch_build_guide): built over qcdf for the
  // random source (a birth's draw), over ncdf for the stratified one (a
  // queue grab's first history id)
  const uint32_t* guide;
  int guide_bits;
  // random source walked in birth-state order (nullable): entry p - order_base
  // of the walk's positions holds (birth state << 32) | (id - order_base), as
  // launch_random_births drew and launch_sort_births sorted them; the walk
  // takes position p's history instead of history p (same id, same stream)
  const uint64_t* order;
  int64_t order_base;
};

// LDS tally tile: off, FP64 partial sums, partial sums in the walk's
// precision (the same as FP64 for FP64 walks), or FP32 partial sums with an
// FP32 compensation term each (FP32 walks)
enum TileKind { kTileOff = 0, kTileF64 = 1, kTileWalk = 2, kTileComp = 3 };
// LDS bytes per tile entry: an int key, the partial sum (and compensation)
inline size_t tile_entry_bytes(int tile, size_t real_bytes) {
  if (tile == kTileComp) return sizeof(int) + 2 * sizeof(float);
  return sizeof(int) + (tile == kTileWalk ? real_bytes : sizeof(double));
}

struct WalkConfig {
  int precision;   // 64 or 32
  int estimator;   // Estimator
  int source;      // SourceMode
  int sampler;     // Sampler
  bool record;
  int tile;        // TileKind (collision estimator; tallies go to tally64)
  int walks_per_lane;  // 2: two histories per lane (strided alias, FP32, no recording)
  int block;       // threads per block
  int grid;        // blocks (rounded up to a multiple of kQueueSlots)
};

// Zeroes the queue counters, then launches the walk.
hipError_t launch_walk(const WalkArgs& a, const WalkConfig& c, hipStream_t s);

// Births of a random source drawn ahead of its walk, for history ids s0 + p,
// p in [0, count) (count < 2^32), streams keyed by *key_dev when set, else
// key: out[p] = (hi << 32) | p with hi the guide bucket of the history's
// first draw when a guide table is given (launch_resolve_births then turns
// it into the state), else the state that draw selects from qcdf.
hipError_t launch_random_births(int64_t s0, int64_t count, uint64_t key, const uint64_t* key_dev,
                                const uint64_t* qcdf, const uint32_t* guide, int guide_bits, int n, uint64_t* out,
                                hipStream_t s);
// The bucket in the high word of each entry replaced by the history's birth
// state (in place; guide required), best run on entries sorted by bucket.
hipError_t launch_resolve_births(int64_t s0, int64_t count, uint64_t key, const uint64_t* key_dev,
                                 const uint64_t* qcdf, const uint32_t* guide, int guide_bits, int n, uint64_t* io,
                                 hipStream_t s);
// Stable radix sort of those entries by their high word's low key_bits bits
// (bits 32 .. 32 + key_bits): with temp == nullptr only *temp_bytes is set
// (the storage needed).
hipError_t launch_sort_births(const uint64_t* in, uint64_t* out, int64_t count, int key_bits, void* temp,
                              size_t* temp_bytes, hipStream_t s);

// Importance features: sums[0] = sum |a_jj|, sums[1] = sum |v_j| (part holds
// 2 * kResidualPartials doubles), then X (n x 8) from A's rows and v.
hipError_t launch_feature_sums(int n, const int* rp, const int* col, const void* val, bool f32, const double* v,
                               double* part, double* sums, hipStream_t s);
hipError_t launch_features(int n, const int* rp, const int* col, const void* val, bool f32, const double* v,
                           const double* sums, float* X, hipStream_t s);
// y_j = a_jj x_j
hipError_t launch_scale_by_diag(int n, const int* rp, const int* col, const void* val, bool f32, const double* x,
                                double* y, hipStream_t s);
// g_j = |a_jj| as float
hipError_t launch_diag_abs(int n, const int* rp, const int* col, const void* val, bool f32, float* g,
                           hipStream_t s);
// g = exp(clip(o, -4, 4)), 1 where mask (nullable) is 0
hipError_t launch_importance(int n, const float* o, const signed char* mask, float* g, hipStream_t s);
// T = log(clip(|x| / mean|x|, e^-4, e^4)), mean from sums[1] / n
hipError_t launch_targets(int n, const double* x, const double* sums, float* T, hipStream_t s);

// Learned-proposal tables from the importance g (n, FP32, > 0); a row that
// reaches g <= 0 or a non-finite g sets FLAG_BAD_IMPORTANCE.
// The learned proposal as alias records (Vose over |h_ij| g_j / S_i per row,
// sequential, rows of <= 64 entries; longer rows set FLAG_LONG_ROW, a bad g
// sets FLAG_BAD_IMPORTANCE).
hipError_t launch_pack_proposal_alias(int n, const int* d_rp, const uint32_t* col, const double* h64, const float* g,
                                      Rec32* lrec32, int* flags, hipStream_t s);
// FLAG_LONG_ROW for every domain row longer than limit
hipError_t launch_flag_long_rows(int n, const int* d_rp, int limit, int* flags, hipStream_t s);
hipError_t launch_build_proposal(int n, const int* d_rp, const uint32_t* col, const double* h64, const float* g,
                                 double* lcdf64, float* lcdf32, double* lfac64, float* lfac32, int* flags,
                                 hipStream_t s);

// Row headers, and (alias != 0) Vose alias tables packed into slot records.
// Rows longer than 64 entries set FLAG_LONG_ROW when alias tables are built.
hipError_t launch_pack_rows(int n, const int* d_rp, const uint32_t* col, const double* h64,
                            const double* W64, const float* W32, uint4* hdr32, uint4* hdr64,
                            Rec32* rec32, uint4* slot64, bool alias, int* flags, hipStream_t s);
// *out = max(*out, longest domain row)
hipError_t launch_max_row_len(int n, const int* d_rp, int* out, hipStream_t s);
// Strided alias tables: Vose over each row padded with zero-probability
// slots to `stride` entries; records {thr (FP32), col, alias col, W32_i}.
// srec8 (nullable, stride <= 32, columns below 2^26 - 1): the 8-byte form;
// srec nullable when srec8 is given.
hipError_t launch_pack_strided(int n, const int* d_rp, const uint32_t* col, const double* h64,
                               const double* W64, const float* W32, int stride, uint4* srec, uint2* srec8,
                               hipStream_t s);

// Target records (either output may be null).  FP64 records also carry the
// entry's own CDF value, so one load serves the search and the transition.
hipError_t launch_pack_targets(int64_t nnz, const int* d_rp, const uint32_t* col, const double* W64,
                               const float* W32, const double* cdf64, uint4* tgt32, Rec32* tgt64,
                               hipStream_t s);

// ---------------------------------------------------------------- source
// *key = splitmix64(seed ^ splitmix64(*counter)); ++*counter  (one thread)
hipError_t launch_next_key(uint64_t seed, unsigned long long* counter, uint64_t* key, hipStream_t s);
// bits of max |rhs| (as uint64) -> *bmax (must be zeroed first)
hipError_t launch_absmax(const double* v, int n, unsigned long long* bmax, hipStream_t s);
// q_i = ceil(ldexp(|rhs_i|, S - e)), bmax < 2^e
hipError_t launch_quantize(const double* rhs, int n, const unsigned long long* bmax,
                           uint64_t* q, hipStream_t s);
// stratified counts n_i = ceil(q_i nh / Q)
hipError_t launch_strat_counts(const uint64_t* q, int n, const uint64_t* qtotal,
                               uint64_t nh, uint64_t* cnt, hipStream_t s);

// ------------------------------------------------------------------ scan
// Inclusive scan of n uint64 values into out (may alias in). tmp holds
// scan_tmp_elems(n) elements.
size_t scan_tmp_elems(int64_t n);
hipError_t launch_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, uint64_t* tmp,
                           hipStream_t s);
// Exclusive scan of n int32 counts into out[0..n] (out[n] = total).
hipError_t launch_scan_i32_excl(const int* in, int* out, int64_t n, int* tmp, hipStream_t s);

// ---------------------------------------------------------------- domain
// state order (internal state k = the caller's unknown ord[k]), rows of w
// elements of elem = 1, 4 or 8 bytes: dst[k] = src[ord[k]] (gather) or
// dst[ord[k]] = src[k] (scatter)
hipError_t launch_gather_rows(int64_t n, int w, size_t elem, const int* ord, const void* src, void* dst,
                              hipStream_t s);
hipError_t launch_scatter_rows(int64_t n, int w, size_t elem, const int* ord, const void* src, void* dst,
                               hipStream_t s);
// P A P^T of a validated CSR matrix; rp_dst holds the permuted row pointers
hipError_t launch_permute_csr(int n, const int* ord, const int* inv, const int* rp_src, const int* col_src,
                              const void* val_src, bool f32, const int* rp_dst, int* col_dst, void* val_dst,
                              hipStream_t s);
hipError_t launch_fill_i32(int* p, int64_t n, int v, hipStream_t s);
hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
hipError_t launch_jacobi(int n, const int* rp, const int* col, const void* val, bool f32,
                         double* dinv, int* flags, hipStream_t s);
// column range and finiteness of a CSR matrix -> flags
hipError_t launch_validate_csr(int n, const int* rp, const int* col, const void* val, bool f32, int* flags,
                               hipStream_t s);
hipError_t launch_count_cols(int n, const int* rp, const int* col, const void* val, bool f32,
                             int* cnt, int* flags, hipStream_t s);
hipError_t launch_scatter_transpose(int n, const int* rp, const int* col, const void* val,
                                    bool f32, const double* dinv, const int* t_rp, int* fill,
                                    int* t_src, double* t_val, hipStream_t s);
// Forward domain: rows of PL A as they are (t_rp = rp, t_src = col,
// t_val = dinv_i a_ij), with the column and finiteness checks of count_cols.
hipError_t launch_copy_rows(int n, const int* rp, const int* col, const void* val, bool f32, const double* dinv,
                            int* t_rp, int* t_src, double* t_val, int* flags, hipStream_t s);
hipError_t launch_domain_rows(int n, const int* t_rp, int* t_src, double* t_val, double omega,
                              double threshold, bool has_prec, int* keep, int* flags,
                              hipStream_t s);
hipError_t launch_domain_fill(int n, const int* t_rp, const int* t_src, const double* t_val,
                              double omega, double threshold, bool has_prec, const int* d_rp,
                              uint32_t* col, double* h64, float* h32, double* cdf64,
                              float* cdf32, double* W64, float* W32, hipStream_t s);

// ---------------------------------------------------- composite operator
// General left / right preconditioners (src/MCLS_LinearProblem_impl.hpp:
// 139-221), one lane per row.  Sort each row by column (repeats set
// FLAG_DUPLICATE).
hipError_t launch_sort_rows(int n, const int* rp, int* col, double* val, int* flags, hipStream_t s);
// C = X Y for row-sorted X, Y: count pass (cnt[i] = entries of row i), then
// fill into c_rp's rows, ascending columns; C_ij accumulates the products in
// the order of X's row.  head holds nnz(X) ints of scratch.
hipError_t launch_spgemm_count(int n, const int* x_rp, const int* x_col, const int* y_rp, const int* y_col, int* head,
                               int* cnt, hipStream_t s);
hipError_t launch_spgemm_fill(int n, const int* x_rp, const int* x_col, const double* x_val, const int* y_rp,
                              const int* y_col, const double* y_val, int* head, const int* c_rp, int* c_col,
                              double* c_val, hipStream_t s);
// Row filter of MT::multiply: keep entries with |v| > (signed row max) *
// threshold that are not exact zeros.
hipError_t launch_filter_count(int n, const int* rp, const double* val, double threshold, int* keep, hipStream_t s);
hipError_t launch_filter_fill(int n, const int* rp, const int* col, const double* val, double threshold,
                              const int* out_rp, int* out_col, double* out_val, hipStream_t s);
// y = P x (FP64 CSR, rows summed left to right)
hipError_t launch_csr_spmv(int n, const int* rp, const int* col, const double* val, const double* x, double* y,
                           hipStream_t s);

// ---------------------------------------------------------------- linalg
// r = PL (b - A x); partial sums of r^2 per block into part
// (residual_blocks(n, lanes) <= kResidualPartials entries).  lanes per row
// from residual_lanes(n, nnz).
constexpr int kResidualPartials = 16384;
int residual_lanes(int64_t n, int64_t nnz);
int residual_blocks(int n, int lanes);
hipError_t launch_prec_residual(int n, int lanes, const int* rp, const int* col, const void* val,
                                bool f32, const double* dinv, const double* b, const double* x,
                                double* r, double* part, hipStream_t s);
// Guide table of a birth search over a cumulative table (qcdf for a random
// source's draws, ncdf for a stratified source's queue grabs): for
// j = 0 .. 2^bits, guide[j] = the first k with qcdf[k] > j << sh (at most
// n - 1), where sh = max(0, bitlength(Q) - bits), Q = qcdf[n - 1] read on the
// device.  The first k with qcdf[k] > t, for t in [0, Q), then lies between
// guide[t >> sh] and guide[(t >> sh) + 1].
int guide_bits_for(int64_t n);
hipError_t launch_build_guide(int n, const uint64_t* qcdf, int bits, uint32_t* guide, hipStream_t s);
// Sliced-ELL copy of A for the residual (slices of 64 rows, entry-major,
// column -1 past a row's end): widths of the (n + 63) / 64 slices, then the
// entries at the 64 * width exclusive prefix offsets.
hipError_t launch_sell_widths(int n, const int* rp, int* width, hipStream_t s);
hipError_t launch_sell_fill(int n, const int* rp, const int* col, const void* val, bool f32, const int* width,
                            const int64_t* off, int* scol, void* sval, hipStream_t s);
// r = PL (b - A x) over the sliced copy, each row summed left to right;
// partial sums of r^2 per block into part (sell_blocks(slices) entries).
int sell_blocks(int nslices);
hipError_t launch_sell_residual(int n, const int* width, const int64_t* off, const int* scol, const void* sval,
                                bool f32, const double* dinv, const double* b, const double* x, double* r,
                                double* part, hipStream_t s);
// out[0] = sum of part[0..m) in a fixed order; out_host (pinned, nullable)
hipError_t launch_sum_partials(const double* part, int m, double* out, hipStream_t s);
// Convergence criteria of the domain operator M (rows of H^T).
// One power step: y = M (x / sqrt(*nrm2_prev)) (nrm2_prev nullable = 1),
// *nrm2 = ||y||^2, *logsum += log ||y||.  mode 0: H, 1: |H|, 2: H* (|h_ij| W_i).
// part holds kResidualPartials doubles.
hipError_t launch_domain_power(int n, const int* d_rp, const uint32_t* col, const double* h64, const double* W64,
                               int mode, const double* x, const double* nrm2_prev, double* y, double* part,
                               double* nrm2, double* logsum, hipStream_t s);
// x_i = 0.5 + u(splitmix(seed ^ splitmix(i)))
hipError_t launch_power_start(int n, uint64_t seed, double* x, hipStream_t s);
// Row-sum and balance statistics: mm (4 words) = bits of {max W, min W,
// min ratio, max ratio}, sums[0] = sum of ratios, sums[1] = rows counted;
// ratio_i = max_k |h_ik| / max_k |h_ki| over non-empty rows; cmax: n words,
// part: 2 * kResidualPartials doubles.
hipError_t launch_balance(int n, const int* d_rp, const uint32_t* col, const double* h64, const double* W64,
                          unsigned long long* cmax, unsigned long long* mm, double* part, double* sums,
                          hipStream_t s);
// ||PL b||_2^2 partials
hipError_t launch_scaled_sq(int n, const double* dinv, const double* b, double* part,
                            hipStream_t s);
// batched systems (block-diagonal operator, rows [off[s], off[s+1]))
// out[s] = sum (dinv v)^2 (square) or sum |v| over system s
hipError_t launch_segment_norm(int nsys, const int64_t* off, const double* v, const double* dinv, bool square,
                               double* out, hipStream_t s);
hipError_t launch_batch_check(int nsys, const double* r2, const double* b2, double tol, int* conv,
                              int* unconverged, hipStream_t s);
hipError_t launch_batch_source(int n, const int* sys_of, const int64_t* off, const double* r, double omega,
                               const int* conv, const double* l1, double* rhs, double* mag, hipStream_t s);
// x += a * y
hipError_t launch_axpy(int n, double a, const double* y, double* x, hipStream_t s);
// y = a * x
hipError_t launch_scale_copy(int n, double a, const double* x, double* y, hipStream_t s);
// y_i = d_i * x_i
hipError_t launch_diag_scale(int n, const double* d, const double* x, double* y, hipStream_t s);
// x += tally * inv (+ rhs for the expected-value estimator); tally zeroed.
// inv = 1/nh, or 1/M read from *mtotal when mtotal != NULL.
hipError_t launch_apply_tally(int n, double* x, double* t64, float* t32, const double* rhs,
                              bool add_rhs, double inv, const uint64_t* mtotal, hipStream_t s);
// x = tally * inv (+ rhs); tally zeroed.
hipError_t launch_store_tally(int n, double* x, double* t64, float* t32, const double* rhs,
                              bool add_rhs, double inv, const uint64_t* mtotal, hipStream_t s);

}  // namespace mcls
