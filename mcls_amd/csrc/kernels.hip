// gfx950 kernels for the adjoint Neumann-Ulam / MCSA path.
//
// Built with -ffp-contract=off: every floating-point operation rounds as
// written, so the domain tables, source weights and walk paths are
// bit-identical to the CPU restatement that checks them.
//
// Reference semantics followed (paths relative to the MCLS source root):
//   src/MCLS_AlmostOptimalDomain_impl.hpp:258-366  domain rows (h, eps drop, CDF, W)
//   src/MCLS_AlmostOptimalDomain.hpp:255-284       processTransition
//   src/MCLS_DomainTransporter_impl.hpp:85-115     tally -> transition -> terminate
//   src/MCLS_AdjointTally.hpp:115-123              collision tally
//   src/MCLS_SamplingTools.hpp:67-80               lower_bound CDF inversion
//   src/MCLS_Xorshift.hpp:103-109                  Xorshift (13,7,17)
//   src/MCLS_LinearProblem_impl.hpp:411-442        rp = PL (b - A x)
//   tpetra/src/MCLS_TpetraHelpers.hpp:196-263      composite threshold filter
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cfloat>
#include <climits>
#include <cmath>
#include <type_traits>

#include "kernels.hpp"

namespace mcls {
namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kRedThreads = 256;
constexpr int kWave = 64;  // gfx950 wavefront

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += kGolden;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t xorshift64(uint64_t& x) {
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return x;
}

// The value x holds in the wave's first active lane, as a wave-uniform value
// (kept in scalar registers, so the branches that depend on it are scalar).
__device__ __forceinline__ uint64_t first_lane_u64(uint64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t history_state(uint64_t key, uint64_t hid) {
  uint64_t s = splitmix64(key ^ splitmix64(hid));
  return s ? s : kGolden;
}

__device__ __forceinline__ double u01_f64(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }
__device__ __forceinline__ float u01_f32(uint64_t r) { return (float)(r >> 40) * 0x1.0p-24f; }

// floor((r >> 11) * Q / 2^53)
__device__ __forceinline__ uint64_t mul_u53(uint64_t r, uint64_t Q) {
  uint64_t a = r >> 11;
  uint64_t lo = a * Q;
  uint64_t hi = __umul64hi(a, Q);
  return (hi << 11) | (lo >> 53);
}

// first k in [0, n) with cdf[k] > t
__device__ __forceinline__ int upper_bound_u64(const uint64_t* __restrict__ cdf, int n, uint64_t t) {
  int lo = 0, len = n;
  while (len > 0) {
    int half = len >> 1;
    if (cdf[lo + half] <= t) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

// The state of a random birth: the first k with qcdf[k] > t, for a draw
// t in [0, total).  With a guide table the search covers only
// [guide[j], guide[j + 1]], j = t's top guide_bits bits of total's range,
// instead of all n entries: the same k, a few probes instead of log2(n).
__device__ __forceinline__ int birth_state(const uint64_t* __restrict__ qcdf, const uint32_t* __restrict__ guide,
                                           int guide_bits, int n, uint64_t total, uint64_t t) {
  if (guide) {
    const int bl = 64 - __clzll((long long)total);
    const int sh = bl > guide_bits ? bl - guide_bits : 0;
    const uint64_t j = t >> sh;
    const int lo = (int)guide[j], hi = (int)guide[j + 1];
    return lo + upper_bound_u64(qcdf + lo, hi - lo + 1, t);
  }
  return upper_bound_u64(qcdf, n, t);
}

// First k in [lo, n) with cdf[k] > t (n when none), found by the whole wave
// in a few dependent round trips instead of log2(n): each round the 64 lanes
// probe 64 evenly spaced entries of the range left and the ballot of
// "cdf <= t" (a prefix of the lanes: cdf is non-decreasing) narrows it
// 64-fold.  Wave-uniform arguments; every lane of the wave must call it and
// every lane gets the answer.
__device__ __forceinline__ int wave_upper_bound_u64(const uint64_t* __restrict__ cdf, int lo, int n, uint64_t t,
                                                    int lane) {
  int hi = n;  // the answer lies in [lo, hi]
  while (hi - lo > 64) {
    const int64_t span = hi - lo;
    const int p = lo + (int)(span * (lane + 1) / 64) - 1;  // in [lo, hi - 1]
    const int c = __popcll(__ballot(cdf[p] <= t));
    const int nlo = c > 0 ? lo + (int)(span * c / 64) : lo;  // one past probe c - 1
    if (c < 64) hi = lo + (int)(span * (c + 1) / 64) - 1;    // probe c: cdf > t there
    lo = nlo;
  }
  return lo + __popcll(__ballot(lane < hi - lo && cdf[lo + lane] <= t));
}

// first k in [0, n) with cdf[k] >= u  (std::lower_bound)
template <typename Real>
__device__ __forceinline__ int lower_bound_cdf(const Real* __restrict__ cdf, int n, Real u) {
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

// The same index as lower_bound_cdf for a non-decreasing cdf whose last entry
// is 1 (u in [0, 1)), found from the guess g = floor(u n): cdf[g - 1] and
// cdf[g] are fetched together, which settles a near-uniform row in one round
// trip; otherwise a bisection runs on the side of g the two values leave
// open.  `overlap` runs between issuing the two loads and using them.
template <typename Real, typename F>
__device__ __forceinline__ int lower_bound_guess(const Real* __restrict__ cdf, int n, Real u, F&& overlap) {
  int g = (int)(u * (Real)n);
  g = g < n - 1 ? g : n - 1;
  const Real c1 = cdf[g];
  const Real c0r = cdf[g > 0 ? g - 1 : 0];
  overlap();
  const Real c0 = g > 0 ? c0r : (Real)-1;
  if (c1 < u) return g + 1 + lower_bound_cdf<Real>(cdf + g + 1, n - g - 1, u);
  if (c0 >= u) return lower_bound_cdf<Real>(cdf, g - 1, u);
  return g;
}

__device__ __forceinline__ void tally_add(double* p, double v) { unsafeAtomicAdd(p, v); }

template <typename Real> struct Tables;
template <> struct Tables<double> {
  __device__ static const double* lcdf(const DomainView& d) { return d.lcdf64; }
  __device__ static const double* lfac(const DomainView& d) { return d.lfac64; }
  __device__ static const double* cdf(const DomainView& d) { return d.cdf64; }
  __device__ static const double* W(const DomainView& d) { return d.W64; }
  __device__ static const double* h(const DomainView& d) { return d.h64; }
  __device__ static double u01(uint64_t r) { return u01_f64(r); }
};
template <> struct Tables<float> {
  __device__ static const float* lcdf(const DomainView& d) { return d.lcdf32; }
  __device__ static const float* lfac(const DomainView& d) { return d.lfac32; }
  __device__ static const float* cdf(const DomainView& d) { return d.cdf32; }
  __device__ static const float* W(const DomainView& d) { return d.W32; }
  __device__ static const float* h(const DomainView& d) { return d.h32; }
  __device__ static float u01(uint64_t r) { return u01_f32(r); }
};

template <typename Real> struct Row;
template <> struct Row<double> {
  __device__ static uint4 hdr(const DomainView& d, int i) { return d.hdr64[i]; }
  __device__ static uint32_t target(const DomainView& d, uint32_t k, uint32_t& beg, uint32_t& len, double& W);
  __device__ static double W(uint4 h) { return __hiloint2double((int)h.w, (int)h.z); }
  // alias draw: slot from the high 32 bits, keep-or-alias from the low 32;
  // `overlap` runs between issuing the slot load and using it
  template <typename F>
  __device__ static uint32_t pick(const DomainView& d, uint32_t start, uint32_t len, uint64_t r, F&& overlap) {
    const uint32_t slot = __umulhi((uint32_t)(r >> 32), len);
    const uint4 s = d.slot64[start + slot];
    overlap();
    const double u2 = (double)(uint32_t)r * 0x1.0p-32;
    return u2 < __hiloint2double((int)s.y, (int)s.x) ? s.z : s.w;
  }
};
template <> struct Row<float> {
  __device__ static uint4 hdr(const DomainView& d, int i) { return d.hdr32[i]; }
  __device__ static float W(uint4 h) { return __uint_as_float(h.z); }
  // target record k: packed column, and that column's row header
  __device__ static uint32_t target(const DomainView& d, uint32_t k, uint32_t& beg, uint32_t& len, float& W) {
    const uint4 t = d.tgt32[k];
    beg = t.y;
    len = t.z;
    W = __uint_as_float(t.w);
    return t.x;
  }
};
__device__ uint32_t Row<double>::target(const DomainView& d, uint32_t k, uint32_t& beg, uint32_t& len, double& W) {
  const Rec32 t = d.tgt64[k];
  beg = t.a.y;
  len = t.a.z;
  W = __hiloint2double((int)t.b.y, (int)t.b.x);
  return t.a.x;
}



// CDF draw (lower_bound of u in the row's CDF) and transition from the FP64
// target records, whose b.zw hold each entry's CDF value; an FP32 walk
// rounds the record's values to float, which gives exactly its cdf32 / W32
// entries.  The search starts at the guess k = floor(u len): record k and
// the CDF value of k - 1 come in one round trip, and when they bracket u
// (cdf[k-1] < u <= cdf[k]) record k is also the transition.  Otherwise k
// moves one entry towards u per round trip (the lines are already cached),
// up to kCdfWalk moves, after which a bisection over the walk precision's
// CDF settles the rest.
template <typename Real, typename F>
__device__ __forceinline__ uint32_t cdf_step(const DomainView& d, uint32_t& beg, uint32_t& len, Real& W, Real u,
                                             F&& overlap) {
  constexpr int kCdfWalk = 3;
  const Rec32* __restrict__ t = d.tgt64 + beg;
  const int n = (int)len;
  int k = (int)(u * (Real)n);
  k = k < n - 1 ? k : n - 1;
  Rec32 r = t[k];
  uint2 l = *(const uint2*)&t[k > 0 ? k - 1 : 0].b.z;
  overlap();
  for (int move = 0;; ++move) {
    const Real ck = (Real)__hiloint2double((int)r.b.w, (int)r.b.z);
    const Real cl = k > 0 ? (Real)__hiloint2double((int)l.y, (int)l.x) : (Real)-1;
    const bool right = ck < u;  // then k < n - 1, as cdf[n - 1] = 1 > u
    if (!right && !(cl >= u)) break;
    if (move == kCdfWalk) {
      const Real* __restrict__ c = Tables<Real>::cdf(d) + beg;
      k = right ? k + 1 + lower_bound_cdf<Real>(c + k + 1, n - k - 1, u) : lower_bound_cdf<Real>(c, k - 1, u);
      r = t[k];
      break;
    }
    k += right ? 1 : -1;  // cl >= u implies k > 0
    r = t[k];
    l = *(const uint2*)&t[k > 0 ? k - 1 : 0].b.z;
  }
  beg = r.a.y;
  len = r.a.z;
  W = (Real)__hiloint2double((int)r.b.y, (int)r.b.x);
  return r.a.x;
}

// --------------------------------------------------------------------------
// Walk kernel.  Each lane owns one live history at a time.  Lanes whose
// history ended are refilled together from a per-XCD work queue: the wave
// takes batches of a.batch consecutive history ids with one atomic, so the
// walks in flight on an XCD were born next to each other (the source emits
// history ids in state order) and share L2 / MALL lines.  Histories follow
// the reference loop order: tally at the current state, sample the
// transition (row CDF inversion or alias slot), multiply the weight by
// W_i sign(h_ij), stop when steps >= L or |w| < cutoff |w0|.
// --------------------------------------------------------------------------
// With TILE, collision tallies first go to a per-wave set-associative tile of
// states in LDS (a.tile_slots entries of {state, partial sum}); a state that
// finds every slot of its set taken by others goes straight to HBM.  A wave flushes its
// tile into the FP64 HBM tally at every work-queue grab and when it exits.
// TILE = kTileF64: FP64 partial sums; kTileWalk: partial sums in the walk's
// precision (FP32 walks: 8 bytes of LDS per entry instead of 12); kTileComp:
// FP32 partial sums with a compensation term per slot (Neumaier: the exact
// rounding error of each LDS addition, recovered from the value the atomic
// returns, is summed beside the partial sum), both added to the FP64 tally
// at the flush, so a grab's sum reaches HBM as if it had been summed in about
// twice FP32's precision (12 bytes per entry).
// Every walk's HBM tally is FP64 (FP32 walks convert their weights on the
// way in), so nothing is summed in FP32 across flushes or in an untiled walk.
// Slots per set of the tile (8 on the 256^3 27-point chosen walk: 30.8 ms
// per launch against 31.4 with 4-way sets, whose sets were full for 15% of
// the tallies, and 35.0 with 16-way sets, whose key reads cost more than
// they keep; 4-way sets were 2% faster than 2-way and 7% faster than direct
// mapping).  A set's keys are read in 16-byte LDS loads.
constexpr int kTileWaysLog2 = 3;
constexpr int kTileWays = 1 << kTileWaysLog2;

// The partial sums of one set, aligned so that a flush reads and clears them
// with single LDS vector accesses (sets start at multiples of 4 entries).
template <typename T>
struct alignas(4 * sizeof(T)) TileSet {
  T s[4];
};

// One tally into a tile slot (workgroup-scope LDS atomics: lanes of one wave
// may hit the same slot).  kTileComp: the atomic returns the sum it found, so
// the lane knows the rounded sum s = fl(old + v) the slot now holds and adds
// the exact error (old + v) - s (Knuth's TwoSum) to the slot's compensation.
template <int TILE, typename TAcc>
__device__ __forceinline__ void tile_add(TAcc* tvals, float* tcomp, int slot, TAcc v) {
  if constexpr (TILE == kTileComp) {
    const float old = __hip_atomic_fetch_add(&tvals[slot], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const float sum = old + v;
    const float bv = sum - old;
    const float err = (old - (sum - bv)) + (v - bv);
    if (err != 0.0f) __hip_atomic_fetch_add(&tcomp[slot], err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    (void)tcomp;
    __hip_atomic_fetch_add(&tvals[slot], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// One wave's LDS tally tile (TILE != kTileOff): a.tile_slots entries of
// {state, partial sum (, compensation)} in kTileWays-way sets, carved out of
// the block's dynamic LDS (keys of every wave, then sums, then compensations).
template <int TILE, typename TAcc>
struct WaveTile {
  int* keys = nullptr;    // state held by each slot, -1 = empty
  TAcc* vals = nullptr;   // partial sums
  float* comp = nullptr;  // kTileComp: compensation terms
  int slots = 0;
  int shift = 0;   // 32 - log2(slots)
  int direct = 0;  // set = state mod sets (else a Fibonacci hash)

  // wave-uniform
  __device__ void init(unsigned char* lds, const WalkArgs& a, int lane) {
    const int wv = threadIdx.x >> 6;
    const size_t nw = (size_t)(blockDim.x >> 6);
    slots = a.tile_slots;
    shift = a.tile_shift;
    direct = a.tile_direct;
    keys = (int*)lds + wv * slots;
    vals = (TAcc*)(lds + nw * slots * sizeof(int)) + wv * slots;
    if constexpr (TILE == kTileComp) comp = (float*)(lds + nw * slots * (sizeof(int) + sizeof(TAcc))) + wv * slots;
    for (int k = lane; k < slots; k += 64) {
      keys[k] = -1;
      vals[k] = 0;
      if constexpr (TILE == kTileComp) comp[k] = 0.0f;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }

  // One tally v at state st.  The state may sit in any slot of its set, whose
  // keys are read together.  Slots are claimed by compare-and-swap, always the
  // first one that read empty (so a set's held slots are a prefix of it and a
  // state holds at most one: lanes of one wave with the same state agree on
  // the slot); a lane that loses that slot to another state tries the next
  // one once.  A full set, or a second loss, sends the tally to the FP64 tally.
  template <typename V>
  __device__ void add(double* __restrict__ tally, int st, V v) {
    const uint32_t s0 = direct ? ((uint32_t)st & ((uint32_t)(slots >> kTileWaysLog2) - 1u)) << kTileWaysLog2
                               : (((uint32_t)st * 2654435761u) >> (shift + kTileWaysLog2)) << kTileWaysLog2;
    static_assert(kTileWays % 4 == 0, "a set's keys are read four at a time");
    int kq[kTileWays];
#pragma unroll
    for (int q = 0; q < kTileWays / 4; ++q) {
      const uint4 kk = *reinterpret_cast<const uint4*>(keys + s0 + 4 * q);
      kq[4 * q] = (int)kk.x;
      kq[4 * q + 1] = (int)kk.y;
      kq[4 * q + 2] = (int)kk.z;
      kq[4 * q + 3] = (int)kk.w;
    }
    // selects, not a first-match chain: st is in at most one slot
    int slot = -1, empty = kTileWays;
#pragma unroll
    for (int j = kTileWays - 1; j >= 0; --j) {
      slot = kq[j] == st ? j : slot;
      empty = kq[j] < 0 ? j : empty;
    }
    if (slot < 0 && empty < kTileWays) {
      int k = atomicCAS(&keys[s0 + empty], -1, st);
      if (k < 0 || k == st) {
        slot = empty;
      } else if (empty + 1 < kTileWays) {
        k = atomicCAS(&keys[s0 + empty + 1], -1, st);
        if (k < 0 || k == st) slot = empty + 1;
      }
    }
    if (slot >= 0)
      tile_add<TILE>(vals, comp, (int)s0 + slot, (TAcc)v);
    else
      tally_add(&tally[st], (double)v);
  }

  // Every held slot's partial sum (and compensation, met in FP64) into the
  // FP64 tally, and the tile cleared.  Wave-uniform: every lane of the wave
  // runs it together; four slots (one set) per lane and pass, keys and sums
  // in vector reads, a set cleared by vector writes when any slot was held.
  __device__ void flush(double* __restrict__ tally, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int k = lane * 4; k < slots; k += 256) {
      const uint4 kk = *reinterpret_cast<const uint4*>(keys + k);
      const int kq[4] = {(int)kk.x, (int)kk.y, (int)kk.z, (int)kk.w};
      // the AND of the keys is negative only when all four slots are empty
      if ((kq[0] & kq[1] & kq[2] & kq[3]) >= 0) {
        const TileSet<TAcc> v = *reinterpret_cast<const TileSet<TAcc>*>(vals + k);
        if constexpr (TILE == kTileComp) {
          const TileSet<float> cv = *reinterpret_cast<const TileSet<float>*>(comp + k);
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (kq[q] >= 0) tally_add(&tally[kq[q]], (double)v.s[q] + (double)cv.s[q]);
          *reinterpret_cast<TileSet<float>*>(comp + k) = TileSet<float>{};
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (kq[q] >= 0) tally_add(&tally[kq[q]], (double)v.s[q]);
        }
        *reinterpret_cast<uint4*>(keys + k) = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        *reinterpret_cast<TileSet<TAcc>*>(vals + k) = TileSet<TAcc>{};
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
};

// Stratified births of one queue grab, history ids [qnext, qend)
// (consecutive): the grab's first and last birth states slo and shi, and,
// when they span at most 64 states (`window`), lane L holding state slo + L's
// last id relative to the grab's start (win_c) and its birth weight
// (win_w0), so that a birth is a shuffle search with no memory access.
struct StratWindow {
  int slo = 0, shi = 0;
  bool window = false;
  int64_t qbase = 0;
  uint32_t win_c = 0xFFFFFFFFu;
  double win_w0 = 0.0;
};

// The grab's window.  The first state (first k with ncdf[k] > qnext) lies in
// [guide[j], guide[j + 1]], j = qnext >> gsh (a.guide built over ncdf,
// nullable): one guide probe and one coalesced load of
// V[m] = ncdf[lo - 1 + m] settle it, and V also holds the window's counts
// unless the window runs past it.  Wave-uniform: every lane of the wave calls
// it with the same grab.
__device__ __forceinline__ void grab_stratified(const WalkArgs& a, int n, int gsh, double totald, int64_t qnext,
                                                int64_t qend, int lane, StratWindow& g) {
  const uint64_t t0 = (uint64_t)qnext, t1 = (uint64_t)(qend - 1);
  int lo = 0, hi = n - 1;  // the first state lies in [lo, hi]
  if (a.guide) {
    const uint64_t j = t0 >> gsh;
    lo = (int)a.guide[j];
    hi = (int)a.guide[j + 1];
  }
  if (hi - lo > 62) lo = wave_upper_bound_u64(a.ncdf, lo, hi + 1, t0, lane);  // no guide, or a long birthless run
  const int km = lo - 1 + lane;
  const uint64_t V = km < 0 ? 0ULL : (km < n ? a.ncdf[km] : ~0ULL);
  // lanes 1..63 hold ncdf[lo..lo + 62], non-decreasing: those <= t0 are a
  // prefix, and ncdf[hi] > t0 with hi <= lo + 62
  g.slo = lo + __popcll(__ballot(lane >= 1 && V <= t0));
  g.qbase = qnext;
  const int dd = g.slo - lo;  // 0 <= dd <= 62
  const int k = g.slo + lane;
  const int s1 = lane + dd + 1;  // V's index of ncdf[k]
  uint64_t ck = __shfl(V, s1 & 63);
  uint64_t pk = __shfl(V, (lane + dd) & 63);  // ncdf[k - 1] (0 for k = 0)
  if (s1 > 63) {  // past V: k >= 63
    ck = k < n ? a.ncdf[k] : ~0ULL;
    pk = k - 1 < n ? a.ncdf[k - 1] : ~0ULL;
  }
  const int c = __popcll(__ballot(ck <= t1));
  g.window = c < 64;
  if (g.window) {
    g.shi = g.slo + c;
    g.win_c = 0xFFFFFFFFu;
    g.win_w0 = 0.0;
    if (lane <= c) {
      const uint64_t nk = ck - pk;
      const uint64_t rel = ck - (uint64_t)g.qbase;
      g.win_c = rel < 0xFFFFFFFEULL ? (uint32_t)rel : 0xFFFFFFFEu;
      g.win_w0 = nk ? a.rhs[k] * (totald / (double)nk) : 0.0;
    }
  } else {
    g.shi = wave_upper_bound_u64(a.ncdf, g.slo + 64, n, t1, lane);
  }
}

// Window birth at position keyb of the grab (history id g.qbase + keyb): its
// state's offset from g.slo (the first window state whose last id passes it,
// by binary lifting over the lanes) and its weight.  Wave-uniform: every lane
// shuffles, each with its own position.
__device__ __forceinline__ void window_pos(const StratWindow& g, uint32_t keyb, int& bpos, double& bw0) {
  bpos = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    const uint32_t cv = __shfl(g.win_c, bpos + step - 1);
    if (cv <= keyb) bpos += step;
  }
  bw0 = __shfl(g.win_w0, bpos);
}

// What a birth at one position of a grab needs before its first step: the
// history's stream state and, in a stratified window, its state's offset
// from the window's first state and its birth weight.  A wave computes them
// for 64 consecutive positions at once (lane L: position 64 blk + L), with
// every lane busy, instead of per birth round with only the newborn lanes.
struct BirthSlot {
  uint64_t r = 0;
  int b = 0;
  double w = 0.0;
};

// Stratified birth outside a window: the first k in [g.slo, g.shi] with
// ncdf[k] > hid, and its weight.
__device__ __forceinline__ int strat_birth(const WalkArgs& a, const StratWindow& g, double totald, int64_t hid,
                                           double& w0) {
  const int st = g.slo + upper_bound_u64(a.ncdf + g.slo, g.shi - g.slo + 1, (uint64_t)hid);
  const uint64_t nk = a.ncdf[st] - (st ? a.ncdf[st - 1] : 0);
  w0 = a.rhs[st] * (totald / (double)nk);
  return st;
}

// The shift of a history id to its guide bucket (as k_build_guide)
__device__ __forceinline__ int guide_shift(uint64_t total, int bits) {
  const int bl = total ? 64 - __clzll((long long)total) : 0;
  return bl > bits ? bl - bits : 0;
}

template <typename Real, int EST, int SRC, int SAMP, bool REC, int TILE>
__global__ __launch_bounds__(256) void k_walk(WalkArgs a) {
  using Acc = double;  // the HBM tally
  using TAcc = typename std::conditional<TILE == kTileWalk || TILE == kTileComp, Real, double>::type;
  const DomainView& d = a.d;
  const int n = d.n;
  const Real* __restrict__ ht = Tables<Real>::h(d);
  Acc* __restrict__ tally = a.tally64;
  const uint32_t* __restrict__ colp = d.col;

  int64_t hi_id = a.last;
  uint64_t total = 0;  // Q (random) or M (stratified)
  if (EST == EST_FORWARD) {
    // every state starts per_state histories; the source tables are unused
  } else if (SRC == SRC_RANDOM) {
    total = a.qcdf[n - 1];
    if (total == 0) return;
  } else {
    total = a.ncdf[n - 1];
    if ((int64_t)total < hi_id) hi_id = (int64_t)total;
  }
  const int64_t span = hi_id - a.first;
  if (span <= 0) return;
  const double totald = (double)total;
  const uint64_t key = a.key_dev ? *a.key_dev : a.key;
  // workgroups are dispatched round-robin over the 8 XCDs: slot = XCD
  const int qs = blockIdx.x % kQueueSlots;
  const int64_t sbeg = a.first + span * qs / kQueueSlots;
  const int64_t send = a.first + span * (qs + 1) / kQueueSlots;
  unsigned long long* ctr = a.queue + qs * kQueueStride;
  const int lane = __lane_id();

  WaveTile<TILE, TAcc> tile;  // TILE: this wave's LDS tally tile
  if constexpr (TILE != kTileOff) {
    extern __shared__ __align__(16) unsigned char walk_lds[];
    tile.init(walk_lds, a, lane);
  }
  // wave-uniform: every lane of the wave runs it together
  auto flush_tile = [&]() {
    if constexpr (TILE != kTileOff) tile.flush(tally, lane);
  };

  // The current row's header is carried from the record that chose the
  // state (FP32 alias records, CDF / learned target records); FP64 alias
  // walks load it per step.
  constexpr bool FUSED = SAMP == SAMPLER_ALIAS && sizeof(Real) == 4;
  // strided alias records hold the current row's W (the 8-byte form reads
  // it from W32 beside the record): no header at all
  constexpr bool STRIDED = SAMP == SAMPLER_ALIAS_STRIDED;
  constexpr bool CARRY = (SAMP != SAMPLER_ALIAS || FUSED) && !STRIDED;
  int64_t qnext = 0, qend = 0;  // wave-uniform batch cursor
  StratWindow grab;             // stratified: birth states of the current batch
  const int gsh = SRC == SRC_STRATIFIED && EST != EST_FORWARD ? guide_shift(total, a.guide_bits) : 0;
  // The births of the current grab, 64 positions at a time: c0 holds block
  // cblk of positions relative to gbase, c1 block cblk + 1 once c1ok
  // (wave-uniform cursor; not with births drawn ahead, whose ids are
  // permuted).
  const bool cached = !(SRC == SRC_RANDOM && EST != EST_FORWARD && a.order);
  BirthSlot c0, c1;
  int64_t gbase = 0, cblk = -1;
  bool c1ok = false;
  auto fill = [&](BirthSlot& c, int64_t blk) {  // wave-uniform
    const int64_t p = blk * 64 + lane;
    c.r = history_state(key, (uint64_t)(gbase + p));
    if (SRC == SRC_STRATIFIED && EST != EST_FORWARD && grab.window) window_pos(grab, (uint32_t)p, c.b, c.w);
  };
  bool exhausted = false;       // wave-uniform
  bool alive = false;
  int64_t hid = 0;
  int state = 0;
  uint32_t beg = 0, len = 0;    // current domain row
  Real Wrow = 0;                // its weight W_i
  Real w = 0, wc = 0;
  int start = 0;                // forward: start state
  double fsum = 0.0;            // forward: the history's sum of w rhs[state]
  int64_t steps = 0;
  uint64_t rng = 0;
  uint64_t my_steps = 0;
  int* path = nullptr;

  for (;;) {
    if (!exhausted) {
      uint64_t need = __ballot(!alive);
      while (need != 0) {
        if (qnext >= qend) {
          flush_tile();
          const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
          unsigned long long b = 0;
          if (lane == leader) b = atomicAdd(ctr, (unsigned long long)a.batch);
          b = first_lane_u64(b);  // the leader is the first active lane
          qnext = sbeg + (int64_t)b;
          qend = qnext + a.batch < send ? qnext + a.batch : send;
          if (qnext >= send) {
            exhausted = true;
            break;
          }
          // the batch's ids are consecutive: bound its birth states once
          // (every lane of the wave is here), so each birth searches only
          // that window of the count table, or none
          if (SRC == SRC_STRATIFIED && EST != EST_FORWARD)
            grab_stratified(a, n, gsh, totald, qnext, qend, lane, grab);
          gbase = qnext;
          cblk = -1;
          c1ok = false;
        }
        const int64_t avail = qend - qnext;
        const int cnt = __popcll(need);
        const int take = (int64_t)cnt < avail ? cnt : (int)avail;
        const int rank = __popcll(need & __lanemask_lt());
        // this round's positions [q, q + take) lie in blocks cblk and cblk + 1
        const int64_t q = qnext - gbase;
        BirthSlot bs;
        if (cached) {
          if ((q >> 6) != cblk) {  // the cursor entered the next block
            if (c1ok)
              c0 = c1;
            else
              fill(c0, q >> 6);
            cblk = q >> 6;
            c1ok = false;
          }
          if (q + take > (cblk + 1) * 64 && !c1ok) {
            fill(c1, cblk + 1);
            c1ok = true;
          }
          // position q + rank sits in lane (q + rank) mod 64, of block cblk
          // when that lane is at or past q mod 64, else of block cblk + 1
          const BirthSlot& m = lane >= (int)(q & 63) ? c0 : c1;
          const int src = (int)((q + rank) & 63);
          bs.r = __shfl(m.r, src);
          if (SRC == SRC_STRATIFIED && EST != EST_FORWARD && grab.window) {
            bs.b = __shfl(m.b, src);
            bs.w = __shfl(m.w, src);
          }
        }
        int bpos = bs.b;
        double bw0 = bs.w;
        if (!alive && rank < take) {
          hid = qnext + rank;
          if (SRC == SRC_RANDOM && EST != EST_FORWARD && a.order) {
            // births drawn ahead and sorted by state: position hid holds
            // this history's id and the state its first draw selected
            const uint64_t kv = a.order[hid - a.order_base];
            hid = a.order_base + (int64_t)(kv & 0xFFFFFFFFull);
            state = (int)(kv >> 32);
          }
          rng = cached ? bs.r : history_state(key, (uint64_t)hid);
          double w0;
          if (EST == EST_FORWARD) {
            state = (int)(hid / a.per_state);
            start = state;
            fsum = 0.0;
            w0 = 1.0;
          } else if (SRC == SRC_RANDOM) {
            const uint64_t t = mul_u53(xorshift64(rng), total);  // the birth draw
            if (!a.order) state = birth_state(a.qcdf, a.guide, a.guide_bits, n, total, t);
            w0 = a.rhs[state] * (totald / (double)a.q[state]);
          } else if (grab.window) {
            state = grab.slo + bpos;
            w0 = bw0;
          } else {
            state = strat_birth(a, grab, totald, hid, w0);
          }
          w = (Real)w0;
          wc = (Real)a.cutoff * (w < 0 ? -w : w);
          steps = 0;
          alive = true;
          if (CARRY) {
            const uint4 hd = Row<Real>::hdr(d, state);
            // Wait for the header here, on the refill path.  Left to the
            // first use after the refill, the wait would also cover the
            // previous step's tally atomic on every step of every lane.
            asm volatile("" ::"v"(hd.x), "v"(hd.y), "v"(hd.z), "v"(hd.w));
            beg = hd.x;
            len = hd.y;
            Wrow = Row<Real>::W(hd);
          }
          if (REC) {
            path = a.paths + (size_t)(hid - a.first) * (size_t)(a.max_len + 1);
            path[0] = state;
          }
        }
        qnext += take;
        need = __ballot(!alive);
      }
    }
    if (!alive) break;  // queue exhausted and no history left on this lane

    if (!CARRY && !STRIDED) {
      const uint4 hd = Row<Real>::hdr(d, state);
      beg = hd.x;
      len = hd.y;
      Wrow = Row<Real>::W(hd);
    }
    // The step's tally at the current state, with the weight it arrived
    // with.  It is issued after the transition's first load, so that the
    // atomic's round trip overlaps that load's instead of preceding it (the
    // counter that tracks outstanding memory operations drains in order).
    // One addition v at state st: into this wave's LDS tile, or the FP64
    // tally when the tile is off or the state's set is full.
    auto tally_to = [&](int st, Real v) {
      if constexpr (TILE != kTileOff)
        tile.add(tally, st, v);
      else
        tally_add(&tally[st], (Acc)v);
    };
    auto tally_step = [&]() {
      if (EST == EST_FORWARD) {
        fsum = fsum + (double)w * a.rhs[state];
      } else if (EST == EST_EXPECTED_VALUE) {
        // the row's scatter, each entry through the tile (a state order keeps
        // a row's columns, and the rows of the wave's walks, in few sets)
        for (uint32_t k = 0; k < len; ++k) tally_to((int)(colp[beg + k] & 0x7FFFFFFFu), w * ht[beg + k]);
      } else {
        tally_to(state, w);
      }
    };
    bool end = !STRIDED && len == 0;  // absorbing row (all of h dropped)
    if (end) {
      tally_step();
    } else {
      const uint64_t r = xorshift64(rng);
      uint32_t cp;
      bool absorbed = false;  // strided: the record says the row is empty
      if constexpr (STRIDED) {
        const uint32_t slot = __umulhi((uint32_t)(r >> 32), d.stride);
        const uint64_t rbase = (uint64_t)(uint32_t)state * d.stride;  // state >= 0
        const float u2 = (float)((uint32_t)r >> 8) * 0x1.0p-24f;
        if (d.srec8) {
          // 8-byte records: the row's W from W32 beside the record, and the
          // alias slot's own record when the draw takes the alias
          const uint2 s8 = d.srec8[rbase + slot];
          const float Wi = d.W32[state];
          tally_step();
          absorbed = s8.x == kStrided8Absorbing;
          uint32_t cx = s8.x;
          if (!absorbed && !(u2 < __uint_as_float(s8.y))) cx = d.srec8[rbase + ((s8.x >> 26) & 31u)].x;
          cp = (cx & 0x80000000u) | (cx & 0x03FFFFFFu);
          if (!absorbed) w = w * (Real)Wi;
        } else {
          const uint4 sr = d.srec[rbase + slot];
          tally_step();
          absorbed = sr.y == kStridedAbsorbing;
          cp = u2 < __uint_as_float(sr.x) ? sr.y : sr.z;
          if (!absorbed) w = w * (Real)__uint_as_float(sr.w);
        }
      } else if constexpr (FUSED) {
        // alias draw: slot from the high 32 bits, keep-or-alias from the low
        // 32; the record also holds both candidates' row headers
        const uint32_t slot = __umulhi((uint32_t)(r >> 32), len);
        const Rec32 sr = d.rec32[beg + slot];
        tally_step();
        const float u2 = (float)((uint32_t)r >> 8) * 0x1.0p-24f;
        const bool keep = u2 < __uint_as_float(sr.a.x);
        cp = keep ? sr.a.y : sr.a.z;
        w = w * Wrow;
        beg = keep ? sr.b.x : sr.b.y;
        len = keep ? (sr.a.w & 0xFFFFu) : (sr.a.w >> 16);
        Wrow = (Real)__uint_as_float(keep ? sr.b.z : sr.b.w);
      } else if constexpr (SAMP == SAMPLER_LEARNED_ALIAS) {
        // the learned proposal by alias: the record holds both candidates'
        // row headers and weight factors S_i / g_j
        const uint32_t slot = __umulhi((uint32_t)(r >> 32), len);
        const Rec32 sr = d.lrec32[beg + slot];
        tally_step();
        const float u2 = (float)((uint32_t)r >> 8) * 0x1.0p-24f;
        const bool keep = u2 < __uint_as_float(sr.a.x);
        cp = keep ? sr.a.y : sr.a.z;
        w = w * (Real)__uint_as_float(keep ? sr.b.z : sr.b.w);
        beg = keep ? sr.b.x : sr.b.y;
        len = keep ? (sr.a.w & 0xFFFFu) : (sr.a.w >> 16);
      } else if constexpr (SAMP == SAMPLER_LEARNED) {
        // importance-reshaped row: P = |h| g / S, weight factor S / g_j
        const uint32_t k = beg + (uint32_t)lower_bound_guess<Real>(Tables<Real>::lcdf(d) + beg, (int)len,
                                                                   Tables<Real>::u01(r), tally_step);
        w = w * Tables<Real>::lfac(d)[k];
        cp = Row<Real>::target(d, k, beg, len, Wrow);
      } else if constexpr (SAMP == SAMPLER_ALIAS) {
        cp = Row<Real>::pick(d, beg, len, r, tally_step);
        w = w * Wrow;
      } else {
        const Real wn = w * Wrow;  // W of the current row; the tally still sees w
        cp = cdf_step<Real>(d, beg, len, Wrow, Tables<Real>::u01(r), tally_step);
        w = wn;
      }
      if (absorbed) {
        end = true;  // the weight is not read again: the walk stops here
      } else {
        if (cp & 0x80000000u) w = -w;
        state = (int)(cp & 0x7FFFFFFFu);
        ++steps;
        ++my_steps;
        if (REC && steps <= a.max_len) path[steps] = state;
        end = steps >= a.history_length || (w < 0 ? -w : w) < wc;
      }
    }
    if (end) {
      alive = false;
      if (EST == EST_FORWARD) tally_add(&tally[start], (Acc)fsum);
      if (REC) {
        for (int k = (int)steps + 1; k <= a.max_len; ++k) path[k] = -1;
        a.wfinal[hid - a.first] = (double)w;
      }
    }
  }
  flush_tile();  // the wave has reconverged after the loop
  if (my_steps) atomicAdd(a.steps, (unsigned long long)my_steps);
}

// Two walks per lane on the strided alias tables (FP32, collision or
// forward estimator, no path recording): each lane carries two histories
// and issues both of their record loads before it waits on either, doubling
// the loads in flight per resident wave, so that half the waves keep as many
// walks in flight with half the history ids grabbed and not yet walked.
// Births (the same grab windows), transitions, tallies and the LDS tile are
// those of k_walk; only the order of the tally additions differs.
template <int EST, int SRC, int TILE>
__global__ __launch_bounds__(256) void k_walk2(WalkArgs a) {
  using Real = float;
  using Acc = double;  // the HBM tally
  using TAcc = typename std::conditional<TILE == kTileWalk || TILE == kTileComp, Real, double>::type;
  constexpr int J = 2;
  const DomainView& d = a.d;
  const int n = d.n;
  Acc* __restrict__ tally = a.tally64;

  int64_t hi_id = a.last;
  uint64_t total = 0;
  if (EST == EST_FORWARD) {
  } else if (SRC == SRC_RANDOM) {
    total = a.qcdf[n - 1];
    if (total == 0) return;
  } else {
    total = a.ncdf[n - 1];
    if ((int64_t)total < hi_id) hi_id = (int64_t)total;
  }
  const int64_t span = hi_id - a.first;
  if (span <= 0) return;
  const double totald = (double)total;
  const uint64_t key = a.key_dev ? *a.key_dev : a.key;
  const int qs = blockIdx.x % kQueueSlots;
  const int64_t sbeg = a.first + span * qs / kQueueSlots;
  const int64_t send = a.first + span * (qs + 1) / kQueueSlots;
  unsigned long long* ctr = a.queue + qs * kQueueStride;
  const int lane = __lane_id();
  const uint32_t stride = d.stride;

  WaveTile<TILE, TAcc> tile;
  if constexpr (TILE != kTileOff) {
    extern __shared__ __align__(16) unsigned char walk_lds[];
    tile.init(walk_lds, a, lane);
  }
  auto flush_tile = [&]() {
    if constexpr (TILE != kTileOff) tile.flush(tally, lane);
  };
  auto tally_at = [&](int st, Real wt) {
    if constexpr (TILE != kTileOff)
      tile.add(tally, st, wt);
    else
      tally_add(&tally[st], (Acc)wt);
  };

  int64_t qnext = 0, qend = 0;
  StratWindow grab;
  const int gsh = SRC == SRC_STRATIFIED && EST != EST_FORWARD ? guide_shift(total, a.guide_bits) : 0;
  // the births of the current grab, 64 positions at a time (as k_walk)
  const bool cached = !(SRC == SRC_RANDOM && EST != EST_FORWARD && a.order);
  BirthSlot c0, c1;
  int64_t gbase = 0, cblk = -1;
  bool c1ok = false;
  auto fill = [&](BirthSlot& c, int64_t blk) {  // wave-uniform
    const int64_t p = blk * 64 + lane;
    c.r = history_state(key, (uint64_t)(gbase + p));
    if (SRC == SRC_STRATIFIED && EST != EST_FORWARD && grab.window) window_pos(grab, (uint32_t)p, c.b, c.w);
  };
  bool exhausted = false;
  bool alive[J] = {false, false};
  int state[J] = {0, 0};
  Real w[J] = {0, 0}, wc[J] = {0, 0};
  int start[J] = {0, 0};
  double fsum[J] = {0.0, 0.0};
  int64_t steps[J] = {0, 0};
  uint64_t rng[J] = {0, 0};
  uint64_t my_steps = 0;

  for (;;) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      if (exhausted) break;
      uint64_t need = __ballot(!alive[j]);
      while (need != 0) {
        if (qnext >= qend) {
          flush_tile();
          const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
          unsigned long long b = 0;
          if (lane == leader) b = atomicAdd(ctr, (unsigned long long)a.batch);
          b = first_lane_u64(b);  // the leader is the first active lane
          qnext = sbeg + (int64_t)b;
          qend = qnext + a.batch < send ? qnext + a.batch : send;
          if (qnext >= send) {
            exhausted = true;
            break;
          }
          if (SRC == SRC_STRATIFIED && EST != EST_FORWARD)
            grab_stratified(a, n, gsh, totald, qnext, qend, lane, grab);
          gbase = qnext;
          cblk = -1;
          c1ok = false;
        }
        const int64_t avail = qend - qnext;
        const int cnt = __popcll(need);
        const int take = (int64_t)cnt < avail ? cnt : (int)avail;
        const int rank = __popcll(need & __lanemask_lt());
        const int64_t q = qnext - gbase;
        BirthSlot bs;
        if (cached) {
          if ((q >> 6) != cblk) {
            if (c1ok)
              c0 = c1;
            else
              fill(c0, q >> 6);
            cblk = q >> 6;
            c1ok = false;
          }
          if (q + take > (cblk + 1) * 64 && !c1ok) {
            fill(c1, cblk + 1);
            c1ok = true;
          }
          const BirthSlot& m = lane >= (int)(q & 63) ? c0 : c1;
          const int src = (int)((q + rank) & 63);
          bs.r = __shfl(m.r, src);
          if (SRC == SRC_STRATIFIED && EST != EST_FORWARD && grab.window) {
            bs.b = __shfl(m.b, src);
            bs.w = __shfl(m.w, src);
          }
        }
        const int bpos = bs.b;
        const double bw0 = bs.w;
        if (!alive[j] && rank < take) {
          int64_t hid = qnext + rank;
          if (SRC == SRC_RANDOM && EST != EST_FORWARD && a.order) {
            const uint64_t kv = a.order[hid - a.order_base];
            hid = a.order_base + (int64_t)(kv & 0xFFFFFFFFull);
            state[j] = (int)(kv >> 32);
          }
          rng[j] = cached ? bs.r : history_state(key, (uint64_t)hid);
          double w0;
          if (EST == EST_FORWARD) {
            state[j] = (int)(hid / a.per_state);
            start[j] = state[j];
            fsum[j] = 0.0;
            w0 = 1.0;
          } else if (SRC == SRC_RANDOM) {
            const uint64_t t = mul_u53(xorshift64(rng[j]), total);  // the birth draw
            if (!a.order) state[j] = birth_state(a.qcdf, a.guide, a.guide_bits, n, total, t);
            w0 = a.rhs[state[j]] * (totald / (double)a.q[state[j]]);
          } else if (grab.window) {
            state[j] = grab.slo + bpos;
            w0 = bw0;
          } else {
            state[j] = strat_birth(a, grab, totald, hid, w0);
          }
          w[j] = (Real)w0;
          wc[j] = (Real)a.cutoff * (w[j] < 0 ? -w[j] : w[j]);
          steps[j] = 0;
          alive[j] = true;
        }
        qnext += take;
        need = __ballot(!alive[j]);
      }
    }
    if (!alive[0] && !alive[1]) break;  // queue exhausted and both walks done

    // both records in flight before either is used
    uint64_t r[J] = {0, 0};
    uint4 sr[J] = {make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u)};
#pragma unroll
    for (int j = 0; j < J; ++j) {
      if (alive[j]) {
        r[j] = xorshift64(rng[j]);
        const uint32_t slot = __umulhi((uint32_t)(r[j] >> 32), stride);
        sr[j] = d.srec[(uint64_t)(uint32_t)state[j] * stride + slot];  // state >= 0
      }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      if (!alive[j]) continue;
      if (EST == EST_FORWARD)
        fsum[j] = fsum[j] + (double)w[j] * a.rhs[state[j]];
      else
        tally_at(state[j], w[j]);
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      if (!alive[j]) continue;
      bool end;
      if (sr[j].y == kStridedAbsorbing) {
        end = true;
      } else {
        const float u2 = (float)((uint32_t)r[j] >> 8) * 0x1.0p-24f;
        const uint32_t cp = u2 < __uint_as_float(sr[j].x) ? sr[j].y : sr[j].z;
        w[j] = w[j] * __uint_as_float(sr[j].w);
        if (cp & 0x80000000u) w[j] = -w[j];
        state[j] = (int)(cp & 0x7FFFFFFFu);
        ++steps[j];
        ++my_steps;
        end = steps[j] >= a.history_length || (w[j] < 0 ? -w[j] : w[j]) < wc[j];
      }
      if (end) {
        alive[j] = false;
        if (EST == EST_FORWARD) tally_add(&tally[start[j]], (Acc)fsum[j]);
      }
    }
  }
  flush_tile();
  if (my_steps) atomicAdd(a.steps, (unsigned long long)my_steps);
}

template <typename Real, int EST, int SRC, int SAMP>
hipError_t launch_walk_s(const WalkArgs& a, const WalkConfig& c, hipStream_t s) {
  if constexpr (SAMP == SAMPLER_ALIAS_STRIDED && sizeof(Real) == 4 && EST != EST_EXPECTED_VALUE) {
    if (c.walks_per_lane == 2 && !c.record) {
      const size_t lds = (size_t)(c.block / 64) * (size_t)a.tile_slots * tile_entry_bytes(c.tile, sizeof(Real));
      if constexpr (EST == EST_COLLISION) {
        if (c.tile == kTileWalk) {
          k_walk2<EST, SRC, kTileWalk><<<c.grid, c.block, lds, s>>>(a);
          return hipGetLastError();
        }
        if (c.tile == kTileComp) {
          k_walk2<EST, SRC, kTileComp><<<c.grid, c.block, lds, s>>>(a);
          return hipGetLastError();
        }
        if (c.tile != kTileOff) {
          k_walk2<EST, SRC, kTileF64><<<c.grid, c.block, lds, s>>>(a);
          return hipGetLastError();
        }
      }
      k_walk2<EST, SRC, kTileOff><<<c.grid, c.block, 0, s>>>(a);
      return hipGetLastError();
    }
  }
  if (c.record) {
    k_walk<Real, EST, SRC, SAMP, true, kTileOff><<<c.grid, c.block, 0, s>>>(a);
  } else if constexpr (EST == EST_COLLISION || EST == EST_EXPECTED_VALUE) {
    const size_t lds = (size_t)(c.block / 64) * (size_t)a.tile_slots * tile_entry_bytes(c.tile, sizeof(Real));
    if constexpr (sizeof(Real) == 4) {
      if (c.tile == kTileWalk) {
        k_walk<Real, EST, SRC, SAMP, false, kTileWalk><<<c.grid, c.block, lds, s>>>(a);
        return hipGetLastError();
      }
      if (c.tile == kTileComp) {
        k_walk<Real, EST, SRC, SAMP, false, kTileComp><<<c.grid, c.block, lds, s>>>(a);
        return hipGetLastError();
      }
    }
    if (c.tile != kTileOff) {
      k_walk<Real, EST, SRC, SAMP, false, kTileF64><<<c.grid, c.block, lds, s>>>(a);
    } else {
      k_walk<Real, EST, SRC, SAMP, false, kTileOff><<<c.grid, c.block, 0, s>>>(a);
    }
  } else {
    k_walk<Real, EST, SRC, SAMP, false, kTileOff><<<c.grid, c.block, 0, s>>>(a);
  }
  return hipGetLastError();
}

template <typename Real, int EST, int SRC>
hipError_t launch_walk_t(const WalkArgs& a, const WalkConfig& c, hipStream_t s) {
  if (c.sampler == SAMPLER_LEARNED) return launch_walk_s<Real, EST, SRC, SAMPLER_LEARNED>(a, c, s);
  if (c.sampler == SAMPLER_LEARNED_ALIAS) {
    if constexpr (sizeof(Real) == 4) return launch_walk_s<Real, EST, SRC, SAMPLER_LEARNED_ALIAS>(a, c, s);
    return hipErrorInvalidValue;  // FP32 records only; the context refuses FP64 first
  }
  if (c.sampler == SAMPLER_ALIAS_STRIDED) {
    // FP32 records, collision or forward estimator; the context refuses the rest first
    if constexpr (sizeof(Real) == 4 && EST != EST_EXPECTED_VALUE)
      return launch_walk_s<Real, EST, SRC, SAMPLER_ALIAS_STRIDED>(a, c, s);
    return hipErrorInvalidValue;
  }
  return c.sampler == SAMPLER_ALIAS ? launch_walk_s<Real, EST, SRC, SAMPLER_ALIAS>(a, c, s)
                                    : launch_walk_s<Real, EST, SRC, SAMPLER_CDF>(a, c, s);
}

template <typename Real>
hipError_t launch_walk_p(const WalkArgs& a, const WalkConfig& c, hipStream_t s) {
  if (c.estimator == EST_FORWARD) return launch_walk_t<Real, EST_FORWARD, SRC_RANDOM>(a, c, s);
  if (c.estimator == EST_EXPECTED_VALUE) {
    return c.source == SRC_STRATIFIED ? launch_walk_t<Real, EST_EXPECTED_VALUE, SRC_STRATIFIED>(a, c, s)
                                      : launch_walk_t<Real, EST_EXPECTED_VALUE, SRC_RANDOM>(a, c, s);
  }
  return c.source == SRC_STRATIFIED ? launch_walk_t<Real, EST_COLLISION, SRC_STRATIFIED>(a, c, s)
                                    : launch_walk_t<Real, EST_COLLISION, SRC_RANDOM>(a, c, s);
}

// Stream key of the next MC solve, advanced on the device so that a captured
// MCSA iteration replays with a fresh key each time:
// key = splitmix64(seed ^ splitmix64(index)), index = (*counter)++.
__global__ void k_next_key(uint64_t seed, unsigned long long* counter, uint64_t* key) {
  const unsigned long long idx = *counter;
  *counter = idx + 1;
  *key = splitmix64(seed ^ splitmix64((uint64_t)idx));
}

// ------------------------------------------------------------------ source
__global__ void k_absmax(const double* __restrict__ v, int n, unsigned long long* bmax) {
  __shared__ unsigned long long sm[kRedThreads];
  unsigned long long m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double a = fabs(v[i]);
    unsigned long long bits = (unsigned long long)__double_as_longlong(a);
    if (bits > m) m = bits;  // non-negative doubles order as integers; NaN is largest
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off && sm[threadIdx.x + off] > sm[threadIdx.x]) sm[threadIdx.x] = sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0 && sm[0]) atomicMax(bmax, sm[0]);
}

__device__ __forceinline__ int ceil_log2_u64(uint64_t v) {
  int l = 0;
  while (((uint64_t)1 << l) < v) ++l;
  return l;
}

__global__ void k_quantize(const double* __restrict__ rhs, int n, const unsigned long long* bmax_bits,
                           uint64_t* __restrict__ q) {
  const double bmax = __longlong_as_double((long long)*bmax_bits);
  const int S = 62 - ceil_log2_u64((uint64_t)(n > 0 ? n : 1));
  int e = 0;
  bool nonzero = bmax > 0.0 && bmax <= DBL_MAX;
  if (nonzero) (void)frexp(bmax, &e);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint64_t qi = 0;
    if (nonzero) qi = (uint64_t)ceil(ldexp(fabs(rhs[i]), S - e));
    q[i] = qi;
  }
}

// floor(num / Q) for num = hi:lo, quotient known to be < 2^53.
__device__ __forceinline__ uint64_t div128_small(uint64_t hi, uint64_t lo, uint64_t Q) {
  double numd = (double)hi * 0x1.0p64 + (double)lo;
  uint64_t e = (uint64_t)(numd / (double)Q);
  // correct e so that e*Q <= num < (e+1)*Q
  for (int it = 0; it < 8; ++it) {
    uint64_t plo = e * Q, phi = __umul64hi(e, Q);
    bool gt = (phi > hi) || (phi == hi && plo > lo);  // e*Q > num
    if (gt) {
      --e;
      continue;
    }
    uint64_t nlo = plo + Q, nhi = phi + (nlo < plo);
    bool le = (nhi < hi) || (nhi == hi && nlo <= lo);  // (e+1)*Q <= num
    if (le) {
      ++e;
      continue;
    }
    break;
  }
  return e;
}

__global__ void k_strat_counts(const uint64_t* __restrict__ q, int n, const uint64_t* qtotal, uint64_t nh,
                               uint64_t* __restrict__ cnt) {
  const uint64_t Q = *qtotal;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint64_t c = 0;
    if (Q) {
      uint64_t lo = q[i] * nh, hi = __umul64hi(q[i], nh);
      uint64_t lo2 = lo + (Q - 1);
      hi += (lo2 < lo);
      c = div128_small(hi, lo2, Q);
    }
    cnt[i] = c;
  }
}

// -------------------------------------------------------------------- scan
template <typename T>
__device__ T block_excl_scan(T v, T* sm, T* total) {
  // Hillis-Steele over kScanThreads lanes; exact for integers.
  const int t = threadIdx.x;
  sm[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    T add = t >= off ? sm[t - off] : T(0);
    __syncthreads();
    sm[t] += add;
    __syncthreads();
  }
  T incl = sm[t];
  *total = sm[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(const T* __restrict__ in, int64_t n,
                                                                 T* __restrict__ sums) {
  __shared__ T sm[kScanThreads];
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  T acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;
    if (i < n) acc += in[i];
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kScanThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = sm[0];
}

// exclusive scan of sums[0..m) in place, single block: each thread sums a
// contiguous chunk, the chunk totals are scanned across the block once, and
// each thread writes its chunk's running prefix (integers: exact in any order)
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(T* sums, int64_t m) {
  __shared__ T sm[kScanThreads];
  const int64_t per = (m + kScanThreads - 1) / kScanThreads;
  const int64_t b = (int64_t)threadIdx.x * per < m ? (int64_t)threadIdx.x * per : m;
  const int64_t e = b + per < m ? b + per : m;
  T s = T(0);
  for (int64_t i = b; i < e; ++i) s += sums[i];
  T tot;
  T ex = block_excl_scan<T>(s, sm, &tot);
  for (int64_t i = b; i < e; ++i) {
    const T v = sums[i];
    sums[i] = ex;
    ex += v;
  }
}

// inclusive scan of each tile plus its tile offset
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T* __restrict__ in, int64_t n,
                                                             const T* __restrict__ offs, T* out) {
  __shared__ T sm[kScanThreads];
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T local = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    v[k] = i < n ? in[i] : T(0);
    local += v[k];
  }
  T tot;
  T ex = block_excl_scan<T>(local, sm, &tot);
  T run = offs[blockIdx.x] + ex;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    run += v[k];
    if (i < n) out[i] = run;
  }
}

// --------------------------------------------------------------- fills
template <typename T>
__global__ void k_fill(T* p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

inline int grid_for(int64_t n, int block, int cap = 65536) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// --------------------------------------------------------------- state order
// A context may number the unknowns in an order of its own: internal state k
// is the caller's unknown ord[k].  Vectors (rows of w elements) cross the
// C ABI by a gather into the internal numbering or a scatter back out.
template <typename T>
__global__ void k_gather_rows(int64_t n, int w, const int* __restrict__ ord, const T* __restrict__ src,
                              T* __restrict__ dst) {
  const int64_t total = n * w;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i / w;
    dst[i] = src[(int64_t)ord[k] * w + (i - k * w)];
  }
}

template <typename T>
__global__ void k_scatter_rows(int64_t n, int w, const int* __restrict__ ord, const T* __restrict__ src,
                               T* __restrict__ dst) {
  const int64_t total = n * w;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i / w;
    dst[(int64_t)ord[k] * w + (i - k * w)] = src[i];
  }
}

// P A P^T: internal row k is the caller's row ord[k] with its columns
// renumbered through inv (inv[ord[k]] = k); entries keep their order.
template <typename V>
__global__ void k_permute_csr(int n, const int* __restrict__ ord, const int* __restrict__ inv,
                              const int* __restrict__ rp_src, const int* __restrict__ col_src,
                              const V* __restrict__ val_src, const int* __restrict__ rp_dst,
                              int* __restrict__ col_dst, V* __restrict__ val_dst) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int r = ord[k];
    const int b = rp_src[r], e = rp_src[r + 1];
    const int o = rp_dst[k] - b;
    for (int p = b; p < e; ++p) {
      col_dst[o + p] = inv[col_src[p]];
      val_dst[o + p] = val_src[p];
    }
  }
}

// --------------------------------------------------------------- domain
template <typename AV>
__device__ __forceinline__ double aval(const void* v, int k) {
  return (double)((const AV*)v)[k];
}

__device__ __forceinline__ void set_flag(int* flags, int which, int row) {
  atomicMin(&flags[which], row + 1);  // flags initialised to INT_MAX
}

template <typename AV>
__global__ void k_jacobi(int n, const int* __restrict__ rp, const int* __restrict__ col, const void* val,
                         double* __restrict__ dinv, int* flags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double dg = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == i) dg = aval<AV>(val, k);
    if (dg == 0.0) set_flag(flags, FLAG_ZERO_DIAG, i);
    dinv[i] = 1.0 / dg;
  }
}

template <typename AV>
__global__ void k_validate_csr(int n, const int* __restrict__ rp, const int* __restrict__ col, const void* val,
                               int* flags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int c = col[k];
      if (c < 0 || c >= n) set_flag(flags, FLAG_BAD_COLUMN, i);
      if (!(fabs(aval<AV>(val, k)) <= DBL_MAX)) set_flag(flags, FLAG_NONFINITE, i);
    }
  }
}

template <typename AV>
__global__ void k_count_cols(int n, const int* __restrict__ rp, const int* __restrict__ col, const void* val,
                             int* __restrict__ cnt, int* flags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int c = col[k];
      double v = aval<AV>(val, k);
      if (!(fabs(v) <= DBL_MAX)) set_flag(flags, FLAG_NONFINITE, i);
      if (c < 0 || c >= n) {
        set_flag(flags, FLAG_BAD_COLUMN, i);
        continue;
      }
      atomicAdd(&cnt[c], 1);
    }
  }
}

template <typename AV>
__global__ void k_scatter_transpose(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                    const void* val, const double* __restrict__ dinv,
                                    const int* __restrict__ t_rp, int* __restrict__ fill,
                                    int* __restrict__ t_src, double* __restrict__ t_val) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      int c = col[k];
      if (c < 0 || c >= n) continue;
      int p = t_rp[c] + atomicAdd(&fill[c], 1);
      t_src[p] = r;
      double a = aval<AV>(val, k);
      t_val[p] = dinv ? a * dinv[r] : a;  // composite (PL A)_{rc}
    }
  }
}

template <typename AV>
__global__ void k_copy_rows(int n, const int* __restrict__ rp, const int* __restrict__ col, const void* val,
                            const double* __restrict__ dinv, int* __restrict__ t_rp, int* __restrict__ t_src,
                            double* __restrict__ t_val, int* flags) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    t_rp[r] = rp[r];
    if (r == n - 1) t_rp[n] = rp[n];
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int c = col[k];
      const double a = aval<AV>(val, k);
      if (!(fabs(a) <= DBL_MAX)) set_flag(flags, FLAG_NONFINITE, r);
      if (c < 0 || c >= n) set_flag(flags, FLAG_BAD_COLUMN, r);
      t_src[k] = c;
      t_val[k] = dinv ? a * dinv[r] : a;  // composite (PL A)_{rc}
    }
  }
}

// Per domain row: order entries by source row (insertion sort; the scatter
// order is arbitrary), check duplicates, count the entries that survive the
// composite threshold and the DBL_EPSILON drop.
// One domain row by one lane (rows longer than a wave): sort the
// transposed row by source index, then count the entries that survive the
// multiply's filter (with a preconditioner) and the DBL_EPSILON drop.
__device__ void domain_row_serial(int i, const int* __restrict__ t_rp, int* __restrict__ t_src,
                                  double* __restrict__ t_val, double omega, double threshold, int has_prec,
                                  int* __restrict__ keep, int* flags) {
  const int beg = t_rp[i], end = t_rp[i + 1];
  for (int p = beg + 1; p < end; ++p) {
    int s = t_src[p];
    double v = t_val[p];
    int q = p - 1;
    while (q >= beg && t_src[q] > s) {
      t_src[q + 1] = t_src[q];
      t_val[q + 1] = t_val[q];
      --q;
    }
    t_src[q + 1] = s;
    t_val[q + 1] = v;
  }
  double thr = 0.0;
  const bool filter = has_prec && end > beg;
  if (filter) {
    double mx = t_val[beg];
    for (int p = beg + 1; p < end; ++p)
      if (t_val[p] > mx) mx = t_val[p];
    thr = mx * threshold;
  }
  int kept = 0;
  for (int p = beg; p < end; ++p) {
    if (p > beg && t_src[p] == t_src[p - 1]) set_flag(flags, FLAG_DUPLICATE, i);
    double m = t_val[p];
    if (filter && fabs(m) <= thr) continue;
    double h = (t_src[p] == i) ? 1.0 - omega * m : -omega * m;
    if (fabs(h) < DBL_EPSILON) continue;
    ++kept;
  }
  keep[i] = kept;
}

// The sorted row's entries h = delta - omega m that survive, their packed
// columns, the sequential |h| prefix sums normalised to the CDF, and W.
__device__ void domain_fill_serial(int i, const int* __restrict__ t_rp, const int* __restrict__ t_src,
                                   const double* __restrict__ t_val, double omega, double threshold, int has_prec,
                                   const int* __restrict__ d_rp, uint32_t* __restrict__ col,
                                   double* __restrict__ h64, float* __restrict__ h32, double* __restrict__ cdf64,
                                   float* __restrict__ cdf32, double* __restrict__ W64, float* __restrict__ W32) {
  const int beg = t_rp[i], end = t_rp[i + 1];
  double thr = 0.0;
  const bool filter = has_prec && end > beg;
  if (filter) {
    double mx = t_val[beg];
    for (int p = beg + 1; p < end; ++p)
      if (t_val[p] > mx) mx = t_val[p];
    thr = mx * threshold;
  }
  int out = d_rp[i];
  const int obeg = out;
  double acc = 0.0;
  for (int p = beg; p < end; ++p) {
    double m = t_val[p];
    if (filter && fabs(m) <= thr) continue;
    int j = t_src[p];
    double h = (j == i) ? 1.0 - omega * m : -omega * m;
    if (fabs(h) < DBL_EPSILON) continue;
    col[out] = (uint32_t)j | (h < 0.0 ? 0x80000000u : 0u);
    h64[out] = h;
    h32[out] = (float)h;
    acc = fabs(h) + acc;
    cdf64[out] = acc;
    ++out;
  }
  W64[i] = acc;
  W32[i] = (float)acc;
  for (int p = obeg; p < out; ++p) {
    double c = cdf64[p] / acc;
    cdf64[p] = c;
    cdf32[p] = (float)c;
  }
}

// Wave-wide maximum over the active lanes' values (-inf elsewhere).
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(v, o);
    v = t > v ? t : v;
  }
  return v;
}

// One domain row per wave (rows of <= 64 entries; longer rows go to lane 0
// and the serial code).  Lane k holds entry k; its sorted position is the
// number of smaller source indices in the row.
__global__ __launch_bounds__(256) void k_domain_rows(int n, const int* __restrict__ t_rp, int* __restrict__ t_src,
                                                     double* __restrict__ t_val, double omega, double threshold,
                                                     int has_prec, int* __restrict__ keep, int* flags) {
  const int lane = __lane_id();
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += nw) {
    const int beg = t_rp[i], z = t_rp[i + 1] - beg;
    if (z > 64) {
      if (lane == 0) domain_row_serial(i, t_rp, t_src, t_val, omega, threshold, has_prec, keep, flags);
      continue;
    }
    const bool on = lane < z;
    const int key = on ? t_src[beg + lane] : INT_MAX;
    const double v = on ? t_val[beg + lane] : 0.0;
    int rank = 0;
    bool dup = false;
    for (int k = 0; k < z; ++k) {
      const int kk = __shfl(key, k);
      rank += kk < key ? 1 : 0;
      dup = dup || (kk == key && k != lane);
    }
    if (on && dup) set_flag(flags, FLAG_DUPLICATE, i);
    if (on) {
      t_src[beg + rank] = key;
      t_val[beg + rank] = v;
    }
    const bool filter = has_prec && z > 0;
    const double thr = filter ? wave_max(on ? v : -INFINITY) * threshold : 0.0;
    const double h = key == i ? 1.0 - omega * v : -omega * v;
    const bool kept = on && !(filter && fabs(v) <= thr) && !(fabs(h) < DBL_EPSILON);
    const int cnt = __popcll(__ballot(kept));
    if (lane == 0) keep[i] = cnt;
  }
}

// One sorted row per wave: kept entries are compacted by ballot rank; each
// lane forms its CDF prefix by adding the kept |h| of lanes 0..itself in
// row order, the same additions in the same order as the serial loop.
__global__ __launch_bounds__(256) void k_domain_fill(int n, const int* __restrict__ t_rp,
                                                     const int* __restrict__ t_src,
                                                     const double* __restrict__ t_val, double omega,
                                                     double threshold, int has_prec, const int* __restrict__ d_rp,
                                                     uint32_t* __restrict__ col, double* __restrict__ h64,
                                                     float* __restrict__ h32, double* __restrict__ cdf64,
                                                     float* __restrict__ cdf32, double* __restrict__ W64,
                                                     float* __restrict__ W32) {
  const int lane = __lane_id();
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += nw) {
    const int beg = t_rp[i], z = t_rp[i + 1] - beg;
    if (z > 64) {
      if (lane == 0)
        domain_fill_serial(i, t_rp, t_src, t_val, omega, threshold, has_prec, d_rp, col, h64, h32, cdf64, cdf32,
                           W64, W32);
      continue;
    }
    const bool on = lane < z;
    const int j = on ? t_src[beg + lane] : 0;
    const double m = on ? t_val[beg + lane] : 0.0;
    const bool filter = has_prec && z > 0;
    const double thr = filter ? wave_max(on ? m : -INFINITY) * threshold : 0.0;
    const double h = j == i ? 1.0 - omega * m : -omega * m;
    const bool kept = on && !(filter && fabs(m) <= thr) && !(fabs(h) < DBL_EPSILON);
    const uint64_t km = __ballot(kept);
    const double ah = kept ? fabs(h) : 0.0;
    double pre = 0.0, tot = 0.0;
    for (int k = 0; k < z; ++k) {
      const double hk = __shfl(ah, k);
      if ((km >> k) & 1ull) {
        tot = hk + tot;
        if (k <= lane) pre = hk + pre;
      }
    }
    if (kept) {
      const int out = d_rp[i] + __popcll(km & __lanemask_lt());
      col[out] = (uint32_t)j | (h < 0.0 ? 0x80000000u : 0u);
      h64[out] = h;
      h32[out] = (float)h;
      const double c = pre / tot;
      cdf64[out] = c;
      cdf32[out] = (float)c;
    }
    if (lane == 0) {
      W64[i] = tot;
      W32[i] = (float)tot;
    }
  }
}

// Row headers and per-row alias tables (Vose's method, sequential per row in
// the same operation order as the CPU restatement, so the tables are
// bit-identical).
__global__ void k_pack_rows(int n, const int* __restrict__ d_rp, const uint32_t* __restrict__ col,
                            const double* __restrict__ h64, const double* __restrict__ W64,
                            const float* __restrict__ W32, uint4* __restrict__ hdr32, uint4* __restrict__ hdr64,
                            Rec32* __restrict__ rec32, uint4* __restrict__ slot64, int alias, int* flags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int beg = d_rp[i];
    const int z = d_rp[i + 1] - beg;
    const double W = W64[i];
    hdr32[i] = make_uint4((uint32_t)beg, (uint32_t)z, __float_as_uint(W32[i]), 0u);
    hdr64[i] = make_uint4((uint32_t)beg, (uint32_t)z, (uint32_t)__double2loint(W), (uint32_t)__double2hiint(W));
    if (!alias) continue;
    if (z > 64) {
      set_flag(flags, FLAG_LONG_ROW, i);
      continue;
    }
    double s[64];
    int small[64], large[64], al[64];
    double thr[64];
    int ns = 0, nl = 0;
    for (int k = 0; k < z; ++k) {
      s[k] = (fabs(h64[beg + k]) / W) * (double)z;
      if (s[k] < 1.0) small[ns++] = k;
      else large[nl++] = k;
    }
    while (ns > 0 && nl > 0) {
      const int l = small[--ns];
      const int g = large[--nl];
      thr[l] = s[l];
      al[l] = g;
      s[g] = (s[g] + s[l]) - 1.0;
      if (s[g] < 1.0) small[ns++] = g;
      else large[nl++] = g;
    }
    while (nl > 0) {
      const int g = large[--nl];
      thr[g] = 1.0;
      al[g] = g;
    }
    while (ns > 0) {
      const int l = small[--ns];
      thr[l] = 1.0;
      al[l] = l;
    }
    for (int k = 0; k < z; ++k) {
      const uint32_t self = col[beg + k], other = col[beg + al[k]];
      const int js = (int)(self & 0x7FFFFFFFu), jo = (int)(other & 0x7FFFFFFFu);
      const int bs = d_rp[js], bo = d_rp[jo];
      const uint32_t ls = (uint32_t)(d_rp[js + 1] - bs), lo = (uint32_t)(d_rp[jo + 1] - bo);
      Rec32 r;
      r.a = make_uint4(__float_as_uint((float)thr[k]), self, other, (ls & 0xFFFFu) | (lo << 16));
      r.b = make_uint4((uint32_t)bs, (uint32_t)bo, __float_as_uint(W32[js]), __float_as_uint(W32[jo]));
      rec32[beg + k] = r;
      slot64[beg + k] = make_uint4((uint32_t)__double2loint(thr[k]), (uint32_t)__double2hiint(thr[k]), self, other);
    }
  }
}

// Longest domain row.
__global__ void k_max_row_len(int n, const int* __restrict__ d_rp, int* out) {
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = max(m, d_rp[i + 1] - d_rp[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_down(m, off));
  if (__lane_id() == 0) atomicMax(out, m);
}

// Strided alias tables: Vose's method (as k_pack_rows) over the row padded
// with zero-probability slots to `stride` entries.  A padding slot is never
// kept (threshold 0); should rounding leave one over at the end, it aliases
// the row's first entry.  Absorbing rows get kStridedAbsorbing records.
__global__ void k_pack_strided(int n, const int* __restrict__ d_rp, const uint32_t* __restrict__ col,
                               const double* __restrict__ h64, const double* __restrict__ W64,
                               const float* __restrict__ W32, int stride, uint4* __restrict__ srec,
                               uint2* __restrict__ srec8) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int beg = d_rp[i];
    const int z = d_rp[i + 1] - beg;
    uint4* out = srec ? srec + (size_t)i * (size_t)stride : nullptr;
    uint2* out8 = srec8 ? srec8 + (size_t)i * (size_t)stride : nullptr;
    if (z == 0) {
      for (int k = 0; k < stride; ++k) {
        if (out) out[k] = make_uint4(0u, kStridedAbsorbing, kStridedAbsorbing, 0u);
        if (out8) out8[k] = make_uint2(kStrided8Absorbing, 0u);
      }
      continue;
    }
    const double W = W64[i];
    double sc[64];
    int small[64], large[64], al[64];
    double thr[64];
    int ns = 0, nl = 0;
    for (int k = 0; k < stride; ++k) {
      sc[k] = k < z ? (fabs(h64[beg + k]) / W) * (double)stride : 0.0;
      if (sc[k] < 1.0) small[ns++] = k;
      else large[nl++] = k;
    }
    while (ns > 0 && nl > 0) {
      const int l = small[--ns];
      const int g = large[--nl];
      thr[l] = sc[l];
      al[l] = g;
      sc[g] = (sc[g] + sc[l]) - 1.0;
      if (sc[g] < 1.0) small[ns++] = g;
      else large[nl++] = g;
    }
    while (nl > 0) {
      const int g = large[--nl];
      thr[g] = 1.0;
      al[g] = g;
    }
    while (ns > 0) {
      const int l = small[--ns];
      thr[l] = l < z ? 1.0 : 0.0;
      al[l] = l < z ? l : 0;
    }
    const uint32_t wbits = __float_as_uint(W32[i]);
    for (int k = 0; k < stride; ++k) {
      const uint32_t other = col[beg + al[k]];
      const uint32_t self = k < z ? col[beg + k] : other;
      if (out) out[k] = make_uint4(__float_as_uint((float)thr[k]), self, other, wbits);
      // 8-byte form: the entry's own column and sign, and its alias slot,
      // whose record holds the alias column and sign
      if (out8)
        out8[k] = make_uint2((self & 0x80000000u) | ((uint32_t)al[k] << 26) | (self & 0x03FFFFFFu),
                             __float_as_uint((float)thr[k]));
    }
  }
}

// FLAG_LONG_ROW for every domain row longer than `limit` entries.
__global__ void k_flag_long_rows(int n, const int* __restrict__ d_rp, int limit, int* flags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (d_rp[i + 1] - d_rp[i] > limit) set_flag(flags, FLAG_LONG_ROW, i);
}

// Alias records of the learned proposal: Vose's method over the reshaped
// row P_ik = |h_ik| g_k / S_i (S_i summed in row order, as k_build_proposal
// does), and for each candidate its weight factor S_i / g_j rounded to FP32,
// the value the CDF form of the proposal multiplies by.
__global__ void k_pack_proposal_alias(int n, const int* __restrict__ d_rp, const uint32_t* __restrict__ col,
                                      const double* __restrict__ h64, const float* __restrict__ g,
                                      Rec32* __restrict__ lrec32, int* flags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int beg = d_rp[i];
    const int z = d_rp[i + 1] - beg;
    double wts[64];
    double S = 0.0;
    bool bad = false;
    for (int k = 0; k < z && k < 64; ++k) {
      const double gj = (double)g[col[beg + k] & 0x7FFFFFFFu];
      if (!(gj > 0.0) || !(gj <= DBL_MAX)) bad = true;
      wts[k] = fabs(h64[beg + k]) * gj;
      S = wts[k] + S;
    }
    if (z > 64 || bad) {
      set_flag(flags, z > 64 ? FLAG_LONG_ROW : FLAG_BAD_IMPORTANCE, i);
      // the caller refuses the tables; the records are still written, as
      // uniform draws with weight factor z |h|, so that every record a
      // walk could read holds in-range row headers
      for (int k = 0; k < z; ++k) {
        const uint32_t self = col[beg + k];
        const int js = (int)(self & 0x7FFFFFFFu);
        const int bs = d_rp[js];
        const uint32_t ls = (uint32_t)(d_rp[js + 1] - bs);
        const float f = (float)(fabs(h64[beg + k]) * (double)z);
        Rec32 r;
        r.a = make_uint4(__float_as_uint(1.0f), self, self, (ls & 0xFFFFu) | (ls << 16));
        r.b = make_uint4((uint32_t)bs, (uint32_t)bs, __float_as_uint(f), __float_as_uint(f));
        lrec32[beg + k] = r;
      }
      continue;
    }
    double s[64], thr[64];
    int small[64], large[64], al[64];
    int ns = 0, nl = 0;
    for (int k = 0; k < z; ++k) {
      s[k] = (wts[k] / S) * (double)z;
      if (s[k] < 1.0) small[ns++] = k;
      else large[nl++] = k;
    }
    while (ns > 0 && nl > 0) {
      const int l = small[--ns];
      const int gg = large[--nl];
      thr[l] = s[l];
      al[l] = gg;
      s[gg] = (s[gg] + s[l]) - 1.0;
      if (s[gg] < 1.0) small[ns++] = gg;
      else large[nl++] = gg;
    }
    while (nl > 0) {
      const int gg = large[--nl];
      thr[gg] = 1.0;
      al[gg] = gg;
    }
    while (ns > 0) {
      const int l = small[--ns];
      thr[l] = 1.0;
      al[l] = l;
    }
    for (int k = 0; k < z; ++k) {
      const uint32_t self = col[beg + k], other = col[beg + al[k]];
      const int js = (int)(self & 0x7FFFFFFFu), jo = (int)(other & 0x7FFFFFFFu);
      const int bs = d_rp[js], bo = d_rp[jo];
      const uint32_t ls = (uint32_t)(d_rp[js + 1] - bs), lo = (uint32_t)(d_rp[jo + 1] - bo);
      const float fs = (float)(S / (double)g[js]), fo = (float)(S / (double)g[jo]);
      Rec32 r;
      r.a = make_uint4(__float_as_uint((float)thr[k]), self, other, (ls & 0xFFFFu) | (lo << 16));
      r.b = make_uint4((uint32_t)bs, (uint32_t)bo, __float_as_uint(fs), __float_as_uint(fo));
      lrec32[beg + k] = r;
    }
  }
}

// Target records: for entry k, the packed column and that column's row header.
__global__ void k_pack_targets(int64_t nnz, const int* __restrict__ d_rp, const uint32_t* __restrict__ col,
                               const double* __restrict__ W64, const float* __restrict__ W32,
                               const double* __restrict__ cdf64, uint4* __restrict__ tgt32,
                               Rec32* __restrict__ tgt64) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = col[k];
    const int j = (int)(c & 0x7FFFFFFFu);
    const int b = d_rp[j];
    const uint32_t l = (uint32_t)(d_rp[j + 1] - b);
    if (tgt32) tgt32[k] = make_uint4(c, (uint32_t)b, l, __float_as_uint(W32[j]));
    if (tgt64) {
      Rec32 r;
      r.a = make_uint4(c, (uint32_t)b, l, 0u);
      r.b = make_uint4((uint32_t)__double2loint(W64[j]), (uint32_t)__double2hiint(W64[j]),
                       (uint32_t)__double2loint(cdf64[k]), (uint32_t)__double2hiint(cdf64[k]));
      tgt64[k] = r;
    }
  }
}

// --------------------------------------------- importance features / targets
// Per-block partial sums of |a_jj| and |v_j| (grid-stride, fixed grid).
template <typename AV>
__global__ __launch_bounds__(kRedThreads) void k_feature_sums(int n, const int* __restrict__ rp,
                                                              const int* __restrict__ col, const void* val,
                                                              const double* __restrict__ v, double* part_d,
                                                              double* part_v) {
  __shared__ double sd[kRedThreads], sv[kRedThreads];
  const AV* __restrict__ a = (const AV*)val;
  double accd = 0.0, accv = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == i) accd = accd + fabs((double)a[k]);
    accv = accv + fabs(v[i]);
  }
  sd[threadIdx.x] = accd;
  sv[threadIdx.x] = accv;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sd[threadIdx.x] = sd[threadIdx.x] + sd[threadIdx.x + o];
      sv[threadIdx.x] = sv[threadIdx.x] + sv[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_d[blockIdx.x] = sd[0];
    part_v[blockIdx.x] = sv[0];
  }
}

// Eight features per state j from row j of A and the MC source v:
//   f0..f3: sums of -a_jk / a_jj over off-diagonal columns k < j-1, k = j-1,
//           k = j+1, k > j+1 (the S, W, E, N couplings of a 2-D 5-point row)
//   f4: sum_k |a_jk| / |a_jj| over k != j (the row sum of |H|)
//   f5: log(|a_jj| / mean |a_ii|)
//   f6: log(|v_j| / mean |v| + 1e-3)
//   f7: 0
template <typename AV>
__global__ void k_features(int n, const int* __restrict__ rp, const int* __restrict__ col, const void* val,
                           const double* __restrict__ v, const double* __restrict__ sums, float* __restrict__ X) {
  const AV* __restrict__ a = (const AV*)val;
  const double mean_d = sums[0] / (double)n, mean_v = sums[1] / (double)n;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    double d = 0.0;
    for (int k = rp[j]; k < rp[j + 1]; ++k)
      if (col[k] == j) d = (double)a[k];
    double f[4] = {0.0, 0.0, 0.0, 0.0}, rs = 0.0;
    const double inv = d != 0.0 ? 1.0 / d : 0.0;
    for (int k = rp[j]; k < rp[j + 1]; ++k) {
      const int c = col[k];
      if (c == j) continue;
      const double t = -(double)a[k] * inv;
      const int cls = c < j - 1 ? 0 : (c == j - 1 ? 1 : (c == j + 1 ? 2 : 3));
      f[cls] += t;
      rs += fabs(t);
    }
    float* x = X + (size_t)j * 8;
    x[0] = (float)f[0];
    x[1] = (float)f[1];
    x[2] = (float)f[2];
    x[3] = (float)f[3];
    x[4] = (float)rs;
    x[5] = (float)(mean_d > 0.0 && d != 0.0 ? log(fabs(d) / mean_d) : 0.0);
    x[6] = (float)log((mean_v > 0.0 ? fabs(v[j]) / mean_v : 0.0) + 1e-3);
    x[7] = 0.f;
  }
}

// y_j = a_jj x_j
template <typename AV>
__global__ void k_scale_by_diag(int n, const int* __restrict__ rp, const int* __restrict__ col, const void* val,
                                const double* __restrict__ x, double* __restrict__ y) {
  const AV* __restrict__ a = (const AV*)val;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    double d = 0.0;
    for (int k = rp[j]; k < rp[j + 1]; ++k)
      if (col[k] == j) d = (double)a[k];
    y[j] = d * x[j];
  }
}

// g_j = |a_jj| (0 when the row has no diagonal, which the proposal builder flags)
template <typename AV>
__global__ void k_diag_abs(int n, const int* __restrict__ rp, const int* __restrict__ col, const void* val,
                           float* __restrict__ g) {
  const AV* __restrict__ a = (const AV*)val;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    double d = 0.0;
    for (int k = rp[j]; k < rp[j + 1]; ++k)
      if (col[k] == j) d = (double)a[k];
    g[j] = (float)fabs(d);
  }
}

// g_j = exp(clip(o_j, -4, 4)); states switched off by the mask use g = 1.
__global__ void k_importance(int n, const float* __restrict__ o, const signed char* __restrict__ mask,
                             float* __restrict__ g) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const float c = fminf(fmaxf(o[j], -4.f), 4.f);
    g[j] = (mask && !mask[j]) ? 1.f : expf(c);
  }
}

// Training target from an adjoint tally x: t_j = log(clip(|x_j| / mean|x|, e^-4, e^4)).
__global__ void k_targets(int n, const double* __restrict__ x, const double* __restrict__ sums,
                          float* __restrict__ T) {
  const double mean = sums[1] / (double)n;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    double r = mean > 0.0 ? fabs(x[j]) / mean : 1.0;
    r = fmin(fmax(r, exp(-4.0)), exp(4.0));
    T[j] = (float)log(r);
  }
}

// Learned-proposal tables, per row sequentially in the same operation order
// as the CPU restatement: acc += |h| g_j; lcdf = acc_k / S; lfac = S / g_j.
__global__ void k_build_proposal(int n, const int* __restrict__ d_rp, const uint32_t* __restrict__ col,
                                 const double* __restrict__ h64, const float* __restrict__ g,
                                 double* __restrict__ lcdf64, float* __restrict__ lcdf32,
                                 double* __restrict__ lfac64, float* __restrict__ lfac32, int* flags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int beg = d_rp[i], end = d_rp[i + 1];
    double acc = 0.0;
    bool bad = false;
    for (int p = beg; p < end; ++p) {
      const double gj = (double)g[col[p] & 0x7FFFFFFFu];
      if (!(gj > 0.0) || !(gj <= DBL_MAX)) bad = true;
      acc = fabs(h64[p]) * gj + acc;
      lcdf64[p] = acc;
    }
    if (bad) {
      set_flag(flags, FLAG_BAD_IMPORTANCE, i);
      continue;
    }
    for (int p = beg; p < end; ++p) {
      const double c = lcdf64[p] / acc;
      const double f = acc / (double)g[col[p] & 0x7FFFFFFFu];
      lcdf64[p] = c;
      lcdf32[p] = (float)c;
      lfac64[p] = f;
      lfac32[p] = (float)f;
    }
  }
}

// ---------------------------------------------------------------- linalg
// r = PL (b - A x) with V lanes per row ("vector CSR"): the lanes of a row
// read its entries with unit stride, so a wave's loads are coalesced; the
// row sum is reduced by a fixed xor butterfly (deterministic), accumulated
// in FP64 for FP32 or FP64 values.  V = 1 is one row per lane with the
// entries summed in stored order.  Grid-stride over rows with a fixed grid,
// so the per-block partial sums of r^2 are reproducible.
template <typename AV, int V>
__global__ __launch_bounds__(kRedThreads) void k_prec_residual(int n, const int* __restrict__ rp,
                                                               const int* __restrict__ col, const void* val,
                                                               const double* __restrict__ dinv,
                                                               const double* __restrict__ b,
                                                               const double* __restrict__ x,
                                                               double* __restrict__ r, double* part) {
  __shared__ double sm[kRedThreads];
  const AV* __restrict__ v = (const AV*)val;
  constexpr int RPB = kRedThreads / V;
  const int g = threadIdx.x / V, j = threadIdx.x % V;
  double sq = 0.0;
  for (int i = blockIdx.x * RPB + g; i < n; i += gridDim.x * RPB) {
    double s = 0.0;
    const int end = rp[i + 1];
    int k = rp[i] + j;
    // the lane's first U entries: their loads are issued together
    constexpr int U = 4;
    double prod[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kk = k + u * V;
      prod[u] = kk < end ? (double)v[kk] * x[col[kk]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s = s + prod[u];
    for (k += U * V; k < end; k += V) s = s + (double)v[k] * x[col[k]];
    if (V > 1)
      for (int off = V / 2; off > 0; off >>= 1) s = s + __shfl_xor(s, off, V);
    if (j == 0) {
      const double t = b[i] - s;
      const double ri = dinv ? dinv[i] * t : t;
      r[i] = ri;
      sq = sq + ri * ri;
    }
  }
  sm[threadIdx.x] = sq;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// ------------------------------------------------ sliced-ELL residual
// A copy of A for the residual: slices of 64 consecutive rows (one wave),
// slice t of width w_t = its longest row, stored entry-major from off[t]:
// entry k of the slice's lane l at off[t] + 64 k + l, column -1 (value 0)
// past the row's end.  A wave then reads 256 contiguous bytes of columns and
// of FP32 values per entry, and sums each row left to right, so r is the
// same sum as the CSR one in stored order.
__global__ void k_sell_widths(int n, const int* __restrict__ rp, int* __restrict__ width) {
  const int t = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int i = t * kWave + lane;
  int len = i < n ? rp[i + 1] - rp[i] : 0;
  for (int off = kWave / 2; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off));
  if (lane == 0 && t * kWave < n) width[t] = len;
}

template <typename AV>
__global__ void k_sell_fill(int n, int nslices, const int* __restrict__ rp, const int* __restrict__ col,
                            const AV* __restrict__ val, const int* __restrict__ width,
                            const int64_t* __restrict__ off, int* __restrict__ scol, AV* __restrict__ sval) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = (int)(g / kWave), lane = (int)(g % kWave);
  if (t >= nslices) return;
  const int i = t * kWave + lane;
  const int beg = i < n ? rp[i] : 0, len = i < n ? rp[i + 1] - beg : 0;
  const int w = width[t];
  int* c = scol + off[t] + lane;
  AV* v = sval + off[t] + lane;
  for (int k = 0; k < w; ++k) {
    c[(int64_t)k * kWave] = k < len ? col[beg + k] : -1;
    v[(int64_t)k * kWave] = k < len ? val[beg + k] : (AV)0;
  }
}

// r = PL (b - A x) over the sliced copy, one lane per row; entries in
// rounds of U whose column and value loads are issued together
// (non-temporal: they are read once), then the U gathers of x.  The grid is
// a multiple of 8 and block b works on the eighth of the slices numbered
// b mod 8 -- the XCD that round-robin dispatch gives it -- so each XCD's L2
// gathers x over one contiguous eighth of the rows and its halo instead of
// the whole vector.  Fixed grid, per-block partial sums of r^2.
template <typename AV, int U>
__global__ __launch_bounds__(kRedThreads) void k_sell_residual(int n, int nslices, const int* __restrict__ width,
                                                               const int64_t* __restrict__ off,
                                                               const int* __restrict__ scol,
                                                               const AV* __restrict__ sval,
                                                               const double* __restrict__ dinv,
                                                               const double* __restrict__ b,
                                                               const double* __restrict__ x,
                                                               double* __restrict__ r, double* part) {
  __shared__ double sm[kRedThreads];
  constexpr int WPB = kRedThreads / kWave;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int xcd = blockIdx.x % 8, per_xcd = gridDim.x / 8, local = blockIdx.x / 8;
  const int span = (nslices + 7) / 8;
  const int t0 = xcd * span, t1 = min(nslices, t0 + span);
  const int tstep = per_xcd * WPB;
  double sq = 0.0;
  int t = t0 + local * WPB + wave;
  if (t < t1) {
    // One software pipeline over all of the wave's rounds: round j's
    // gathers are issued before round j + 1's column and value loads, also
    // where round j + 1 opens the wave's next slice, so the stream of A
    // never waits for the gathers' answers.  The next slice's width and
    // offset are read a slice ahead.  All control flow is wave-uniform.
    int w = width[t];
    int64_t o = off[t];
    int tn = t + tstep;
    int wn = tn < t1 ? width[tn] : 0;
    int64_t on = tn < t1 ? off[tn] : 0;
    // b and PL of a slice's rows are loaded with its first round, so that
    // finishing the slice waits for nothing issued after them; they and r
    // stream past the L2 (non-temporal), which keeps it for the gathers of x
    int i = t * kWave + lane;
    double bi = i < n ? __builtin_nontemporal_load(b + i) : 0.0;
    double di = dinv && i < n ? __builtin_nontemporal_load(dinv + i) : 1.0;
    int cc[U];
    AV vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t e = o + (int64_t)u * kWave + lane;
      cc[u] = u < w ? __builtin_nontemporal_load(scol + e) : -1;
      vv[u] = u < w ? __builtin_nontemporal_load(sval + e) : (AV)0;
    }
    int k = 0;
    double s = 0.0;
    while (true) {
      double xx[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xx[u] = cc[u] >= 0 ? x[cc[u]] : 0.0;
      const bool last = k + U >= w;  // this round ends the slice
      const int k2 = last ? 0 : k + U;
      const int w2 = last ? (tn < t1 ? wn : 0) : w;
      const int64_t o2 = last ? on : o;
      int cn[U];
      AV vn[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kk = k2 + u;
        const int64_t e = o2 + (int64_t)kk * kWave + lane;
        cn[u] = kk < w2 ? __builtin_nontemporal_load(scol + e) : -1;
        vn[u] = kk < w2 ? __builtin_nontemporal_load(sval + e) : (AV)0;
      }
      const int in = tn * kWave + lane;
      double bn = 0.0, dn = 1.0;
      if (last && tn < t1 && in < n) {
        bn = __builtin_nontemporal_load(b + in);
        if (dinv) dn = __builtin_nontemporal_load(dinv + in);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (cc[u] >= 0) s = __dadd_rn(s, __dmul_rn((double)vv[u], xx[u]));
      if (last) {
        if (i < n) {
          const double d = __dadd_rn(bi, -s);
          const double ri = dinv ? __dmul_rn(di, d) : d;
          __builtin_nontemporal_store(ri, r + i);
          sq = sq + ri * ri;
        }
        s = 0.0;
        if (tn >= t1) break;
        i = in;
        bi = bn;
        di = dn;
        t = tn;
        w = wn;
        o = on;
        tn = t + tstep;
        wn = tn < t1 ? width[tn] : 0;
        on = tn < t1 ? off[tn] : 0;
      }
      k = k2;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cc[u] = cn[u];
        vv[u] = vn[u];
      }
    }
  }
  sm[threadIdx.x] = sq;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(kRedThreads) void k_scaled_sq(int n, const double* __restrict__ dinv,
                                                           const double* __restrict__ b, double* part) {
  __shared__ double sm[kRedThreads];
  double sq = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double t = dinv ? dinv[i] * b[i] : b[i];
    sq = sq + t * t;
  }
  sm[threadIdx.x] = sq;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// fixed-order reduction of m partials by one block
__global__ __launch_bounds__(kRedThreads) void k_sum_partials(const double* __restrict__ part, int m,
                                                              double* out) {
  __shared__ double sm[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < m; i += kRedThreads) acc = acc + part[i];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

// ------------------------------------------------- convergence criteria
// y = M (x / sqrt(*nrm2_prev)) for the domain operator M (rows of H^T):
// mode 0: M = H, 1: |H|, 2: H* with H*_ij = |h_ij| W_i; per-block partial
// sums of y^2 into part (grid-stride, fixed grid: deterministic).
__global__ __launch_bounds__(kRedThreads) void k_domain_power(int n, const int* __restrict__ d_rp,
                                                              const uint32_t* __restrict__ col,
                                                              const double* __restrict__ h64,
                                                              const double* __restrict__ W64, int mode,
                                                              const double* __restrict__ x,
                                                              const double* __restrict__ nrm2_prev,
                                                              double* __restrict__ y, double* __restrict__ part) {
  __shared__ double sm[kRedThreads];
  const double p = nrm2_prev ? *nrm2_prev : 1.0;
  const double scale = p > 0.0 ? 1.0 / sqrt(p) : 1.0;
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double v = 0.0;
    const double wi = mode == 2 ? W64[i] : 1.0;
    for (int k = d_rp[i]; k < d_rp[i + 1]; ++k) {
      const double h = h64[k];
      const double m = mode == 0 ? h : fabs(h) * wi;
      v = v + m * x[col[k] & 0x7FFFFFFFu];
    }
    v = v * scale;
    y[i] = v;
    acc = acc + v * v;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// *logsum += log(sqrt(*nrm2)) (one thread)
__global__ void k_log_accum(const double* nrm2, double* logsum) {
  const double v = *nrm2;
  *logsum = *logsum + (v > 0.0 ? 0.5 * log(v) : -INFINITY);
}

// Deterministic positive start vector x_i in [0.5, 1.5).
__global__ void k_power_start(int n, uint64_t seed, double* x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] = 0.5 + u01_f64(splitmix64(seed ^ splitmix64((uint64_t)i)));
}

// Column maxima of |h| (bits of non-negative doubles order as integers).
__global__ void k_col_absmax(int n, const int* __restrict__ d_rp, const uint32_t* __restrict__ col,
                             const double* __restrict__ h64, unsigned long long* __restrict__ cmax) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int k = d_rp[i]; k < d_rp[i + 1]; ++k)
      atomicMax(&cmax[col[k] & 0x7FFFFFFFu], (unsigned long long)__double_as_longlong(fabs(h64[k])));
}

// Per non-empty row i: ratio_i = max_k |h_ik| / colmax_i; out[0] = max W,
// out[1] = min W, out[2] = min ratio, out[3] = max ratio, out[4] = sum of
// ratios, out[5] = rows counted (block results combined by atomics on the
// bits of non-negative doubles for min/max; sums per block into part).
__global__ __launch_bounds__(kRedThreads) void k_balance(int n, const int* __restrict__ d_rp,
                                                         const double* __restrict__ h64,
                                                         const double* __restrict__ W64,
                                                         const unsigned long long* __restrict__ cmax,
                                                         unsigned long long* __restrict__ mm,
                                                         double* __restrict__ part) {
  __shared__ double ssum[kRedThreads], scnt[kRedThreads];
  double sum = 0.0, cnt = 0.0;
  unsigned long long wmax = 0, wmin = ~0ULL, rmin = ~0ULL, rmax = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned long long wb = (unsigned long long)__double_as_longlong(W64[i]);
    if (wb > wmax) wmax = wb;
    if (wb < wmin) wmin = wb;
    if (d_rp[i + 1] == d_rp[i]) continue;
    double rm = 0.0;
    for (int k = d_rp[i]; k < d_rp[i + 1]; ++k) rm = fmax(rm, fabs(h64[k]));
    const double cm = __longlong_as_double((long long)cmax[i]);
    const double r = rm / cm;
    const unsigned long long rb = (unsigned long long)__double_as_longlong(r);
    if (rb < rmin) rmin = rb;
    if (rb > rmax) rmax = rb;
    sum = sum + r;
    cnt = cnt + 1.0;
  }
  if (wmax) atomicMax(&mm[0], wmax);
  atomicMin(&mm[1], wmin);
  atomicMin(&mm[2], rmin);
  if (rmax) atomicMax(&mm[3], rmax);
  ssum[threadIdx.x] = sum;
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + off];
      scnt[threadIdx.x] = scnt[threadIdx.x] + scnt[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ssum[0];
    part[kResidualPartials + blockIdx.x] = scnt[0];
  }
}

// ------------------------------------------------------- batched systems
// One block per system s (rows [off[s], off[s+1])): out[s] = sum_i (d_i v_i)^2
// (d = dinv or 1) when square, else sum_i |v_i|; fixed-order block tree.
template <bool SQUARE>
__global__ __launch_bounds__(kRedThreads) void k_segment_norm(int nsys, const int64_t* __restrict__ off,
                                                              const double* __restrict__ v,
                                                              const double* __restrict__ dinv,
                                                              double* __restrict__ out) {
  __shared__ double sm[kRedThreads];
  for (int sys = blockIdx.x; sys < nsys; sys += gridDim.x) {
    double acc = 0.0;
    for (int64_t i = off[sys] + threadIdx.x; i < off[sys + 1]; i += kRedThreads) {
      const double t = dinv ? dinv[i] * v[i] : v[i];
      acc = SQUARE ? acc + t * t : acc + fabs(t);
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[sys] = sm[0];
    __syncthreads();
  }
}

// Per-system convergence test ||rp_s||_2 <= tol ||PL b_s||_2 (homogeneous
// systems use 1, src/MCLS_MCSASolverManager_impl.hpp:312-318); conv[s] set
// once converged; *unconverged = number of systems still iterating.
__global__ __launch_bounds__(kRedThreads) void k_batch_check(int nsys, const double* __restrict__ r2,
                                                             const double* __restrict__ b2, double tol,
                                                             int* __restrict__ conv, int* unconverged) {
  __shared__ int sm[kRedThreads];
  int cnt = 0;
  for (int s = threadIdx.x; s < nsys; s += kRedThreads) {
    double bn = sqrt(b2[s]);
    if (fabs(bn) < 10.0 * DBL_EPSILON) bn = 1.0;
    const bool ok = sqrt(r2[s]) <= tol * bn;
    if (ok) conv[s] = 1;
    cnt += conv[s] ? 0 : 1;
  }
  sm[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *unconverged = sm[0];
}

// MC source of a batch: rhs_i = omega rp_i, zero for converged systems;
// mag_i = |rhs_i| n_s / ||rhs_s||_1 sets the birth allocation (histories
// per system proportional to its unknown count; any allocation is unbiased
// because the birth weight divides by the sampled probability).
__global__ void k_batch_source(int n, const int* __restrict__ sys_of, const int64_t* __restrict__ off,
                               const double* __restrict__ r, double omega, const int* __restrict__ conv,
                               const double* __restrict__ l1, double* __restrict__ rhs, double* __restrict__ mag) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int s = sys_of[i];
    double v = conv[s] ? 0.0 : r[i] * omega;
    rhs[i] = v;
    const double norm1 = l1[s] * fabs(omega);
    mag[i] = (v != 0.0 && norm1 > 0.0) ? fabs(v) * ((double)(off[s + 1] - off[s]) / norm1) : 0.0;
  }
}

__global__ void k_axpy(int n, double a, const double* __restrict__ y, double* __restrict__ x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = x[i] + a * y[i];
}

__global__ void k_scale_copy(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = x[i] * a;
}

__global__ void k_diag_scale(int n, const double* __restrict__ d, const double* __restrict__ x,
                             double* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = d[i] * x[i];
}

template <bool ACCUM>
__global__ void k_tally_out(int n, double* __restrict__ x, double* __restrict__ t64, float* __restrict__ t32,
                            const double* __restrict__ rhs, int add_rhs, double inv,
                            const uint64_t* mtotal) {
  if (mtotal) {
    uint64_t M = *mtotal;
    inv = M ? 1.0 / (double)M : 0.0;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double t;
    if (t64) {
      t = t64[i];
      t64[i] = 0.0;
    } else {
      t = (double)t32[i];
      t32[i] = 0.0f;
    }
    double dx = t * inv;
    if (add_rhs) dx = dx + rhs[i];
    x[i] = ACCUM ? x[i] + dx : dx;
  }
}

// --------------------------------------------------- composite operator
// Composite operators with general left / right preconditioners
// (src/MCLS_LinearProblem_impl.hpp:139-221): sorted transposes, row-by-row
// sparse products and the per-row filter of MT::multiply.  One lane per
// row: these run once per operator, not per MCSA iteration.

// Sort each row's (col, val) pairs by column; a repeated column sets
// FLAG_DUPLICATE.
__global__ void k_sort_rows(int n, const int* __restrict__ rp, int* __restrict__ col, double* __restrict__ val,
                            int* flags) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int beg = rp[i], end = rp[i + 1];
    for (int p = beg + 1; p < end; ++p) {
      const int c = col[p];
      const double v = val[p];
      int q = p - 1;
      while (q >= beg && col[q] > c) {
        col[q + 1] = col[q];
        val[q + 1] = val[q];
        --q;
      }
      col[q + 1] = c;
      val[q + 1] = v;
    }
    for (int p = beg + 1; p < end; ++p)
      if (col[p] == col[p - 1]) set_flag(flags, FLAG_DUPLICATE, i);
  }
}

// Row i of C = X Y (rows of X and Y ascending): a merge of the rows Y_k for
// the entries k of X's row, smallest column first.  C_ij is the first
// product X_ik Y_kj in the order of X's row, plus each later one in turn
// (the Gustavson order of a row-by-row product).  head (nnz(X) entries) is
// each list's cursor.  FILL = false counts the row's entries into cnt.
template <bool FILL>
__global__ void k_spgemm(int n, const int* __restrict__ x_rp, const int* __restrict__ x_col,
                         const double* __restrict__ x_val, const int* __restrict__ y_rp,
                         const int* __restrict__ y_col, const double* __restrict__ y_val, int* __restrict__ head,
                         int* __restrict__ cnt, const int* __restrict__ c_rp, int* __restrict__ c_col,
                         double* __restrict__ c_val) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int xb = x_rp[i], xe = x_rp[i + 1];
    for (int p = xb; p < xe; ++p) head[p] = y_rp[x_col[p]];
    int out = FILL ? c_rp[i] : 0;
    for (;;) {
      int jmin = INT_MAX;
      for (int p = xb; p < xe; ++p) {
        const int h = head[p];
        if (h < y_rp[x_col[p] + 1]) {
          const int j = y_col[h];
          jmin = j < jmin ? j : jmin;
        }
      }
      if (jmin == INT_MAX) break;
      double sum = 0.0;
      bool first = true;
      for (int p = xb; p < xe; ++p) {
        const int h = head[p];
        if (h < y_rp[x_col[p] + 1] && y_col[h] == jmin) {
          if (FILL) {
            const double prod = x_val[p] * y_val[h];
            sum = first ? prod : sum + prod;
          }
          first = false;
          head[p] = h + 1;
        }
      }
      if (FILL) {
        c_col[out] = jmin;
        c_val[out] = sum;
      }
      ++out;
    }
    if (!FILL) cnt[i] = out;
  }
}

// TpetraMatrixHelpers::filter (tpetra/src/MCLS_TpetraHelpers.hpp:196-263):
// per row, drop |v| <= (signed row max) * threshold and exact zeros.
template <bool FILL>
__global__ void k_filter_rows(int n, const int* __restrict__ rp, const int* __restrict__ col,
                              const double* __restrict__ val, double threshold, int* __restrict__ keep,
                              const int* __restrict__ out_rp, int* __restrict__ out_col,
                              double* __restrict__ out_val) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int beg = rp[i], end = rp[i + 1];
    double thr = 0.0;
    if (end > beg) {
      double mx = val[beg];
      for (int p = beg + 1; p < end; ++p)
        if (val[p] > mx) mx = val[p];
      thr = mx * threshold;
    }
    int out = FILL ? out_rp[i] : 0;
    for (int p = beg; p < end; ++p) {
      const double v = val[p];
      if (fabs(v) <= thr || v == 0.0) continue;
      if (FILL) {
        out_col[out] = col[p];
        out_val[out] = v;
      }
      ++out;
    }
    if (!FILL) keep[i] = out;
  }
}

// y = P x for an FP64 CSR matrix, each row summed left to right.
__global__ void k_csr_spmv(int n, const int* __restrict__ rp, const int* __restrict__ col,
                           const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) acc = acc + val[k] * x[col[k]];
    y[i] = acc;
  }
}

}  // namespace

// ============================================================== launchers

hipError_t launch_walk(const WalkArgs& a, const WalkConfig& c, hipStream_t s) {
  k_fill<unsigned long long><<<1, kQueueWords, 0, s>>>(a.queue, kQueueWords, 0ULL);
  return c.precision == 32 ? launch_walk_p<float>(a, c, s) : launch_walk_p<double>(a, c, s);
}

hipError_t launch_feature_sums(int n, const int* rp, const int* col, const void* val, bool f32, const double* v,
                               double* part, double* sums, hipStream_t s) {
  const int g = residual_blocks(n, 1);
  if (f32)
    k_feature_sums<float><<<g, kRedThreads, 0, s>>>(n, rp, col, val, v, part, part + kResidualPartials);
  else
    k_feature_sums<double><<<g, kRedThreads, 0, s>>>(n, rp, col, val, v, part, part + kResidualPartials);
  k_sum_partials<<<1, kRedThreads, 0, s>>>(part, g, sums);
  k_sum_partials<<<1, kRedThreads, 0, s>>>(part + kResidualPartials, g, sums + 1);
  return hipGetLastError();
}

hipError_t launch_features(int n, const int* rp, const int* col, const void* val, bool f32, const double* v,
                           const double* sums, float* X, hipStream_t s) {
  if (f32)
    k_features<float><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, v, sums, X);
  else
    k_features<double><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, v, sums, X);
  return hipGetLastError();
}

hipError_t launch_scale_by_diag(int n, const int* rp, const int* col, const void* val, bool f32, const double* x,
                                double* y, hipStream_t s) {
  if (f32)
    k_scale_by_diag<float><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, x, y);
  else
    k_scale_by_diag<double><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, x, y);
  return hipGetLastError();
}

hipError_t launch_diag_abs(int n, const int* rp, const int* col, const void* val, bool f32, float* g,
                           hipStream_t s) {
  if (f32)
    k_diag_abs<float><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, g);
  else
    k_diag_abs<double><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, g);
  return hipGetLastError();
}

hipError_t launch_importance(int n, const float* o, const signed char* mask, float* g, hipStream_t s) {
  k_importance<<<grid_for(n, 256), 256, 0, s>>>(n, o, mask, g);
  return hipGetLastError();
}

hipError_t launch_targets(int n, const double* x, const double* sums, float* T, hipStream_t s) {
  k_targets<<<grid_for(n, 256), 256, 0, s>>>(n, x, sums, T);
  return hipGetLastError();
}

hipError_t launch_build_proposal(int n, const int* d_rp, const uint32_t* col, const double* h64, const float* g,
                                 double* lcdf64, float* lcdf32, double* lfac64, float* lfac32, int* flags,
                                 hipStream_t s) {
  k_build_proposal<<<grid_for(n, 256), 256, 0, s>>>(n, d_rp, col, h64, g, lcdf64, lcdf32, lfac64, lfac32, flags);
  return hipGetLastError();
}

hipError_t launch_pack_targets(int64_t nnz, const int* d_rp, const uint32_t* col, const double* W64,
                               const float* W32, const double* cdf64, uint4* tgt32, Rec32* tgt64,
                               hipStream_t s) {
  if (nnz <= 0) return hipSuccess;
  k_pack_targets<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, d_rp, col, W64, W32, cdf64, tgt32, tgt64);
  return hipGetLastError();
}

hipError_t launch_pack_rows(int n, const int* d_rp, const uint32_t* col, const double* h64, const double* W64,
                            const float* W32, uint4* hdr32, uint4* hdr64, Rec32* rec32, uint4* slot64,
                            bool alias, int* flags, hipStream_t s) {
  k_pack_rows<<<grid_for(n, 256), 256, 0, s>>>(n, d_rp, col, h64, W64, W32, hdr32, hdr64, rec32, slot64,
                                                alias ? 1 : 0, flags);
  return hipGetLastError();
}

hipError_t launch_flag_long_rows(int n, const int* d_rp, int limit, int* flags, hipStream_t s) {
  k_flag_long_rows<<<grid_for(n, 256), 256, 0, s>>>(n, d_rp, limit, flags);
  return hipGetLastError();
}

hipError_t launch_pack_proposal_alias(int n, const int* d_rp, const uint32_t* col, const double* h64, const float* g,
                                      Rec32* lrec32, int* flags, hipStream_t s) {
  k_pack_proposal_alias<<<grid_for(n, 256), 256, 0, s>>>(n, d_rp, col, h64, g, lrec32, flags);
  return hipGetLastError();
}

hipError_t launch_next_key(uint64_t seed, unsigned long long* counter, uint64_t* key, hipStream_t s) {
  k_next_key<<<1, 1, 0, s>>>(seed, counter, key);
  return hipGetLastError();
}

hipError_t launch_absmax(const double* v, int n, unsigned long long* bmax, hipStream_t s) {
  k_absmax<<<grid_for(n, kRedThreads, 2048), kRedThreads, 0, s>>>(v, n, bmax);
  return hipGetLastError();
}

hipError_t launch_quantize(const double* rhs, int n, const unsigned long long* bmax, uint64_t* q,
                           hipStream_t s) {
  k_quantize<<<grid_for(n, 256), 256, 0, s>>>(rhs, n, bmax, q);
  return hipGetLastError();
}

hipError_t launch_strat_counts(const uint64_t* q, int n, const uint64_t* qtotal, uint64_t nh, uint64_t* cnt,
                               hipStream_t s) {
  k_strat_counts<<<grid_for(n, 256), 256, 0, s>>>(q, n, qtotal, nh, cnt);
  return hipGetLastError();
}

size_t scan_tmp_elems(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

template <typename T>
static hipError_t scan_incl(const T* in, T* out, int64_t n, T* tmp, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  k_scan_tile_sums<T><<<(unsigned)tiles, kScanThreads, 0, s>>>(in, n, tmp);
  k_scan_sums<T><<<1, kScanThreads, 0, s>>>(tmp, tiles);
  k_scan_apply<T><<<(unsigned)tiles, kScanThreads, 0, s>>>(in, n, tmp, out);
  return hipGetLastError();
}

hipError_t launch_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, uint64_t* tmp, hipStream_t s) {
  return scan_incl<unsigned long long>((const unsigned long long*)in, (unsigned long long*)out, n,
                                       (unsigned long long*)tmp, s);
}

hipError_t launch_scan_i32_excl(const int* in, int* out, int64_t n, int* tmp, hipStream_t s) {
  k_fill<int><<<1, 1, 0, s>>>(out, 1, 0);
  return scan_incl<int>(in, out + 1, n, tmp, s);
}

hipError_t launch_max_row_len(int n, const int* d_rp, int* out, hipStream_t s) {
  k_max_row_len<<<grid_for(n, 256, 4096), 256, 0, s>>>(n, d_rp, out);
  return hipGetLastError();
}
hipError_t launch_pack_strided(int n, const int* d_rp, const uint32_t* col, const double* h64, const double* W64,
                               const float* W32, int stride, uint4* srec, uint2* srec8, hipStream_t s) {
  if (stride < 1 || stride > 64 || (!srec && !srec8)) return hipErrorInvalidValue;
  if (srec8 && (stride > 32 || n >= (1 << 26) - 1)) return hipErrorInvalidValue;
  k_pack_strided<<<grid_for(n, 256), 256, 0, s>>>(n, d_rp, col, h64, W64, W32, stride, srec, srec8);
  return hipGetLastError();
}
hipError_t launch_gather_rows(int64_t n, int w, size_t elem, const int* ord, const void* src, void* dst,
                              hipStream_t s) {
  const int g = grid_for(n * w, 256);
  if (elem == 8)
    k_gather_rows<uint64_t><<<g, 256, 0, s>>>(n, w, ord, (const uint64_t*)src, (uint64_t*)dst);
  else if (elem == 4)
    k_gather_rows<uint32_t><<<g, 256, 0, s>>>(n, w, ord, (const uint32_t*)src, (uint32_t*)dst);
  else if (elem == 1)
    k_gather_rows<uint8_t><<<g, 256, 0, s>>>(n, w, ord, (const uint8_t*)src, (uint8_t*)dst);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
hipError_t launch_scatter_rows(int64_t n, int w, size_t elem, const int* ord, const void* src, void* dst,
                               hipStream_t s) {
  const int g = grid_for(n * w, 256);
  if (elem == 8)
    k_scatter_rows<uint64_t><<<g, 256, 0, s>>>(n, w, ord, (const uint64_t*)src, (uint64_t*)dst);
  else if (elem == 4)
    k_scatter_rows<uint32_t><<<g, 256, 0, s>>>(n, w, ord, (const uint32_t*)src, (uint32_t*)dst);
  else if (elem == 1)
    k_scatter_rows<uint8_t><<<g, 256, 0, s>>>(n, w, ord, (const uint8_t*)src, (uint8_t*)dst);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
hipError_t launch_permute_csr(int n, const int* ord, const int* inv, const int* rp_src, const int* col_src,
                              const void* val_src, bool f32, const int* rp_dst, int* col_dst, void* val_dst,
                              hipStream_t s) {
  if (f32)
    k_permute_csr<float><<<grid_for(n, 256), 256, 0, s>>>(n, ord, inv, rp_src, col_src, (const float*)val_src,
                                                          rp_dst, col_dst, (float*)val_dst);
  else
    k_permute_csr<double><<<grid_for(n, 256), 256, 0, s>>>(n, ord, inv, rp_src, col_src, (const double*)val_src,
                                                           rp_dst, col_dst, (double*)val_dst);
  return hipGetLastError();
}
hipError_t launch_fill_i32(int* p, int64_t n, int v, hipStream_t s) {
  k_fill<int><<<grid_for(n, 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t s) {
  k_fill<double><<<grid_for(n, 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
  k_fill<unsigned long long><<<grid_for(n, 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_jacobi(int n, const int* rp, const int* col, const void* val, bool f32, double* dinv,
                         int* flags, hipStream_t s) {
  if (f32)
    k_jacobi<float><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, dinv, flags);
  else
    k_jacobi<double><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, dinv, flags);
  return hipGetLastError();
}

hipError_t launch_validate_csr(int n, const int* rp, const int* col, const void* val, bool f32, int* flags,
                               hipStream_t s) {
  if (f32)
    k_validate_csr<float><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, flags);
  else
    k_validate_csr<double><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, flags);
  return hipGetLastError();
}

hipError_t launch_count_cols(int n, const int* rp, const int* col, const void* val, bool f32, int* cnt,
                             int* flags, hipStream_t s) {
  if (f32)
    k_count_cols<float><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, cnt, flags);
  else
    k_count_cols<double><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, cnt, flags);
  return hipGetLastError();
}

hipError_t launch_scatter_transpose(int n, const int* rp, const int* col, const void* val, bool f32,
                                    const double* dinv, const int* t_rp, int* fill, int* t_src,
                                    double* t_val, hipStream_t s) {
  if (f32)
    k_scatter_transpose<float><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, dinv, t_rp, fill, t_src, t_val);
  else
    k_scatter_transpose<double><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, dinv, t_rp, fill, t_src, t_val);
  return hipGetLastError();
}

hipError_t launch_copy_rows(int n, const int* rp, const int* col, const void* val, bool f32, const double* dinv,
                            int* t_rp, int* t_src, double* t_val, int* flags, hipStream_t s) {
  if (f32)
    k_copy_rows<float><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, dinv, t_rp, t_src, t_val, flags);
  else
    k_copy_rows<double><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, dinv, t_rp, t_src, t_val, flags);
  return hipGetLastError();
}

hipError_t launch_domain_rows(int n, const int* t_rp, int* t_src, double* t_val, double omega, double threshold,
                              bool has_prec, int* keep, int* flags, hipStream_t s) {
  k_domain_rows<<<grid_for((int64_t)n * 64, 256), 256, 0, s>>>(n, t_rp, t_src, t_val, omega, threshold,
                                                                 has_prec ? 1 : 0, keep, flags);
  return hipGetLastError();
}

hipError_t launch_domain_fill(int n, const int* t_rp, const int* t_src, const double* t_val, double omega,
                              double threshold, bool has_prec, const int* d_rp, uint32_t* col, double* h64,
                              float* h32, double* cdf64, float* cdf32, double* W64, float* W32, hipStream_t s) {
  k_domain_fill<<<grid_for((int64_t)n * 64, 256), 256, 0, s>>>(n, t_rp, t_src, t_val, omega, threshold,
                                                                 has_prec ? 1 : 0, d_rp, col, h64, h32, cdf64, cdf32,
                                                                 W64, W32);
  return hipGetLastError();
}

int residual_lanes(int64_t n, int64_t nnz) {
  const double z = n > 0 ? (double)nnz / (double)n : 0.0;
  // each lane takes up to 4 entries of a row in one round of loads
  if (z < 3.0) return 1;
  if (z <= 12.0) return 4;
  if (z <= 32.0) return 8;
  if (z <= 64.0) return 16;
  return 32;
}

int residual_blocks(int n, int lanes) {
  int64_t g = ((int64_t)n * lanes + kRedThreads - 1) / kRedThreads;
  if (g < 1) g = 1;
  return (int)(g < kResidualPartials ? g : kResidualPartials);
}

template <typename AV>
static hipError_t prec_residual_t(int n, int lanes, const int* rp, const int* col, const void* val,
                                  const double* dinv, const double* b, const double* x, double* r, double* part,
                                  hipStream_t s) {
  const int g = residual_blocks(n, lanes);
  switch (lanes) {
    case 1: k_prec_residual<AV, 1><<<g, kRedThreads, 0, s>>>(n, rp, col, val, dinv, b, x, r, part); break;
    case 4: k_prec_residual<AV, 4><<<g, kRedThreads, 0, s>>>(n, rp, col, val, dinv, b, x, r, part); break;
    case 8: k_prec_residual<AV, 8><<<g, kRedThreads, 0, s>>>(n, rp, col, val, dinv, b, x, r, part); break;
    case 16: k_prec_residual<AV, 16><<<g, kRedThreads, 0, s>>>(n, rp, col, val, dinv, b, x, r, part); break;
    default: k_prec_residual<AV, 32><<<g, kRedThreads, 0, s>>>(n, rp, col, val, dinv, b, x, r, part); break;
  }
  return hipGetLastError();
}

hipError_t launch_prec_residual(int n, int lanes, const int* rp, const int* col, const void* val, bool f32,
                                const double* dinv, const double* b, const double* x, double* r, double* part,
                                hipStream_t s) {
  return f32 ? prec_residual_t<float>(n, lanes, rp, col, val, dinv, b, x, r, part, s)
             : prec_residual_t<double>(n, lanes, rp, col, val, dinv, b, x, r, part, s);
}

// guide[j] for j = 0 .. 2^bits (see birth_state); Q = 0: all zero
__global__ void k_build_guide(int n, const uint64_t* __restrict__ qcdf, int bits, uint32_t* __restrict__ guide) {
  const uint64_t total = qcdf[n - 1];
  const int m = 1 << bits;
  const int bl = total ? 64 - __clzll((long long)total) : 0;
  const int sh = bl > bits ? bl - bits : 0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j <= m; j += gridDim.x * blockDim.x) {
    const int k = total ? upper_bound_u64(qcdf, n, (uint64_t)j << sh) : 0;
    guide[j] = (uint32_t)(k < n - 1 ? k : n - 1);
  }
}

int guide_bits_for(int64_t n) {
  int b = 0;
  while (b < 62 && ((int64_t)1 << b) < n) ++b;  // ceil(log2 n)
  b -= 2;                                        // about four states per entry
  return b < 4 ? 4 : (b > 22 ? 22 : b);
}

hipError_t launch_build_guide(int n, const uint64_t* qcdf, int bits, uint32_t* guide, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int m = 1 << bits;
  k_build_guide<<<grid_for(m + 1, 256), 256, 0, s>>>(n, qcdf, bits, guide);
  return hipGetLastError();
}

// Random births drawn ahead of the walk (see launch_random_births).  With a
// guide table, pass 1 keys each history by its draw's guide bucket - pure
// arithmetic on its stream, no memory read - so that after sorting by bucket
// pass 2 resolves the birth states in bucket order, where the guide and
// cumulative-weight reads of neighbouring histories share lines.  Without a
// guide the state is drawn and searched directly.
__device__ __forceinline__ uint64_t birth_draw(uint64_t kk, int64_t hid, uint64_t total) {
  uint64_t rng = history_state(kk, (uint64_t)hid);
  return mul_u53(xorshift64(rng), total);  // the walk's birth draw
}

__global__ void k_random_births(int64_t s0, int64_t count, uint64_t key, const uint64_t* __restrict__ key_dev,
                                const uint64_t* __restrict__ qcdf, const uint32_t* __restrict__ guide, int guide_bits,
                                int n, uint64_t* __restrict__ out) {
  const uint64_t total = qcdf[n - 1];
  const uint64_t kk = key_dev ? *key_dev : key;
  const int gsh = guide_shift(total, guide_bits);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < count; p += (int64_t)gridDim.x * blockDim.x) {
    uint64_t hi = 0;
    if (total) {
      const uint64_t t = birth_draw(kk, s0 + p, total);
      hi = guide ? (t >> gsh) : (uint64_t)birth_state(qcdf, guide, guide_bits, n, total, t);
    }
    out[p] = (hi << 32) | (uint64_t)p;
  }
}

// pass 2 (with a guide): the bucket in the high word becomes the state
__global__ void k_resolve_births(int64_t s0, int64_t count, uint64_t key, const uint64_t* __restrict__ key_dev,
                                 const uint64_t* __restrict__ qcdf, const uint32_t* __restrict__ guide,
                                 int guide_bits, int n, uint64_t* __restrict__ io) {
  const uint64_t total = qcdf[n - 1];
  const uint64_t kk = key_dev ? *key_dev : key;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t kv = io[i];
    const uint64_t p = kv & 0xFFFFFFFFull;
    int st = 0;
    if (total) st = birth_state(qcdf, guide, guide_bits, n, total, birth_draw(kk, s0 + (int64_t)p, total));
    io[i] = ((uint64_t)(uint32_t)st << 32) | p;
  }
}

hipError_t launch_random_births(int64_t s0, int64_t count, uint64_t key, const uint64_t* key_dev,
                                const uint64_t* qcdf, const uint32_t* guide, int guide_bits, int n, uint64_t* out,
                                hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (count >= ((int64_t)1 << 32) || n <= 0) return hipErrorInvalidValue;
  k_random_births<<<grid_for(count, 256), 256, 0, s>>>(s0, count, key, key_dev, qcdf, guide, guide_bits, n, out);
  return hipGetLastError();
}

hipError_t launch_resolve_births(int64_t s0, int64_t count, uint64_t key, const uint64_t* key_dev,
                                 const uint64_t* qcdf, const uint32_t* guide, int guide_bits, int n, uint64_t* io,
                                 hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (count >= ((int64_t)1 << 32) || n <= 0 || !guide) return hipErrorInvalidValue;
  k_resolve_births<<<grid_for(count, 256), 256, 0, s>>>(s0, count, key, key_dev, qcdf, guide, guide_bits, n, io);
  return hipGetLastError();
}

hipError_t launch_sort_births(const uint64_t* in, uint64_t* out, int64_t count, int key_bits, void* temp,
                              size_t* temp_bytes, hipStream_t s) {
  if (count < 0 || count >= ((int64_t)1 << 32) || key_bits < 1 || key_bits > 31) return hipErrorInvalidValue;
  // stable LSD passes over the key bits only: within a key the entries keep
  // their id order
  return rocprim::radix_sort_keys(temp, *temp_bytes, in, out, (uint32_t)count, 32u, 32u + (unsigned)key_bits, s);
}

constexpr int kSellRound = 16;  // entries per lane per round of loads

int sell_blocks(int nslices) {
  // a multiple of 8 (one share per XCD), at most kResidualPartials
  int64_t g = ((int64_t)nslices * kWave + kRedThreads - 1) / kRedThreads;
  g = (g + 7) / 8 * 8;
  if (g < 8) g = 8;
  return (int)(g < kResidualPartials ? g : kResidualPartials);
}

hipError_t launch_sell_widths(int n, const int* rp, int* width, hipStream_t s) {
  const int nslices = (n + kWave - 1) / kWave;
  if (nslices == 0) return hipSuccess;
  k_sell_widths<<<(nslices + 3) / 4, 4 * kWave, 0, s>>>(n, rp, width);
  return hipGetLastError();
}

hipError_t launch_sell_fill(int n, const int* rp, const int* col, const void* val, bool f32, const int* width,
                            const int64_t* off, int* scol, void* sval, hipStream_t s) {
  const int nslices = (n + kWave - 1) / kWave;
  if (nslices == 0) return hipSuccess;
  const int64_t threads = (int64_t)nslices * kWave;
  const int g = (int)((threads + 255) / 256);
  if (f32)
    k_sell_fill<float><<<g, 256, 0, s>>>(n, nslices, rp, col, (const float*)val, width, off, scol, (float*)sval);
  else
    k_sell_fill<double><<<g, 256, 0, s>>>(n, nslices, rp, col, (const double*)val, width, off, scol,
                                          (double*)sval);
  return hipGetLastError();
}

hipError_t launch_sell_residual(int n, const int* width, const int64_t* off, const int* scol, const void* sval,
                                bool f32, const double* dinv, const double* b, const double* x, double* r,
                                double* part, hipStream_t s) {
  const int nslices = (n + kWave - 1) / kWave;
  const int g = sell_blocks(nslices);
  if (f32)
    k_sell_residual<float, kSellRound><<<g, kRedThreads, 0, s>>>(n, nslices, width, off, scol, (const float*)sval,
                                                                 dinv, b, x, r, part);
  else
    k_sell_residual<double, kSellRound><<<g, kRedThreads, 0, s>>>(n, nslices, width, off, scol,
                                                                  (const double*)sval, dinv, b, x, r, part);
  return hipGetLastError();
}

hipError_t launch_domain_power(int n, const int* d_rp, const uint32_t* col, const double* h64, const double* W64,
                               int mode, const double* x, const double* nrm2_prev, double* y, double* part,
                               double* nrm2, double* logsum, hipStream_t s) {
  const int g = residual_blocks(n, 1);
  k_domain_power<<<g, kRedThreads, 0, s>>>(n, d_rp, col, h64, W64, mode, x, nrm2_prev, y, part);
  k_sum_partials<<<1, kRedThreads, 0, s>>>(part, g, nrm2);
  k_log_accum<<<1, 1, 0, s>>>(nrm2, logsum);
  return hipGetLastError();
}

hipError_t launch_power_start(int n, uint64_t seed, double* x, hipStream_t s) {
  k_power_start<<<grid_for(n, 256), 256, 0, s>>>(n, seed, x);
  return hipGetLastError();
}

hipError_t launch_balance(int n, const int* d_rp, const uint32_t* col, const double* h64, const double* W64,
                          unsigned long long* cmax, unsigned long long* mm, double* part, double* sums,
                          hipStream_t s) {
  k_fill<unsigned long long><<<grid_for(n, 256), 256, 0, s>>>(cmax, n, 0ULL);
  k_col_absmax<<<grid_for(n, 256), 256, 0, s>>>(n, d_rp, col, h64, cmax);
  const unsigned long long init[4] = {0ULL, ~0ULL, ~0ULL, 0ULL};
  for (int k = 0; k < 4; ++k) k_fill<unsigned long long><<<1, 1, 0, s>>>(mm + k, 1, init[k]);
  const int g = residual_blocks(n, 1);
  k_balance<<<g, kRedThreads, 0, s>>>(n, d_rp, h64, W64, cmax, mm, part);
  k_sum_partials<<<1, kRedThreads, 0, s>>>(part, g, sums);
  k_sum_partials<<<1, kRedThreads, 0, s>>>(part + kResidualPartials, g, sums + 1);
  return hipGetLastError();
}

hipError_t launch_sum_partials(const double* part, int m, double* out, hipStream_t s) {
  k_sum_partials<<<1, kRedThreads, 0, s>>>(part, m, out);
  return hipGetLastError();
}

hipError_t launch_scaled_sq(int n, const double* dinv, const double* b, double* part, hipStream_t s) {
  k_scaled_sq<<<residual_blocks(n, 1), kRedThreads, 0, s>>>(n, dinv, b, part);
  return hipGetLastError();
}

hipError_t launch_segment_norm(int nsys, const int64_t* off, const double* v, const double* dinv, bool square,
                               double* out, hipStream_t s) {
  const int g = nsys < 65535 ? nsys : 65535;
  if (square)
    k_segment_norm<true><<<g, kRedThreads, 0, s>>>(nsys, off, v, dinv, out);
  else
    k_segment_norm<false><<<g, kRedThreads, 0, s>>>(nsys, off, v, dinv, out);
  return hipGetLastError();
}

hipError_t launch_batch_check(int nsys, const double* r2, const double* b2, double tol, int* conv,
                              int* unconverged, hipStream_t s) {
  k_batch_check<<<1, kRedThreads, 0, s>>>(nsys, r2, b2, tol, conv, unconverged);
  return hipGetLastError();
}

hipError_t launch_batch_source(int n, const int* sys_of, const int64_t* off, const double* r, double omega,
                               const int* conv, const double* l1, double* rhs, double* mag, hipStream_t s) {
  k_batch_source<<<grid_for(n, 256), 256, 0, s>>>(n, sys_of, off, r, omega, conv, l1, rhs, mag);
  return hipGetLastError();
}

hipError_t launch_axpy(int n, double a, const double* y, double* x, hipStream_t s) {
  k_axpy<<<grid_for(n, 256), 256, 0, s>>>(n, a, y, x);
  return hipGetLastError();
}

hipError_t launch_scale_copy(int n, double a, const double* x, double* y, hipStream_t s) {
  k_scale_copy<<<grid_for(n, 256), 256, 0, s>>>(n, a, x, y);
  return hipGetLastError();
}

hipError_t launch_diag_scale(int n, const double* d, const double* x, double* y, hipStream_t s) {
  k_diag_scale<<<grid_for(n, 256), 256, 0, s>>>(n, d, x, y);
  return hipGetLastError();
}

hipError_t launch_apply_tally(int n, double* x, double* t64, float* t32, const double* rhs, bool add_rhs,
                              double inv, const uint64_t* mtotal, hipStream_t s) {
  k_tally_out<true><<<grid_for(n, 256), 256, 0, s>>>(n, x, t64, t32, rhs, add_rhs ? 1 : 0, inv, mtotal);
  return hipGetLastError();
}

hipError_t launch_store_tally(int n, double* x, double* t64, float* t32, const double* rhs, bool add_rhs,
                              double inv, const uint64_t* mtotal, hipStream_t s) {
  k_tally_out<false><<<grid_for(n, 256), 256, 0, s>>>(n, x, t64, t32, rhs, add_rhs ? 1 : 0, inv, mtotal);
  return hipGetLastError();
}

hipError_t launch_sort_rows(int n, const int* rp, int* col, double* val, int* flags, hipStream_t s) {
  k_sort_rows<<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, flags);
  return hipGetLastError();
}

hipError_t launch_spgemm_count(int n, const int* x_rp, const int* x_col, const int* y_rp, const int* y_col, int* head,
                               int* cnt, hipStream_t s) {
  k_spgemm<false><<<grid_for(n, 256), 256, 0, s>>>(n, x_rp, x_col, nullptr, y_rp, y_col, nullptr, head, cnt,
                                                   nullptr, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_spgemm_fill(int n, const int* x_rp, const int* x_col, const double* x_val, const int* y_rp,
                              const int* y_col, const double* y_val, int* head, const int* c_rp, int* c_col,
                              double* c_val, hipStream_t s) {
  k_spgemm<true><<<grid_for(n, 256), 256, 0, s>>>(n, x_rp, x_col, x_val, y_rp, y_col, y_val, head, nullptr, c_rp,
                                                  c_col, c_val);
  return hipGetLastError();
}

hipError_t launch_filter_count(int n, const int* rp, const double* val, double threshold, int* keep, hipStream_t s) {
  k_filter_rows<false><<<grid_for(n, 256), 256, 0, s>>>(n, rp, nullptr, val, threshold, keep, nullptr, nullptr,
                                                        nullptr);
  return hipGetLastError();
}

hipError_t launch_filter_fill(int n, const int* rp, const int* col, const double* val, double threshold,
                              const int* out_rp, int* out_col, double* out_val, hipStream_t s) {
  k_filter_rows<true><<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, threshold, nullptr, out_rp, out_col,
                                                       out_val);
  return hipGetLastError();
}

hipError_t launch_csr_spmv(int n, const int* rp, const int* col, const double* val, const double* x, double* y,
                           hipStream_t s) {
  k_csr_spmv<<<grid_for(n, 256), 256, 0, s>>>(n, rp, col, val, x, y);
  return hipGetLastError();
}

}  // namespace mcls
