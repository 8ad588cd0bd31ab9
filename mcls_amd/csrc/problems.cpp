// Synthetic workload systems (host only).  See include/mcls_problems.h.
//
// Every random quantity is a counter-based draw u(tag, k) so that rows can be
// generated in parallel and independently, and the same bits come out on any
// machine.
#include "mcls_problems.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

inline uint64_t splitmix64(uint64_t z) {
  z += kGolden;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

enum : uint64_t {
  TAG_B = 0xB0,
  TAG_D = 0xD0,
  TAG_SIGMA = 0x51,
  TAG_Q = 0x90,
  TAG_JITTER = 0x11,
  TAG_S = 0x55,
  TAG_GRF = 0x6F,
  TAG_BATCH = 0xBA
};

struct Stream {
  uint64_t key;
  Stream(uint64_t seed, uint64_t tag) : key(splitmix64(seed ^ splitmix64(tag))) {}
  double u(uint64_t k) const {
    return (double)(splitmix64(key ^ splitmix64(k)) >> 11) * 0x1.0p-53;
  }
};

template <class F>
void parallel_rows(int64_t n, int nthreads, F&& f) {
  if (nthreads <= 1 || n < 4096) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    int64_t a = n * t / nthreads, b = n * (t + 1) / nthreads;
    th.emplace_back([&f, a, b] { f(a, b); });
  }
  for (auto& x : th) x.join();
}

struct Grid {
  int64_t nx, ny, nz;
  int64_t n() const { return nx * ny * nz; }
};

// Row lengths of the structured operators (neighbours inside the grid).
int64_t row_len_5pt(const Grid& g, int64_t i) {
  int64_t x = i % g.nx, y = i / g.nx;
  return 1 + (x > 0) + (x < g.nx - 1) + (y > 0) + (y < g.ny - 1);
}
int64_t row_len_7pt(const Grid& g, int64_t i) {
  int64_t x = i % g.nx, y = (i / g.nx) % g.ny, z = i / (g.nx * g.ny);
  return 1 + (x > 0) + (x < g.nx - 1) + (y > 0) + (y < g.ny - 1) + (z > 0) + (z < g.nz - 1);
}
int64_t row_len_27pt(const Grid& g, int64_t i) {
  int64_t x = i % g.nx, y = (i / g.nx) % g.ny, z = i / (g.nx * g.ny);
  int64_t cx = 1 + (x > 0) + (x < g.nx - 1);
  int64_t cy = 1 + (y > 0) + (y < g.ny - 1);
  int64_t cz = 1 + (z > 0) + (z < g.nz - 1);
  return cx * cy * cz;
}

bool grid_of(const mcls_problem_spec* s, Grid* g) {
  switch (s->kind) {
    case MCLS_PROBLEM_POISSON2D_5PT:
    case MCLS_PROBLEM_ANISO2D_5PT:
    case MCLS_PROBLEM_VARCOEF2D_5PT:
      *g = {s->nx, s->ny, 1};
      return s->nx > 0 && s->ny > 0;
    case MCLS_PROBLEM_DIFFUSION3D_7PT:
    case MCLS_PROBLEM_STENCIL3D_27PT:
      *g = {s->nx, s->ny > 0 ? s->ny : s->nx, s->nz > 0 ? s->nz : s->nx};
      return s->nx > 0;
    default:
      return false;
  }
}

int64_t row_len(const mcls_problem_spec* s, const Grid& g, int64_t i) {
  switch (s->kind) {
    case MCLS_PROBLEM_DIFFUSION3D_7PT: return row_len_7pt(g, i);
    case MCLS_PROBLEM_STENCIL3D_27PT: return row_len_27pt(g, i);
    default: return row_len_5pt(g, i);
  }
}

// Gaussian random field by random Fourier features (Rahimi & Recht 2007):
// g(x) = sigma sqrt(2/M) sum_m cos(w_m . x + phi_m), w_m ~ N(0, I/ell^2),
// phi_m ~ U[0, 2 pi): covariance sigma^2 exp(-|r|^2 / (2 ell^2)).
struct Grf2D {
  std::vector<double> wx, wy, ph;
  double amp;
  Grf2D(uint64_t seed, double ell, double sigma, int M) : wx(M), wy(M), ph(M) {
    Stream s(seed, TAG_GRF);
    for (int m = 0; m < M; ++m) {
      double u1 = s.u(4 * m) , u2 = s.u(4 * m + 1), u3 = s.u(4 * m + 2), u4 = s.u(4 * m + 3);
      double r = std::sqrt(-2.0 * std::log(1.0 - u1));
      wx[m] = r * std::cos(2.0 * M_PI * u2) / ell;
      wy[m] = r * std::sin(2.0 * M_PI * u2) / ell;
      (void)u3;
      ph[m] = 2.0 * M_PI * u4;
    }
    amp = sigma * std::sqrt(2.0 / M);
  }
  double operator()(double x, double y) const {
    double acc = 0.0;
    for (size_t m = 0; m < wx.size(); ++m) acc += std::cos(wx[m] * x + wy[m] * y + ph[m]);
    return amp * acc;
  }
};

}  // namespace

extern "C" {

double mcls_problem_uniform(uint64_t seed, uint64_t tag, uint64_t k) {
  return Stream(seed, tag).u(k);
}

int64_t mcls_problem_dims(const mcls_problem_spec* spec, int64_t* n_out) {
  Grid g;
  if (!spec || !grid_of(spec, &g)) return -1;
  int64_t n = g.n();
  if (n_out) *n_out = n;
  switch (spec->kind) {
    case MCLS_PROBLEM_DIFFUSION3D_7PT:
      return 7 * n - 2 * (g.ny * g.nz + g.nx * g.nz + g.nx * g.ny);
    case MCLS_PROBLEM_STENCIL3D_27PT:
      return (3 * g.nx - 2) * (3 * g.ny - 2) * (3 * g.nz - 2);
    default:
      return 5 * n - 2 * (g.nx + g.ny);
  }
}

int mcls_problem_build(const mcls_problem_spec* spec, int64_t* row_ptr,
                       int32_t* col, double* val64, float* val32, double* b,
                       int nthreads) {
  Grid g;
  if (!spec || !grid_of(spec, &g) || !row_ptr || !col || !b) return -1;
  const int64_t n = g.n();
  if (n > INT32_MAX) return -1;
  // Row pointers: lengths in parallel, prefix sum serially.
  parallel_rows(n, nthreads, [&](int64_t a, int64_t e) {
    for (int64_t i = a; i < e; ++i) row_ptr[i + 1] = row_len(spec, g, i);
  });
  row_ptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) row_ptr[i + 1] += row_ptr[i];

  const Stream sb(spec->seed, TAG_B);
  const int kind = spec->kind;

  if (kind == MCLS_PROBLEM_STENCIL3D_27PT) {
    const Stream sj(spec->seed, TAG_JITTER), ss(spec->seed, TAG_S);
    const double slo = spec->p0, shi = spec->p1, jlo = spec->p2, jhi = spec->p3;
    parallel_rows(n, nthreads, [&](int64_t a, int64_t e) {
      float vals[27];
      int32_t cols[27];
      for (int64_t i = a; i < e; ++i) {
        int64_t x = i % g.nx, y = (i / g.nx) % g.ny, z = i / (g.nx * g.ny);
        int m = 0, diag_pos = -1;
        double off_sum = 0.0;
        for (int dz = -1; dz <= 1; ++dz) {
          if (z + dz < 0 || z + dz >= g.nz) continue;
          for (int dy = -1; dy <= 1; ++dy) {
            if (y + dy < 0 || y + dy >= g.ny) continue;
            for (int dx = -1; dx <= 1; ++dx) {
              if (x + dx < 0 || x + dx >= g.nx) continue;
              int64_t j = i + dx + g.nx * (dy + g.ny * dz);
              cols[m] = (int32_t)j;
              if (j == i) {
                diag_pos = m;
                vals[m] = 0.0f;
              } else {
                uint64_t lo = (uint64_t)std::min(i, j), hi = (uint64_t)std::max(i, j);
                double v = -(jlo + (jhi - jlo) * sj.u(lo * (uint64_t)n + hi));
                vals[m] = (float)v;
                off_sum += std::fabs((double)vals[m]);
              }
              ++m;
            }
          }
        }
        double s = slo + (shi - slo) * ss.u((uint64_t)i);
        vals[diag_pos] = (float)(off_sum / s);
        int64_t beg = row_ptr[i];
        for (int k = 0; k < m; ++k) {
          col[beg + k] = cols[k];
          if (val32) val32[beg + k] = vals[k];
          if (val64) val64[beg + k] = (double)vals[k];
        }
        b[i] = sb.u((uint64_t)i);
      }
    });
    return 0;
  }

  // FP64-native kinds.
  const Stream sd(spec->seed, TAG_D), ssg(spec->seed, TAG_SIGMA), sq(spec->seed, TAG_Q);
  std::vector<double> kappa;
  if (kind == MCLS_PROBLEM_VARCOEF2D_5PT) {
    kappa.resize((size_t)n);
    int M = spec->p2 > 0 ? (int)spec->p2 : 256;
    Grf2D grf(spec->seed, spec->p0 > 0 ? spec->p0 : 0.1, spec->p1 > 0 ? spec->p1 : 1.0, M);
    double hx = 1.0 / (double)g.nx, hy = 1.0 / (double)g.ny;
    parallel_rows(n, nthreads, [&](int64_t a, int64_t e) {
      for (int64_t i = a; i < e; ++i) {
        int64_t x = i % g.nx, y = i / g.nx;
        kappa[(size_t)i] = std::exp(grf((x + 0.5) * hx, (y + 0.5) * hy));
      }
    });
  }
  const int64_t bs = spec->p0 > 0 ? (int64_t)spec->p0 : 8;
  auto Dcell = [&](int64_t x, int64_t y, int64_t z) {
    int64_t nbx = (g.nx + bs - 1) / bs, nby = (g.ny + bs - 1) / bs;
    int64_t blk = (x / bs) + nbx * ((y / bs) + nby * (z / bs));
    return 0.5 * std::pow(4.0, sd.u((uint64_t)blk));
  };
  parallel_rows(n, nthreads, [&](int64_t a, int64_t e) {
    for (int64_t i = a; i < e; ++i) {
      int64_t beg = row_ptr[i];
      int m = 0;
      auto put = [&](int64_t j, double v) {
        col[beg + m] = (int32_t)j;
        if (val64) val64[beg + m] = v;
        if (val32) val32[beg + m] = (float)v;
        ++m;
      };
      if (kind == MCLS_PROBLEM_POISSON2D_5PT || kind == MCLS_PROBLEM_ANISO2D_5PT) {
        double ex = 1.0, ey = kind == MCLS_PROBLEM_ANISO2D_5PT ? spec->p0 : 1.0;
        int64_t x = i % g.nx, y = i / g.nx;
        if (y > 0) put(i - g.nx, -ey);
        if (x > 0) put(i - 1, -ex);
        put(i, 2.0 * ex + 2.0 * ey);
        if (x < g.nx - 1) put(i + 1, -ex);
        if (y < g.ny - 1) put(i + g.nx, -ey);
        b[i] = sb.u((uint64_t)i);
      } else if (kind == MCLS_PROBLEM_VARCOEF2D_5PT) {
        int64_t x = i % g.nx, y = i / g.nx;
        double ih2 = (double)g.nx * (double)g.nx;
        double ki = kappa[(size_t)i];
        auto face = [&](int64_t j) { double kj = kappa[(size_t)j]; return 2.0 * ki * kj / (ki + kj) * ih2; };
        double tym = y > 0 ? face(i - g.nx) : 2.0 * ki * ih2;
        double txm = x > 0 ? face(i - 1) : 2.0 * ki * ih2;
        double txp = x < g.nx - 1 ? face(i + 1) : 2.0 * ki * ih2;
        double typ = y < g.ny - 1 ? face(i + g.nx) : 2.0 * ki * ih2;
        if (y > 0) put(i - g.nx, -tym);
        if (x > 0) put(i - 1, -txm);
        put(i, tym + txm + txp + typ);
        if (x < g.nx - 1) put(i + 1, -txp);
        if (y < g.ny - 1) put(i + g.nx, -typ);
        b[i] = sb.u((uint64_t)i);
      } else {  // DIFFUSION3D_7PT
        int64_t x = i % g.nx, y = (i / g.nx) % g.ny, z = i / (g.nx * g.ny);
        double ih2 = (double)g.nx * (double)g.nx;
        double di = Dcell(x, y, z);
        auto face = [&](bool inside, int64_t xx, int64_t yy, int64_t zz) {
          if (!inside) return 2.0 * di * ih2;
          double dj = Dcell(xx, yy, zz);
          return 2.0 * di * dj / (di + dj) * ih2;
        };
        double tzm = face(z > 0, x, y, z - 1), tym = face(y > 0, x, y - 1, z);
        double txm = face(x > 0, x - 1, y, z), txp = face(x < g.nx - 1, x + 1, y, z);
        double typ = face(y < g.ny - 1, x, y + 1, z), tzp = face(z < g.nz - 1, x, y, z + 1);
        double sig = 0.01 + 0.09 * ssg.u((uint64_t)i);
        int64_t sxy = g.nx * g.ny;
        if (z > 0) put(i - sxy, -tzm);
        if (y > 0) put(i - g.nx, -tym);
        if (x > 0) put(i - 1, -txm);
        put(i, tzm + tym + txm + txp + typ + tzp + sig);
        if (x < g.nx - 1) put(i + 1, -txp);
        if (y < g.ny - 1) put(i + g.nx, -typ);
        if (z < g.nz - 1) put(i + sxy, -tzp);
        b[i] = sq.u((uint64_t)i);
      }
    }
  });
  return 0;
}

int mcls_problem_batch_member(uint64_t seed, int32_t index, int32_t lo,
                              int32_t hi, double eps_hi,
                              mcls_problem_spec* out) {
  if (!out || hi < lo || lo < 1 || eps_hi < 1.0) return -1;
  Stream s(seed, TAG_BATCH);
  int32_t n = lo + (int32_t)std::floor(s.u(2 * (uint64_t)index) * (double)(hi - lo + 1));
  if (n > hi) n = hi;
  double eps = std::exp(s.u(2 * (uint64_t)index + 1) * std::log(eps_hi));
  std::memset(out, 0, sizeof(*out));
  out->kind = MCLS_PROBLEM_ANISO2D_5PT;
  out->nx = n;
  out->ny = n;
  out->nz = 1;
  out->seed = splitmix64(seed ^ splitmix64(0x5E5E0000ULL + (uint64_t)index));
  out->p0 = eps;
  return 0;
}

}  // extern "C"
