/*
 * mcls_problems.h -- synthetic linear systems for the MCSA workloads.
 *
 * No counterpart entry point in the reference: MCLS builds its systems in
 * example drivers (example/poisson/poisson_driver.cpp:80-114, a 1-D 3-point
 * operator; example/neutron_diffusion/DiffusionProblem.cpp:62-85, a 2-D
 * 9-point diffusion operator).  These generators produce the workload systems
 * in plain CSR (int64 row pointers, int32 columns, sorted unique columns per
 * row) with every random value drawn from a counter-based SplitMix64 stream,
 * so any row can be regenerated independently and the CPU and GPU see the
 * same input bits.  Host-only library: libmcls_problems.so.
 */
#ifndef MCLS_PROBLEMS_H
#define MCLS_PROBLEMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum mcls_problem_kind {
  /* 2-D 5-point Laplacian on an nx x ny interior grid, Dirichlet boundary
   * eliminated: a_ii = 4, neighbours -1.  b ~ U[0,1). */
  MCLS_PROBLEM_POISSON2D_5PT = 1,
  /* 3-D single-group diffusion-reaction, cell-centred 7-point finite volume
   * on nx^3 cells, h = 1/nx, homogeneous Dirichlet at the outer faces.
   * D piecewise constant on p0^3-cell blocks, log-uniform in [0.5, 2];
   * Sigma_a ~ U[0.01, 0.1] and q ~ U[0,1) per cell; harmonic-mean face
   * coefficients.  b = q. */
  MCLS_PROBLEM_DIFFUSION3D_7PT = 2,
  /* 2-D anisotropic 5-point operator -u_xx - eps u_yy (eps = p0):
   * a_ii = 2 + 2 eps, x neighbours -1, y neighbours -eps.  b ~ U[0,1). */
  MCLS_PROBLEM_ANISO2D_5PT = 3,
  /* 2-D variable-coefficient diffusion, 5-point finite volume on nx x ny
   * cells, kappa = exp(g) with g a zero-mean Gaussian random field of
   * squared-exponential covariance (correlation length p0, std p1) drawn
   * by 512 random Fourier features; harmonic-mean faces; Dirichlet.
   * b ~ U[0,1). */
  MCLS_PROBLEM_VARCOEF2D_5PT = 4,
  /* 3-D 27-point stencil on nx^3 points, Dirichlet eliminated.
   * Off-diagonals -U[p2, p3] drawn per unordered edge (A symmetric);
   * a_ii = sum_j |a_ij| / s_i with s_i ~ U[p0, p1], so the row sum of |H|
   * under point Jacobi is s_i.  b ~ U[0,1).  Natural dtype FP32. */
  MCLS_PROBLEM_STENCIL3D_27PT = 5
};

typedef struct mcls_problem_spec {
  int32_t kind;
  int32_t nx, ny, nz;
  uint64_t seed;
  double p0, p1, p2, p3;
} mcls_problem_spec;

/* Number of unknowns (*n_out) and stored entries (return value); negative on
 * a bad spec. */
int64_t mcls_problem_dims(const mcls_problem_spec* spec, int64_t* n_out);

/* Fill CSR arrays (row_ptr n+1, col nnz) and the right-hand side b (n).
 * val64 and/or val32 may be NULL; when both are given val64 holds the
 * FP32 values widened for the 27-point kind and val32 the FP64 values
 * rounded for the others.  nthreads <= 0 means one thread.  Returns 0. */
int mcls_problem_build(const mcls_problem_spec* spec, int64_t* row_ptr,
                       int32_t* col, double* val64, float* val32, double* b,
                       int nthreads);

/* Parameters of system `index` of the batched anisotropic family:
 * n ~ U{lo..hi}, eps log-uniform in [1, eps_hi], per-system seed. */
int mcls_problem_batch_member(uint64_t seed, int32_t index, int32_t lo,
                              int32_t hi, double eps_hi,
                              mcls_problem_spec* out);

/* Counter-based uniform in [0,1): (splitmix64(key ^ splitmix64(k)) >> 11)
 * * 2^-53, key = splitmix64(seed ^ tag).  Exposed for tests. */
double mcls_problem_uniform(uint64_t seed, uint64_t tag, uint64_t k);

#ifdef __cplusplus
}
#endif
#endif
