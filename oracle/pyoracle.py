"""ctypes bindings to the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  See
mcls_oracle.h for the semantics and the reference files each function follows.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

EST_COLLISION = 0
EST_EXPECTED_VALUE = 1
EST_FORWARD = 2
SRC_RANDOM = 0
SRC_STRATIFIED = 1
SAMPLER_CDF = 0
SAMPLER_ALIAS = 1
SAMPLER_LEARNED = 2
SAMPLER_LEARNED_ALIAS = 3
SAMPLER_ALIAS_STRIDED = 4

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class MCParams(C.Structure):
    _fields_ = [
        ("num_histories", C.c_int64),
        ("history_length", C.c_int64),
        ("weight_cutoff", C.c_double),
        ("precision", C.c_int64),
        ("estimator", C.c_int64),
        ("source", C.c_int64),
        ("seed", C.c_uint64),
        ("num_threads", C.c_int64),
        ("sampler", C.c_int64),
        ("importance", C.c_void_p),
    ]


class SolverParams(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("max_iters", C.c_int64),
        ("check_freq", C.c_int64),
        ("smoother_steps", C.c_int64),
        ("richardson_omega", C.c_double),
        ("neumann_omega", C.c_double),
        ("threshold", C.c_double),
        ("num_sets", C.c_int64),
        ("reference_mode", C.c_int64),
        ("mc", MCParams),
        ("ref_quirk", C.c_int64),
    ]


class SolverResult(C.Structure):
    _fields_ = [
        ("iters", C.c_int64),
        ("converged", C.c_int64),
        ("achieved_tol", C.c_double),
        ("histories", C.c_int64),
        ("steps", C.c_int64),
        ("seconds_mc", C.c_double),
        ("seconds_total", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    u64, i64, i32, f64, f32 = C.c_uint64, C.c_int64, C.c_int, C.c_double, C.c_float
    vp = C.c_void_p
    L.orc_splitmix64.argtypes = [u64]
    L.orc_splitmix64.restype = u64
    L.orc_xorshift64.argtypes = [C.POINTER(u64)]
    L.orc_xorshift64.restype = u64
    L.orc_history_state.argtypes = [u64, u64, u64]
    L.orc_history_state.restype = u64
    L.orc_u01_f64.argtypes = [u64]
    L.orc_u01_f64.restype = f64
    L.orc_u01_f32.argtypes = [u64]
    L.orc_u01_f32.restype = f32
    L.orc_lower_bound_f64.argtypes = [_f64p, i32, f64]
    L.orc_lower_bound_f64.restype = i32
    L.orc_lower_bound_f32.argtypes = [_f32p, i32, f32]
    L.orc_lower_bound_f32.restype = i32
    L.orc_upper_bound_u64.argtypes = [_u64p, i64, u64]
    L.orc_upper_bound_u64.restype = i64
    L.orc_build_domain.argtypes = [i32, _i64p, _i32p, _f64p, vp, f64, f64,
                                   _i64p, _i32p, _f64p, _f64p, _f64p]
    L.orc_build_domain.restype = i64
    L.orc_build_domain_forward.argtypes = L.orc_build_domain.argtypes
    L.orc_build_domain_forward.restype = i64
    L.orc_forward_per_state.argtypes = [i32, i64]
    L.orc_forward_per_state.restype = i64
    L.orc_build_proposal.argtypes = [i32, _i64p, _i32p, _f64p, _f32p, _f64p, _f64p]
    L.orc_build_proposal.restype = i32
    L.orc_build_alias.argtypes = [i32, _i64p, _f64p, _f64p, _f64p, _i32p]
    L.orc_build_alias.restype = i64
    L.orc_alias_stride.argtypes = [i32, _i64p]
    L.orc_alias_stride.restype = i32
    L.orc_build_alias_strided.argtypes = [i32, _i64p, _f64p, _f64p, i32, _f64p, _i32p]
    L.orc_build_alias_strided.restype = i64
    L.orc_build_proposal_alias.argtypes = [i32, _i64p, _i32p, _f64p, _f32p, _f64p, _i32p, _f64p]
    L.orc_build_proposal_alias.restype = i64
    L.orc_jacobi_dinv.argtypes = [i32, _i64p, _i32p, _f64p, _f64p]
    L.orc_jacobi_dinv.restype = i32
    L.orc_source_quantize.argtypes = [i32, _f64p, _u64p, _u64p]
    L.orc_source_quantize.restype = u64
    L.orc_source_birth.argtypes = [i32, _f64p, _u64p, _u64p, u64, C.POINTER(u64), C.POINTER(f64)]
    L.orc_source_birth.restype = i32
    dom = [i32, _i64p, _i32p, _f64p, _f64p, _f64p]
    L.orc_mc_solve.argtypes = dom + [_f64p, C.POINTER(MCParams), u64, _f64p, C.POINTER(i64)]
    L.orc_mc_solve.restype = i32
    L.orc_mc_solve_window.argtypes = dom + [_f64p, C.c_void_p, C.POINTER(MCParams), u64, i64, i64, _f64p,
                                            C.POINTER(i64)]
    L.orc_mc_solve_window.restype = i32
    L.orc_record_paths.argtypes = dom + [_f64p, C.POINTER(MCParams), u64, i64, i64, i32,
                                         _i32p, _f64p]
    L.orc_record_paths.restype = i32
    L.orc_spmv.argtypes = [i32, _i64p, _i32p, _f64p, _f64p, _f64p]
    L.orc_spmv.restype = None
    L.orc_transpose.argtypes = [i32, _i64p, _i32p, _f64p, _i64p, _i32p, _f64p]
    L.orc_transpose.restype = None
    L.orc_spgemm.argtypes = [i32, _i64p, _i32p, _f64p, _i64p, _i32p, _f64p, f64, i64,
                             _i64p, _i32p, _f64p]
    L.orc_spgemm.restype = i64
    L.orc_prec_residual.argtypes = [i32, _i64p, _i32p, _f64p, vp, _f64p, _f64p, _f64p]
    L.orc_prec_residual.restype = None
    for nm in ("orc_norm2", "orc_norm1", "orc_norminf"):
        getattr(L, nm).argtypes = [i32, _f64p]
        getattr(L, nm).restype = f64
    L.orc_mcsa_solve.argtypes = [i32, _i64p, _i32p, _f64p, vp, _f64p, _f64p,
                                 C.POINTER(SolverParams), u64, C.POINTER(SolverResult), vp]
    L.orc_mcsa_solve.restype = i32
    L.orc_fixed_point_solve.argtypes = [i32, _i64p, _i32p, _f64p, vp, _f64p, _f64p,
                                        C.POINTER(SolverParams), C.POINTER(SolverResult)]
    L.orc_fixed_point_solve.restype = i32
    L.orc_ref_build_source.argtypes = [i32, _f64p, f64, i32, i32, C.POINTER(u64), _i64p]
    L.orc_ref_build_source.restype = i64
    L.orc_ref_mc_solve.argtypes = dom + [_f64p, C.POINTER(MCParams), i32, i32, C.POINTER(u64),
                                         _f64p, C.POINTER(i64)]
    L.orc_ref_mc_solve.restype = i64
    L.orc_ref_transport.argtypes = [_i64p, _i32p, _f64p, _f64p, _f64p, i64, f64,
                                    C.POINTER(u64), C.POINTER(i32), C.POINTER(f64),
                                    C.POINTER(i64), _f64p]
    L.orc_ref_transport.restype = i32
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Csr:
    """Plain host CSR (int64 row pointers, int32 columns, float64 values)."""

    def __init__(self, rp, col, val, n=None):
        self.rp = np.ascontiguousarray(rp, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = int(n if n is not None else len(self.rp) - 1)

    @classmethod
    def from_dense(cls, A, keep_zeros=False):
        A = np.asarray(A, dtype=np.float64)
        rp, col, val = [0], [], []
        for i in range(A.shape[0]):
            for j in range(A.shape[1]):
                if A[i, j] != 0.0 or (keep_zeros and keep_zeros(i, j)):
                    col.append(j)
                    val.append(A[i, j])
            rp.append(len(col))
        return cls(rp, col, val, A.shape[0])

    @classmethod
    def from_scipy(cls, S):
        S = S.tocsr()
        S.sort_indices()
        return cls(S.indptr, S.indices, S.data, S.shape[0])

    def to_dense(self):
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            for k in range(self.rp[i], self.rp[i + 1]):
                A[i, self.col[k]] += self.val[k]
        return A


class Domain:
    def __init__(self, n, rp, col, h, cdf, W):
        self.n, self.rp, self.col, self.h, self.cdf, self.W = n, rp, col, h, cdf, W

    @property
    def nnz(self):
        return int(self.rp[-1])

    def args(self):
        return (self.n, self.rp, self.col, self.h, self.cdf, self.W)


def jacobi_dinv(A):
    d = np.zeros(A.n)
    rc = lib().orc_jacobi_dinv(A.n, A.rp, A.col, A.val, d)
    if rc != 0:
        raise ValueError(f"zero diagonal in row {-rc - 1}")
    return d


def build_domain(A, dinv=None, omega=1.0, threshold=0.0, forward=False):
    """Adjoint domain (rows of H^T), or with forward=True the forward
    method's domain (rows of H)."""
    nnz = int(A.rp[-1])
    cap = max(nnz, 1)
    rp = np.zeros(A.n + 1, np.int64)
    col = np.zeros(cap, np.int32)
    h = np.zeros(cap)
    cdf = np.zeros(cap)
    W = np.zeros(A.n)
    dv = None if dinv is None else np.ascontiguousarray(dinv, np.float64)
    fn = lib().orc_build_domain_forward if forward else lib().orc_build_domain
    got = fn(A.n, A.rp, A.col, A.val, _ptr(dv), omega, threshold, rp, col, h, cdf, W)
    if got < 0:
        raise ValueError({-1: "column out of range", -2: "duplicate column"}.get(got, got))
    return Domain(A.n, rp, col[:got].copy(), h[:got].copy(), cdf[:got].copy(), W)


def mc_params(num_histories, history_length=10, weight_cutoff=0.0, precision=64,
              estimator=EST_COLLISION, source=SRC_RANDOM, seed=0, num_threads=1,
              sampler=SAMPLER_CDF, importance=None):
    """importance: float32 g (n) for the learned proposal (kept referenced)."""
    imp = None
    if importance is not None:
        imp = np.ascontiguousarray(importance, np.float32)
        if sampler != SAMPLER_LEARNED_ALIAS:
            sampler = SAMPLER_LEARNED
    p = MCParams(int(num_histories), int(history_length), float(weight_cutoff),
                 int(precision), int(estimator), int(source), int(seed) & (2**64 - 1),
                 int(num_threads), int(sampler), None if imp is None else imp.ctypes.data)
    p._importance_ref = imp
    return p


def build_proposal(dom, g):
    g = np.ascontiguousarray(g, np.float32)
    lcdf = np.zeros(max(dom.nnz, 1))
    lfac = np.zeros(max(dom.nnz, 1))
    rc = lib().orc_build_proposal(dom.n, dom.rp, dom.col, dom.h, g, lcdf, lfac)
    if rc != 0:
        raise ValueError(f"non-positive importance reached from row {-rc - 1}")
    return lcdf[: dom.nnz], lfac[: dom.nnz]


def build_proposal_alias(dom, g):
    g = np.ascontiguousarray(g, np.float32)
    m = max(dom.nnz, 1)
    thr, alias, lfac = np.zeros(m), np.zeros(m, np.int32), np.zeros(m)
    rc = lib().orc_build_proposal_alias(dom.n, dom.rp, dom.col, dom.h, g, thr, alias, lfac)
    if rc != 0:
        raise ValueError(f"row {-rc - 1}: longer than 64 entries or a bad importance")
    return thr[: dom.nnz], alias[: dom.nnz], lfac[: dom.nnz]


def build_alias(dom):
    thr = np.zeros(max(dom.nnz, 1))
    alias = np.zeros(max(dom.nnz, 1), np.int32)
    rc = lib().orc_build_alias(dom.n, dom.rp, dom.h, dom.W, thr, alias)
    if rc != 0:
        raise ValueError(f"row {-rc - 1} longer than 64 entries")
    return thr[: dom.nnz], alias[: dom.nnz]


def alias_stride(dom):
    """The longest row's length, at least 1 (0 when a row exceeds 64)."""
    return int(lib().orc_alias_stride(dom.n, dom.rp))


def build_alias_strided(dom):
    """Strided alias tables: (stride, thr[n, stride], alias[n, stride])."""
    stride = alias_stride(dom)
    if stride == 0:
        raise ValueError("a row is longer than 64 entries")
    thr = np.zeros(dom.n * stride)
    alias = np.zeros(dom.n * stride, np.int32)
    rc = lib().orc_build_alias_strided(dom.n, dom.rp, dom.h, dom.W, stride, thr, alias)
    if rc != 0:
        raise ValueError(f"strided alias tables failed ({rc})")
    return stride, thr.reshape(dom.n, stride), alias.reshape(dom.n, stride)


def mc_solve(dom, b, params, solve_index=0):
    x = np.zeros(dom.n)
    steps = C.c_int64(0)
    b = np.ascontiguousarray(b, np.float64)
    lib().orc_mc_solve(*dom.args(), b, C.byref(params), int(solve_index), x, C.byref(steps))
    return x, int(steps.value)


def mc_solve_window(dom, b, params, first, count, solve_index=0, mag=None):
    """Histories [first, first+count) of the solve, normalised by its full
    count, no expected-value birth term; births by |mag| when given."""
    x = np.zeros(dom.n)
    steps = C.c_int64(0)
    b = np.ascontiguousarray(b, np.float64)
    m = None if mag is None else np.ascontiguousarray(mag, np.float64)
    lib().orc_mc_solve_window(*dom.args(), b, _ptr(m), C.byref(params), int(solve_index), int(first), int(count), x,
                              C.byref(steps))
    return x, int(steps.value)


def record_paths(dom, b, params, first, count, max_len, solve_index=0):
    paths = np.zeros((count, max_len + 1), np.int32)
    w = np.zeros(count)
    b = np.ascontiguousarray(b, np.float64)
    lib().orc_record_paths(*dom.args(), b, C.byref(params), int(solve_index), int(first),
                           int(count), int(max_len), paths, w)
    return paths, w


def source_quantize(b):
    b = np.ascontiguousarray(b, np.float64)
    q = np.zeros(len(b), np.uint64)
    cdf = np.zeros(len(b), np.uint64)
    Q = lib().orc_source_quantize(len(b), b, q, cdf)
    return int(Q), q, cdf


def solver_params(tol=1e-8, max_iters=1000, check_freq=1, smoother_steps=1,
                  richardson_omega=1.0, neumann_omega=1.0, threshold=0.0, num_sets=1,
                  reference_mode=0, mc=None, ref_quirk=0):
    return SolverParams(float(tol), int(max_iters), int(check_freq), int(smoother_steps),
                        float(richardson_omega), float(neumann_omega), float(threshold),
                        int(num_sets), int(reference_mode), mc if mc is not None else mc_params(1),
                        int(ref_quirk))


def mcsa_solve(A, b, x0=None, dinv=None, params=None, solve_index0=0):
    params = params if params is not None else solver_params()
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    res = SolverResult()
    hist = np.zeros(params.max_iters + 1)
    b = np.ascontiguousarray(b, np.float64)
    dv = None if dinv is None else np.ascontiguousarray(dinv, np.float64)
    rc = lib().orc_mcsa_solve(A.n, A.rp, A.col, A.val, _ptr(dv), b, x, C.byref(params),
                              int(solve_index0), C.byref(res), _ptr(hist))
    if rc != 0:
        raise ValueError(f"oracle mcsa failed: {rc}")
    return x, res, hist[: min(res.iters, params.max_iters) + 1]


def fixed_point_solve(A, b, x0=None, dinv=None, params=None):
    params = params if params is not None else solver_params()
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    res = SolverResult()
    b = np.ascontiguousarray(b, np.float64)
    dv = None if dinv is None else np.ascontiguousarray(dinv, np.float64)
    lib().orc_fixed_point_solve(A.n, A.rp, A.col, A.val, _ptr(dv), b, x, C.byref(params),
                                C.byref(res))
    return x, res


def spmv(A, x):
    y = np.zeros(A.n)
    lib().orc_spmv(A.n, A.rp, A.col, A.val, np.ascontiguousarray(x, np.float64), y)
    return y


def prec_residual(A, b, x, dinv=None):
    r = np.zeros(A.n)
    dv = None if dinv is None else np.ascontiguousarray(dinv, np.float64)
    lib().orc_prec_residual(A.n, A.rp, A.col, A.val, _ptr(dv),
                            np.ascontiguousarray(b, np.float64),
                            np.ascontiguousarray(x, np.float64), r)
    return r


def ref_build_source(b, nh_total, stratified=False, ref_quirk=False, rng_state=1):
    b = np.ascontiguousarray(b, np.float64)
    counts = np.zeros(len(b), np.int64)
    st = C.c_uint64(rng_state)
    nh = lib().orc_ref_build_source(len(b), b, float(nh_total), int(stratified),
                                    int(ref_quirk), C.byref(st), counts)
    return int(nh), counts


def ref_transport(dom, state, weight, history_length, weight_cutoff=0.0, rng_state=1, x=None):
    x = np.zeros(dom.n) if x is None else x
    st = C.c_uint64(rng_state)
    s = C.c_int(state)
    w = C.c_double(weight)
    steps = C.c_int64(0)
    ev = lib().orc_ref_transport(dom.rp, dom.col, dom.h, dom.cdf, dom.W, int(history_length),
                                 float(weight_cutoff), C.byref(st), C.byref(s), C.byref(w),
                                 C.byref(steps), x)
    return ev, s.value, w.value, steps.value, x


def ref_mc_solve(dom, b, params, stratified=False, ref_quirk=False, rng_state=1):
    x = np.zeros(dom.n)
    st = C.c_uint64(rng_state)
    steps = C.c_int64(0)
    nh = lib().orc_ref_mc_solve(*dom.args(), np.ascontiguousarray(b, np.float64),
                                C.byref(params), int(stratified), int(ref_quirk), C.byref(st),
                                x, C.byref(steps))
    return x, int(nh), int(steps.value)


# ------------------------------------------------------ composite operator --
def transpose(A):
    """MT::copyTranspose: rows in ascending source row."""
    nnz = int(A.rp[-1])
    t_rp = np.zeros(A.n + 1, np.int64)
    t_col = np.zeros(max(nnz, 1), np.int32)
    t_val = np.zeros(max(nnz, 1))
    lib().orc_transpose(A.n, A.rp, A.col, A.val, t_rp, t_col, t_val)
    return Csr(t_rp, t_col[:nnz], t_val[:nnz], A.n)


def spgemm(X, Y, threshold=0.0):
    """MT::multiply(X, false, Y, false, threshold): Gustavson product in the
    order of X's row entries, then the row filter (mcls_oracle.h)."""
    ylen = np.diff(Y.rp)
    cap = int(ylen[X.col].sum()) if len(X.col) else 0
    c_rp = np.zeros(X.n + 1, np.int64)
    c_col = np.zeros(max(cap, 1), np.int32)
    c_val = np.zeros(max(cap, 1))
    got = lib().orc_spgemm(X.n, X.rp, X.col, X.val, Y.rp, Y.col, Y.val, float(threshold), cap,
                           c_rp, c_col, c_val)
    if got < 0:
        raise RuntimeError("spgemm capacity")
    return Csr(c_rp, c_col[:got], c_val[:got], X.n)


def composite(A, PL=None, PR=None, threshold=0.0, forward=False):
    """LinearProblem::getCompositeOperator (forward) and
    getTransposeCompositeOperator (adjoint), src/MCLS_LinearProblem_impl.hpp:139-221:
    the same products in the same order, each filtered.  Without a
    preconditioner: A (forward) or A^T (adjoint), unfiltered."""
    if forward:
        if PL is not None and PR is not None:
            return spgemm(PL, spgemm(A, PR, threshold), threshold)
        if PL is not None:
            return spgemm(PL, A, threshold)
        if PR is not None:
            return spgemm(A, PR, threshold)
        return A
    AT = transpose(A)
    if PL is not None and PR is not None:
        return spgemm(transpose(PR), spgemm(AT, transpose(PL), threshold), threshold)
    if PR is not None:
        return spgemm(transpose(PR), AT, threshold)
    if PL is not None:
        return spgemm(AT, transpose(PL), threshold)
    return AT


def diag_csr(d):
    n = len(d)
    return Csr(np.arange(n + 1), np.arange(n), np.asarray(d, np.float64), n)


def build_domain_general(A, PL=None, PR=None, omega=1.0, threshold=0.0, forward=False):
    """Domain rows from the composite operator with general (sparse) left and
    right preconditioners: the rows of the (transposed, for the adjoint
    method) composite, with h = delta - omega m and the DBL_EPSILON drop
    (src/MCLS_AlmostOptimalDomain_impl.hpp:296-364); the composite's filter
    has already run inside its products."""
    if PL is None and PR is None:
        return build_domain(A, None, omega, threshold, forward)
    Cm = composite(A, PL, PR, threshold, forward)
    return build_domain(Cm, None, omega, 0.0, forward=True)


def norm2(v):
    """Sequential 2-norm (orc_norm2), as the C solvers compute it."""
    v = np.ascontiguousarray(v, np.float64)
    return float(lib().orc_norm2(len(v), v))


class PrecSystem:
    """A with a left preconditioner (sparse PL, or diagonal dinv) and an
    optional sparse right preconditioner PR (src/MCLS_LinearProblem_impl.hpp)."""

    def __init__(self, A, PL=None, PR=None, dinv=None):
        self.A, self.PL, self.PR = A, PL, PR
        self.dinv = None if dinv is None else np.asarray(dinv, np.float64)

    def left(self, v):
        if self.PL is not None:
            return spmv(self.PL, v)
        return v * self.dinv if self.dinv is not None else np.array(v, np.float64)

    def right(self, v):
        return spmv(self.PR, v) if self.PR is not None else np.array(v, np.float64)

    def prec_residual(self, b, y):
        """updatePrecResidual (:411-442): PL (b - A PR y)."""
        return self.left(np.asarray(b, np.float64) - spmv(self.A, self.right(y)))

    def left_csr(self):
        if self.PL is not None:
            return self.PL
        return None if self.dinv is None else diag_csr(self.dinv)

    def domain(self, omega=1.0, threshold=0.0, forward=False):
        return build_domain_general(self.A, self.left_csr(), self.PR, omega, threshold, forward)


def mc_solve_general(sys, b, params, omega=1.0, threshold=0.0, forward=False, solve_index=0):
    """MonteCarloSolverManager::solve with preconditioners
    (src/MCLS_MonteCarloSolverManager_impl.hpp:200-233, 266-380):
    source omega PL b, walk over the composite's domain, then x = PR x."""
    dom = sys.domain(omega, threshold, forward)
    x, steps = mc_solve(dom, sys.left(b) * omega, params, solve_index)
    return sys.right(x), steps


def mcsa_solve_general(sys, b, x0=None, params=None, solve_index0=0, forward=False):
    """MCSASolverManager::solve (src/MCLS_MCSASolverManager_impl.hpp:280-420)
    with general preconditioners: iterate on y (x = PR y), residual
    PL (b - A PR y), MC correction over the composite's domain with the
    internal solver (no PL on its source, no PR on its result), and the
    final recovery x = PR y.  Same stream indices as mcsa_solve."""
    params = params if params is not None else solver_params()
    y = np.zeros(sys.A.n) if x0 is None else np.array(x0, np.float64)
    b = np.asarray(b, np.float64)
    dom = sys.domain(params.neumann_omega, params.threshold, forward)
    source_norm = norm2(sys.left(b))
    if abs(source_norm) < 10.0 * np.finfo(float).eps:
        source_norm = 1.0
    crit = params.tol * source_norm
    r = sys.prec_residual(b, y)
    rnorm = norm2(r)
    hist = [rnorm / source_norm]
    it = 0
    go = True
    while go:
        it += 1
        for _ in range(params.smoother_steps):
            y = y + params.richardson_omega * r
            r = sys.prec_residual(b, y)
        dx, _ = mc_solve(dom, r * params.neumann_omega, params.mc, solve_index0 + it - 1)
        y = y + dx
        r = sys.prec_residual(b, y)
        rnorm = norm2(r)
        hist.append(rnorm / source_norm)
        if it % max(params.check_freq, 1) == 0:
            go = rnorm > crit and it < params.max_iters
    return sys.right(y), {"iters": it, "converged": rnorm <= crit, "achieved_tol": rnorm / source_norm}, hist
