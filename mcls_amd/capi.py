"""ctypes bindings to the C ABI of the gfx950 solver (include/mcls_gpu.h) and
of the workload generators (include/mcls_problems.h).

This is plumbing for the tests and the benchmark: every computation happens
in libmcls_gpu.so.  Loading fails loudly when the library has not been built;
there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
GPU_LIB = os.path.join(LIB_DIR, "libmcls_gpu.so")
PROBLEMS_LIB = os.path.join(LIB_DIR, "libmcls_problems.so")

MCLS_OK = 0
MCLS_ERR_INVALID_ARG = 1
MCLS_ERR_HIP = 2
MCLS_ERR_MATRIX = 3
MCLS_ERR_STATE = 4
MCLS_ERR_UNSUPPORTED = 5
MCLS_ERR_NUMERIC = 6
MCLS_ERR_NO_DEVICE = 7
MCLS_F64, MCLS_F32 = 0, 1
MCLS_HOST, MCLS_DEVICE = 0, 1

# every symbol include/mcls_gpu.h declares
GPU_SYMBOLS = [
    "mcls_gpu_create", "mcls_gpu_destroy", "mcls_gpu_last_error", "mcls_gpu_synchronize",
    "mcls_gpu_set_matrix_csr", "mcls_gpu_set_left_prec_diag",
    "mcls_gpu_set_left_prec_point_jacobi", "mcls_gpu_clear_left_prec",
    "mcls_gpu_get_left_prec_diag", "mcls_gpu_set_param_i64", "mcls_gpu_set_param_f64",
    "mcls_gpu_set_param_str", "mcls_gpu_build_domain", "mcls_gpu_domain_nnz",
    "mcls_gpu_get_domain", "mcls_gpu_mc_solve", "mcls_gpu_mcsa_solve",
    "mcls_gpu_prec_residual", "mcls_gpu_get_stats", "mcls_gpu_record_paths",
    "mcls_gpu_mc_solve_raw", "mcls_gpu_set_systems", "mcls_gpu_get_system_residuals",
    "mcls_gpu_autotune", "mcls_gpu_convergence_criteria",
    "mcls_gpu_set_importance", "mcls_gpu_set_importance_diagonal", "mcls_gpu_get_importance",
    "mcls_gpu_set_importance_model",
    "mcls_gpu_set_importance_mask", "mcls_gpu_importance_features", "mcls_gpu_importance_targets",
    "mcls_gpu_set_left_prec_matrix", "mcls_gpu_set_right_prec_matrix", "mcls_gpu_clear_right_prec",
    "mcls_gpu_set_state_order", "mcls_gpu_mc_solve_noalloc", "mcls_gpu_get_residual_history",
    "mcls_gpu_mc_solve_window",
]
MCLS_NOALLOC_MAX_N = 1048576
MLP_SYMBOLS = [
    "mcls_mlp_create", "mcls_mlp_destroy", "mcls_mlp_get_params", "mcls_mlp_set_params",
    "mcls_mlp_forward", "mcls_mlp_loss_grad", "mcls_mlp_dataset_clear", "mcls_mlp_dataset_append",
    "mcls_mlp_dataset_rows", "mcls_mlp_train",
]
MLP_PARAMS = 8961
PROBLEM_SYMBOLS = ["mcls_problem_dims", "mcls_problem_build", "mcls_problem_batch_member",
                   "mcls_problem_uniform"]

POISSON2D_5PT, DIFFUSION3D_7PT, ANISO2D_5PT, VARCOEF2D_5PT, STENCIL3D_27PT = 1, 2, 3, 4, 5


class GpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("converged", C.c_int64),
        ("achieved_tol", C.c_double),
        ("source_norm", C.c_double),
        ("histories", C.c_int64),
        ("steps", C.c_int64),
        ("walk_launches", C.c_int64),
        ("ms_total", C.c_double),
        ("ms_walk", C.c_double),
        ("ms_source", C.c_double),
        ("ms_residual", C.c_double),
        ("ms_domain", C.c_double),
        ("domain_nnz", C.c_int64),
        ("solve_index", C.c_int64),
        ("last_walk_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [
        ("block", C.c_int64),
        ("blocks_per_cu", C.c_int64),
        ("batch", C.c_int64),
        ("tally_tile", C.c_int64),
        ("ms_default", C.c_double),
        ("ms_best", C.c_double),
        ("candidates", C.c_int64),
        ("ms_confirmed", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Criteria(C.Structure):
    _fields_ = [
        ("rho_h", C.c_double),
        ("rho_habs", C.c_double),
        ("rho_hstar", C.c_double),
        ("inf_norm", C.c_double),
        ("min_row_sum", C.c_double),
        ("balance_min", C.c_double),
        ("balance_max", C.c_double),
        ("balance_ave", C.c_double),
        ("iterations", C.c_int64 * 3),
        ("converged", C.c_int64 * 3),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_[:8]}
        d["iterations"] = list(self.iterations)
        d["converged"] = list(self.converged)
        return d


class ProblemSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("seed", C.c_uint64),
        ("p0", C.c_double),
        ("p1", C.c_double),
        ("p2", C.c_double),
        ("p3", C.c_double),
    ]


_gpu = None
_prob = None


def gpu_lib():
    global _gpu
    if _gpu is not None:
        return _gpu
    if not os.path.exists(GPU_LIB):
        raise RuntimeError(f"{GPU_LIB} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(GPU_LIB)
    vp, i64, u64, i32, f64 = C.c_void_p, C.c_int64, C.c_uint64, C.c_int, C.c_double
    L.mcls_gpu_create.argtypes = [C.POINTER(vp), i32, C.c_uint]
    L.mcls_gpu_destroy.argtypes = [vp]
    L.mcls_gpu_last_error.restype = C.c_char_p
    L.mcls_gpu_synchronize.argtypes = [vp]
    L.mcls_gpu_set_matrix_csr.argtypes = [vp, i64, i64, vp, vp, vp, i32, i32]
    L.mcls_gpu_set_state_order.argtypes = [vp, i64, vp, i32]
    L.mcls_gpu_set_left_prec_diag.argtypes = [vp, vp, i32]
    L.mcls_gpu_set_left_prec_point_jacobi.argtypes = [vp]
    L.mcls_gpu_clear_left_prec.argtypes = [vp]
    L.mcls_gpu_get_left_prec_diag.argtypes = [vp, vp, i32]
    L.mcls_gpu_set_left_prec_matrix.argtypes = [vp, vp, vp, vp, i32]
    L.mcls_gpu_set_right_prec_matrix.argtypes = [vp, vp, vp, vp, i32]
    L.mcls_gpu_clear_right_prec.argtypes = [vp]
    L.mcls_gpu_set_param_i64.argtypes = [vp, C.c_char_p, i64]
    L.mcls_gpu_set_param_f64.argtypes = [vp, C.c_char_p, f64]
    L.mcls_gpu_set_param_str.argtypes = [vp, C.c_char_p, C.c_char_p]
    L.mcls_gpu_build_domain.argtypes = [vp]
    L.mcls_gpu_domain_nnz.argtypes = [vp, C.POINTER(i64)]
    L.mcls_gpu_get_domain.argtypes = [vp, vp, vp, vp, vp, vp]
    L.mcls_gpu_mc_solve.argtypes = [vp, vp, vp, i32]
    L.mcls_gpu_mcsa_solve.argtypes = [vp, vp, vp, i32]
    L.mcls_gpu_mc_solve_noalloc.argtypes = [vp, vp, vp]
    L.mcls_gpu_get_residual_history.argtypes = [vp, vp, i64, C.POINTER(i64)]
    L.mcls_gpu_mc_solve_window.argtypes = [vp, vp, u64, i64, i64, vp]
    L.mcls_gpu_prec_residual.argtypes = [vp, vp, vp, vp, C.POINTER(f64), i32]
    L.mcls_gpu_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.mcls_gpu_record_paths.argtypes = [vp, vp, u64, i64, i64, C.c_int32, vp, vp]
    L.mcls_gpu_mc_solve_raw.argtypes = [vp, vp, u64, vp]
    L.mcls_gpu_set_systems.argtypes = [vp, i64, vp]
    L.mcls_gpu_get_system_residuals.argtypes = [vp, vp]
    L.mcls_gpu_set_importance.argtypes = [vp, vp, i32]
    L.mcls_gpu_set_importance_diagonal.argtypes = [vp]
    L.mcls_gpu_autotune.argtypes = [vp, vp, i32, C.c_int64, C.POINTER(Tuning)]
    L.mcls_gpu_convergence_criteria.argtypes = [vp, C.c_int64, C.c_double, C.POINTER(Criteria)]
    L.mcls_gpu_get_importance.argtypes = [vp, vp, i32]
    L.mcls_gpu_set_importance_model.argtypes = [vp, vp]
    L.mcls_gpu_set_importance_mask.argtypes = [vp, vp, i32]
    L.mcls_gpu_importance_features.argtypes = [vp, vp, vp, i32]
    L.mcls_gpu_importance_targets.argtypes = [vp, vp, u64, vp, i32]
    L.mcls_mlp_create.argtypes = [C.POINTER(vp), i32, u64]
    L.mcls_mlp_destroy.argtypes = [vp]
    L.mcls_mlp_get_params.argtypes = [vp, vp]
    L.mcls_mlp_set_params.argtypes = [vp, vp]
    L.mcls_mlp_forward.argtypes = [vp, vp, i64, vp, i32]
    L.mcls_mlp_loss_grad.argtypes = [vp, vp, vp, i64, C.POINTER(f64), vp, i32]
    L.mcls_mlp_dataset_clear.argtypes = [vp]
    L.mcls_mlp_dataset_append.argtypes = [vp, vp, vp, i64, i32]
    L.mcls_mlp_dataset_rows.argtypes = [vp, C.POINTER(i64)]
    L.mcls_mlp_train.argtypes = [vp, i32, i64, C.c_float, u64, C.POINTER(f64)]
    for name in GPU_SYMBOLS + MLP_SYMBOLS:
        f = getattr(L, name)
        if name != "mcls_gpu_last_error":
            f.restype = i32
    _gpu = L
    return L


def problems_lib():
    global _prob
    if _prob is not None:
        return _prob
    if not os.path.exists(PROBLEMS_LIB):
        raise RuntimeError(f"{PROBLEMS_LIB} is missing: build it first")
    L = C.CDLL(PROBLEMS_LIB)
    vp = C.c_void_p
    L.mcls_problem_dims.argtypes = [C.POINTER(ProblemSpec), C.POINTER(C.c_int64)]
    L.mcls_problem_dims.restype = C.c_int64
    L.mcls_problem_build.argtypes = [C.POINTER(ProblemSpec), vp, vp, vp, vp, vp, C.c_int]
    L.mcls_problem_build.restype = C.c_int
    L.mcls_problem_batch_member.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32,
                                            C.c_double, C.POINTER(ProblemSpec)]
    L.mcls_problem_batch_member.restype = C.c_int
    L.mcls_problem_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    L.mcls_problem_uniform.restype = C.c_double
    _prob = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_hip = None


def hip_runtime():
    """The HIP runtime the solver library is linked against (loaded with it),
    for plain device buffers; nothing else in this package allocates device
    memory from Python."""
    global _hip
    if _hip is None:
        gpu_lib()  # loads libamdhip64.so.7 as the library's dependency
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


class DeviceArray:
    """n doubles in device memory (hipMalloc), for solves whose vectors stay
    resident on the GPU."""

    _H2D, _D2H = 1, 2

    def __init__(self, n):
        self.L = hip_runtime()
        self.n = int(n)
        self.ptr = C.c_void_p()
        if self.L.hipMalloc(C.byref(self.ptr), 8 * max(self.n, 1)) != 0:
            raise GpuError(MCLS_ERR_HIP, "hipMalloc failed")

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        if self.L.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), 8 * self.n, self._H2D) != 0:
            raise GpuError(MCLS_ERR_HIP, "hipMemcpy to device failed")

    def download(self):
        a = np.empty(self.n)
        if self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, 8 * self.n, self._D2H) != 0:
            raise GpuError(MCLS_ERR_HIP, "hipMemcpy to host failed")
        return a

    def zero(self):
        if self.L.hipMemset(self.ptr, 0, 8 * self.n) != 0 or self.L.hipDeviceSynchronize() != 0:
            raise GpuError(MCLS_ERR_HIP, "hipMemset failed")

    def close(self):
        if self.ptr:
            self.L.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Problem:
    """A generated system: CSR (int64 rp, int32 col, f64 and/or f32 values), b."""

    def __init__(self, spec, rp, col, val64, val32, b):
        self.spec, self.rp, self.col, self.val64, self.val32, self.b = spec, rp, col, val64, val32, b

    @property
    def n(self):
        return len(self.rp) - 1

    @property
    def nnz(self):
        return int(self.rp[-1])

    def values(self):
        return self.val64 if self.val64 is not None else self.val32.astype(np.float64)


def make_spec(kind, nx, ny=0, nz=0, seed=0, p0=0.0, p1=0.0, p2=0.0, p3=0.0):
    return ProblemSpec(kind, nx, ny, nz, seed & (2**64 - 1), p0, p1, p2, p3)


def brick_order(nx, ny, nz, bx, by, bz):
    """State order of an nx x ny x nz grid (unknown x + nx (y + ny z))
    visited brick by brick: bricks of bx x by x bz in x-fastest order, the
    points of a brick x-fastest inside it.  Bricks at the high faces are cut
    to fit.  Returns order[k] = the unknown stored k-th (int32)."""
    if min(nx, ny, nz, bx, by, bz) < 1:
        raise ValueError("grid and brick sizes must be >= 1")
    x = np.arange(nx, dtype=np.int64)
    y = np.arange(ny, dtype=np.int64)
    z = np.arange(nz, dtype=np.int64)
    # sort key: (brick z, brick y, brick x, z, y, x) of every unknown
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    nbx, nby = -(-nx // bx), -(-ny // by)
    brick = (Z // bz * nby + Y // by) * nbx + X // bx
    key = np.lexsort(((X % bx).ravel(), (Y % by).ravel(), (Z % bz).ravel(), brick.ravel()))
    return key.astype(np.int32)


def nested_brick_order(nx, ny, nz, levels):
    """State order of an nx x ny x nz grid visited brick by brick on several
    levels: `levels` lists brick sizes (bx, by, bz) from the coarsest to the
    finest, each dividing the one before it; the grid is cut into bricks of
    the first size in x-fastest order, each of those into bricks of the next
    size in x-fastest order, and so on, the points of a finest brick
    x-fastest.  One level is brick_order.  Returns order[k] = the unknown
    stored k-th (int32)."""
    levels = [tuple(int(v) for v in lv) for lv in levels]
    if not levels or min(nx, ny, nz) < 1 or any(min(lv) < 1 for lv in levels):
        raise ValueError("grid and brick sizes must be >= 1")
    for outer, inner in zip(levels, levels[1:]):
        if any(o % i for o, i in zip(outer, inner)):
            raise ValueError("each brick size must divide the one before it")
    Z, Y, X = np.meshgrid(np.arange(nz, dtype=np.int64), np.arange(ny, dtype=np.int64),
                          np.arange(nx, dtype=np.int64), indexing="ij")
    keys = []  # most significant first
    nb = [-(-nz // levels[0][2]), -(-ny // levels[0][1]), -(-nx // levels[0][0])]
    keys.append(((Z // levels[0][2]) * nb[1] + Y // levels[0][1]) * nb[2] + X // levels[0][0])
    for outer, inner in zip(levels, levels[1:]):
        keys += [(Z % outer[2]) // inner[2], (Y % outer[1]) // inner[1], (X % outer[0]) // inner[0]]
    last = levels[-1]
    keys += [Z % last[2], Y % last[1], X % last[0]]
    return np.lexsort([k.ravel() for k in reversed(keys)]).astype(np.int32)


def build_problem(spec, want64=True, want32=False, nthreads=None):
    L = problems_lib()
    n = C.c_int64(0)
    nnz = L.mcls_problem_dims(C.byref(spec), C.byref(n))
    if nnz < 0:
        raise ValueError("bad problem spec")
    n = n.value
    rp = np.empty(n + 1, np.int64)
    col = np.empty(nnz, np.int32)
    v64 = np.empty(nnz, np.float64) if want64 else None
    v32 = np.empty(nnz, np.float32) if want32 else None
    b = np.empty(n, np.float64)
    nt = nthreads if nthreads is not None else min(16, os.cpu_count() or 1)
    rc = L.mcls_problem_build(C.byref(spec), _p(rp), _p(col), _p(v64), _p(v32), _p(b), nt)
    if rc != 0:
        raise ValueError("problem build failed")
    return Problem(spec, rp, col, v64, v32, b)


def batch_member(seed, index, lo=32, hi=256, eps_hi=100.0):
    spec = ProblemSpec()
    if problems_lib().mcls_problem_batch_member(seed, index, lo, hi, eps_hi, C.byref(spec)) != 0:
        raise ValueError("bad batch member arguments")
    return spec


class Context:
    """One mcls_gpu_context (one stream, one linear problem)."""

    def __init__(self, device=0):
        self.L = gpu_lib()
        self.h = C.c_void_p()
        self._check(self.L.mcls_gpu_create(C.byref(self.h), device, 0))
        self.n = 0
        self.nsys = 0

    def _check(self, rc):
        if rc != MCLS_OK:
            raise GpuError(rc, self.L.mcls_gpu_last_error().decode())
        return rc

    def close(self):
        if self.h:
            self.L.mcls_gpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- problem
    def set_matrix(self, rp, col, val):
        rp = np.ascontiguousarray(rp, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val)
        dt = MCLS_F32 if val.dtype == np.float32 else MCLS_F64
        if dt == MCLS_F64:
            val = val.astype(np.float64, copy=False)
        n = len(rp) - 1
        self._check(self.L.mcls_gpu_set_matrix_csr(self.h, n, int(rp[-1]), _p(rp), _p(col), _p(val),
                                                   dt, MCLS_HOST))
        self.n = n

    def set_state_order(self, order):
        """Number the unknowns in `order` (internal state k = unknown
        order[k]); None returns to the caller's numbering.  Drops the
        operator: call before set_matrix / set_problem."""
        if order is None:
            self._check(self.L.mcls_gpu_set_state_order(self.h, 0, None, MCLS_HOST))
            return
        order = np.ascontiguousarray(order, np.int32)
        self._order_keep = order
        self._check(self.L.mcls_gpu_set_state_order(self.h, len(order), _p(order), MCLS_HOST))

    def set_problem(self, prob, dtype="f64"):
        val = prob.val32 if dtype == "f32" else prob.values()
        self.set_matrix(prob.rp, prob.col, val)

    def set_point_jacobi(self):
        self._check(self.L.mcls_gpu_set_left_prec_point_jacobi(self.h))

    def set_left_prec_diag(self, dinv):
        dinv = np.ascontiguousarray(dinv, np.float64)
        self._check(self.L.mcls_gpu_set_left_prec_diag(self.h, _p(dinv), MCLS_HOST))

    def clear_left_prec(self):
        self._check(self.L.mcls_gpu_clear_left_prec(self.h))

    @staticmethod
    def _prec_arrays(rp, col, val):
        return (np.ascontiguousarray(rp, np.int64), np.ascontiguousarray(col, np.int32),
                np.ascontiguousarray(val, np.float64))

    def set_left_prec_matrix(self, rp, col, val):
        """LinearProblem::setLeftPrec with a sparse PL (n x n CSR, ascending rows)."""
        rp, col, val = self._prec_arrays(rp, col, val)
        self._check(self.L.mcls_gpu_set_left_prec_matrix(self.h, _p(rp), _p(col), _p(val), MCLS_HOST))

    def set_right_prec_matrix(self, rp, col, val):
        """LinearProblem::setRightPrec (n x n CSR, ascending rows)."""
        rp, col, val = self._prec_arrays(rp, col, val)
        self._check(self.L.mcls_gpu_set_right_prec_matrix(self.h, _p(rp), _p(col), _p(val), MCLS_HOST))

    def clear_right_prec(self):
        self._check(self.L.mcls_gpu_clear_right_prec(self.h))

    def get_left_prec_diag(self):
        d = np.empty(self.n)
        self._check(self.L.mcls_gpu_get_left_prec_diag(self.h, _p(d), MCLS_HOST))
        return d

    def set_params(self, params):
        for k, v in params.items():
            kb = k.encode()
            if isinstance(v, str):
                self._check(self.L.mcls_gpu_set_param_str(self.h, kb, v.encode()))
            elif isinstance(v, (bool, int, np.integer)):
                self._check(self.L.mcls_gpu_set_param_i64(self.h, kb, int(v)))
            else:
                self._check(self.L.mcls_gpu_set_param_f64(self.h, kb, float(v)))

    # --- domain
    def build_domain(self):
        self._check(self.L.mcls_gpu_build_domain(self.h))

    def domain(self):
        nnz = C.c_int64(0)
        self._check(self.L.mcls_gpu_domain_nnz(self.h, C.byref(nnz)))
        m = nnz.value
        rp = np.empty(self.n + 1, np.int64)
        col = np.empty(max(m, 1), np.int32)
        h = np.empty(max(m, 1))
        cdf = np.empty(max(m, 1))
        W = np.empty(self.n)
        self._check(self.L.mcls_gpu_get_domain(self.h, _p(rp), _p(col), _p(h), _p(cdf), _p(W)))
        return rp, col[:m], h[:m], cdf[:m], W

    # --- solves
    def mc_solve(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty(self.n)
        self._check(self.L.mcls_gpu_mc_solve(self.h, _p(b), _p(x), MCLS_HOST))
        return x

    def mc_solve_raw(self, rhs, solve_index):
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.empty(self.n)
        self._check(self.L.mcls_gpu_mc_solve_raw(self.h, _p(rhs), int(solve_index), _p(x)))
        return x

    def mc_solve_window(self, rhs, solve_index, first, count):
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.empty(self.n)
        self._check(self.L.mcls_gpu_mc_solve_window(self.h, _p(rhs), int(solve_index), int(first), int(count), _p(x)))
        return x

    def mcsa_solve(self, b, x0=None):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64)
        self._check(self.L.mcls_gpu_mcsa_solve(self.h, _p(b), _p(x), MCLS_HOST))
        return x

    def mc_solve_noalloc(self, b_ptr, x_ptr):
        """mcls_gpu_mc_solve_noalloc: b, x are device pointers (n doubles)."""
        self._check(self.L.mcls_gpu_mc_solve_noalloc(self.h, b_ptr, x_ptr))

    def mcsa_solve_device(self, b_ptr, x_ptr):
        """MCSA with b and x (initial guess in, solution out) already in
        device memory: raw pointers to n doubles each (e.g. a torch tensor's
        data_ptr(), written before this call)."""
        self._check(self.L.mcls_gpu_mcsa_solve(self.h, C.c_void_p(int(b_ptr)), C.c_void_p(int(x_ptr)),
                                               MCLS_DEVICE))

    def prec_residual(self, b, x):
        b = np.ascontiguousarray(b, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        r = np.empty(self.n)
        nrm = C.c_double(0)
        self._check(self.L.mcls_gpu_prec_residual(self.h, _p(b), _p(x), _p(r), C.byref(nrm), MCLS_HOST))
        return r, nrm.value

    def record_paths(self, rhs, solve_index, first, count, max_len):
        rhs = np.ascontiguousarray(rhs, np.float64)
        paths = np.empty((count, max_len + 1), np.int32)
        w = np.empty(count)
        self._check(self.L.mcls_gpu_record_paths(self.h, _p(rhs), int(solve_index), int(first),
                                                 int(count), int(max_len), _p(paths), _p(w)))
        return paths, w

    def set_systems(self, offsets):
        """Block-diagonal batch: system s on rows [offsets[s], offsets[s+1])."""
        if offsets is None or len(offsets) == 0:
            self._check(self.L.mcls_gpu_set_systems(self.h, 0, None))
            self.nsys = 0
            return
        off = np.ascontiguousarray(offsets, np.int64)
        self._check(self.L.mcls_gpu_set_systems(self.h, len(off) - 1, _p(off)))
        self.nsys = len(off) - 1

    def system_residuals(self):
        out = np.empty(self.nsys)
        self._check(self.L.mcls_gpu_get_system_residuals(self.h, _p(out)))
        return out

    # --- learned proposal
    def set_importance(self, g):
        g = np.ascontiguousarray(g, np.float32)
        self._check(self.L.mcls_gpu_set_importance(self.h, _p(g), MCLS_HOST))

    def importance(self):
        g = np.empty(self.n, np.float32)
        self._check(self.L.mcls_gpu_get_importance(self.h, _p(g), MCLS_HOST))
        return g

    def autotune(self, b, pilot_histories=0):
        b = np.ascontiguousarray(b, np.float64)
        t = Tuning()
        self._check(self.L.mcls_gpu_autotune(self.h, _p(b), MCLS_HOST, int(pilot_histories), C.byref(t)))
        return t.as_dict()

    def convergence_criteria(self, max_iters=5000, tol=1e-6):
        c = Criteria()
        self._check(self.L.mcls_gpu_convergence_criteria(self.h, int(max_iters), float(tol), C.byref(c)))
        return c.as_dict()

    def set_importance_diagonal(self):
        self._check(self.L.mcls_gpu_set_importance_diagonal(self.h))

    def set_importance_model(self, mlp):
        self._model = mlp
        self._check(self.L.mcls_gpu_set_importance_model(self.h, mlp.h if mlp is not None else None))

    def set_importance_mask(self, mask):
        if mask is None:
            self._check(self.L.mcls_gpu_set_importance_mask(self.h, None, MCLS_HOST))
            return
        m = np.ascontiguousarray(mask, np.int8)
        self._check(self.L.mcls_gpu_set_importance_mask(self.h, _p(m), MCLS_HOST))

    def importance_features(self, v):
        v = np.ascontiguousarray(v, np.float64)
        X = np.empty((self.n, 8), np.float32)
        self._check(self.L.mcls_gpu_importance_features(self.h, _p(v), _p(X), MCLS_HOST))
        return X

    def importance_targets(self, v, solve_index):
        v = np.ascontiguousarray(v, np.float64)
        T = np.empty(self.n, np.float32)
        self._check(self.L.mcls_gpu_importance_targets(self.h, _p(v), int(solve_index), _p(T), MCLS_HOST))
        return T

    def residual_history(self):
        """||rp|| / ||PL b|| before and after each iteration of the last MCSA solve."""
        m = C.c_int64(0)
        self._check(self.L.mcls_gpu_get_residual_history(self.h, None, 0, C.byref(m)))
        h = np.empty(max(m.value, 1))
        self._check(self.L.mcls_gpu_get_residual_history(self.h, _p(h), m.value, C.byref(m)))
        return h[: m.value]

    def stats(self):
        s = Stats()
        self._check(self.L.mcls_gpu_get_stats(self.h, C.byref(s)))
        return s

    def synchronize(self):
        self._check(self.L.mcls_gpu_synchronize(self.h))


class Mlp:
    """The importance network (include/mcls_mlp.h)."""

    def __init__(self, device=0, seed=0):
        self.L = gpu_lib()
        self.h = C.c_void_p()
        self._check(self.L.mcls_mlp_create(C.byref(self.h), device, seed & (2**64 - 1)))

    def _check(self, rc):
        if rc != MCLS_OK:
            raise GpuError(rc, self.L.mcls_gpu_last_error().decode())

    def close(self):
        if self.h:
            self.L.mcls_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(MLP_PARAMS, np.float32)
        self._check(self.L.mcls_mlp_get_params(self.h, _p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_PARAMS
        self._check(self.L.mcls_mlp_set_params(self.h, _p(p)))

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty(len(X), np.float32)
        self._check(self.L.mcls_mlp_forward(self.h, _p(X), len(X), _p(out), MCLS_HOST))
        return out

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_double(0)
        self._check(self.L.mcls_mlp_loss_grad(self.h, _p(X), _p(T), len(X), C.byref(loss), _p(g), MCLS_HOST))
        return loss.value, g

    def dataset_clear(self):
        self._check(self.L.mcls_mlp_dataset_clear(self.h))

    def dataset_append(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        self._check(self.L.mcls_mlp_dataset_append(self.h, _p(X), _p(T), len(X), MCLS_HOST))

    def dataset_rows(self):
        r = C.c_int64(0)
        self._check(self.L.mcls_mlp_dataset_rows(self.h, C.byref(r)))
        return r.value

    def train(self, epochs, batch_rows, lr, seed=0):
        loss = C.c_double(0)
        self._check(self.L.mcls_mlp_train(self.h, int(epochs), int(batch_rows), C.c_float(lr),
                                          seed & (2**64 - 1), C.byref(loss)))
        return loss.value
