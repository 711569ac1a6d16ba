"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

The oracle restates Alien's SimpleCSR Krylov path on the CPU (see
alien_oracle.h). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

ORC_CG, ORC_BICGS = 0, 1
PREC_DIAG, PREC_CHEB, PREC_ILU0, PREC_NONE, PREC_FILU0, PREC_NEUMANN, PREC_FACTOR = 0, 1, 2, 3, 4, 5, 6

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_u8 = C.POINTER(C.c_uint8)


class SolveParams(C.Structure):
    _fields_ = [("solver", C.c_int), ("precond", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int),
                ("poly_factor", C.c_double), ("poly_order", C.c_int), ("poly_max_iter", C.c_int),
                ("nthreads", C.c_int), ("filu_factor_niter", C.c_int), ("filu_solver_niter", C.c_int),
                ("g_ro", _ip), ("g_cols", _ip), ("g_vals", _dp), ("gt_ro", _ip), ("gt_cols", _ip), ("gt_vals", _dp)]


class SolveResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("error", C.c_int), ("value", C.c_double),
                ("true_rel_res", C.c_double), ("cheb_alpha", C.c_double), ("cheb_beta", C.c_double),
                ("setup_s", C.c_double), ("solve_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_SIGS = [
    ("orc_dcol", None, [C.c_int, _ip, _ip, _ip]),
    ("orc_spmv_csr", None, [C.c_int, _ip, _ip, _dp, _dp, _dp]),
    ("orc_spmv_bsr", None, [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp]),
    ("orc_spmv_vblock", None, [C.c_int, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp]),
    ("orc_addlmult", None, [C.c_int, _ip, _ip, _ip, _dp, C.c_double, _dp, _dp]),
    ("orc_addumult", None, [C.c_int, _ip, _ip, _ip, _dp, C.c_double, _dp, _dp]),
    ("orc_invdiag_csr", None, [C.c_int, _ip, _dp, _dp]),
    ("orc_bellpack_nslices", C.c_int, [C.c_int, C.c_int]),
    ("orc_bellpack_bro", None, [C.c_int, C.c_int, _ip, _ip]),
    ("orc_bellpack_fill", None, [C.c_int, C.c_int, _ip, _ip, _dp, _ip, _ip, _dp]),
    ("orc_bellpack_masks", None, [C.c_int, C.c_int, _ip, _ip, _ip, _u8, _u8]),
    ("orc_bellpack_spmv", None, [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp]),
    ("orc_bellpack_invdiag", None, [C.c_int, C.c_int, _ip, _ip, _dp, _dp]),
    ("orc_dot", C.c_double, [C.c_int, _dp, _dp]),
    ("orc_ilu0_factor", None, [C.c_int, _ip, _ip, _ip, _dp]),
    ("orc_filu0_factor", None, [C.c_int, _ip, _ip, _ip, _dp, C.c_int, _dp]),
    ("orc_ilu0_solve", None, [C.c_int, _ip, _ip, _ip, _dp, _dp, _dp]),
    ("orc_levels_lower", C.c_int, [C.c_int, _ip, _ip, _ip, _ip]),
    ("orc_levels_upper", C.c_int, [C.c_int, _ip, _ip, _ip, _ip]),
    ("orc_color_greedy", C.c_int, [C.c_int, _ip, _ip, _ip]),
    ("orc_solve", C.c_int, [C.POINTER(SolveParams), C.c_int, _ip, _ip, _dp, _dp, _dp, C.POINTER(SolveResult)]),
    ("orc_cheb_bounds", None, [C.c_int, _ip, _ip, _dp, C.c_double, C.c_int, _dp, _dp]),
    ("orc_cheb_apply", None, [C.c_int, _ip, _ip, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp]),
    ("orc_time_cg_iters", C.c_double, [C.c_int, _ip, _ip, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("orc_time_spmv", C.c_double, [C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_int]),
]

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for n, r, a in _SIGS:
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def dcol(ro, cols):
    d = np.empty(len(ro) - 1, np.int32)
    lib().orc_dcol(len(ro) - 1, _i(ro), _i(cols), _i(d))
    return d


def spmv(ro, cols, vals, x):
    y = np.empty(len(ro) - 1)
    lib().orc_spmv_csr(len(ro) - 1, _i(ro), _i(cols), _d(vals), _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def spmv_bsr(ro, cols, vals, x, bs=2):
    y = np.empty((len(ro) - 1) * bs)
    lib().orc_spmv_bsr(len(ro) - 1, bs, _i(ro), _i(cols), _d(vals), _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def spmv_vblock(ro, cols, block_cols, block_size, block_offset, vals, x):
    """Variable-block product; y has block_offset[-1] + block_size[-1] entries."""
    nb = len(ro) - 1
    n = int(block_offset[nb - 1] + block_size[nb - 1]) if nb else 0
    y = np.empty(n)
    lib().orc_spmv_vblock(nb, _i(ro), _i(cols), _i(block_cols), _i(block_size), _i(block_offset), _d(vals),
                          _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def addmult(ro, cols, vals, alpha, x, y, upper):
    d = dcol(ro, cols)
    y = np.array(y, dtype=np.float64, copy=True)
    f = lib().orc_addumult if upper else lib().orc_addlmult
    f(len(ro) - 1, _i(ro), _i(cols), _i(d), _d(vals), alpha, _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def invdiag_csr(ro, cols, vals):
    d = dcol(ro, cols)
    y = np.empty(len(ro) - 1)
    lib().orc_invdiag_csr(len(ro) - 1, _i(d), _d(vals), _d(y))
    return y


def bellpack(ro, cols, vals=None, slice_=1024):
    n = len(ro) - 1
    ns = lib().orc_bellpack_nslices(n, slice_)
    bro = np.empty(ns + 1, np.int32)
    lib().orc_bellpack_bro(n, slice_, _i(ro), _i(bro))
    total = int(bro[ns]) * slice_
    bc = np.empty(total, np.int32)
    bv = np.empty(total, np.float64) if vals is not None else None
    lib().orc_bellpack_fill(n, slice_, _i(ro), _i(cols), None if vals is None else _d(vals), _i(bro), _i(bc),
                            None if bv is None else _d(bv))
    return bro, bc, bv


def bellpack_masks(ro, cols, bro, slice_=1024):
    n = len(ro) - 1
    d = dcol(ro, cols)
    total = int(bro[-1]) * slice_
    lm = np.empty(total, np.uint8)
    um = np.empty(total, np.uint8)
    lib().orc_bellpack_masks(n, slice_, _i(ro), _i(d), _i(bro), lm.ctypes.data_as(_u8), um.ctypes.data_as(_u8))
    return lm, um


def bellpack_spmv(n, bro, bc, bv, x, slice_=1024):
    y = np.empty(n)
    lib().orc_bellpack_spmv(n, slice_, _i(bro), _i(bc), _d(bv), _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def bellpack_invdiag(n, bro, bc, bv, preset=1.0, slice_=1024):
    y = np.full(n, preset)
    lib().orc_bellpack_invdiag(n, slice_, _i(bro), _i(bc), _d(bv), _d(y))
    return y


def dot(x, y):
    return lib().orc_dot(len(x), _d(np.ascontiguousarray(x, np.float64)), _d(np.ascontiguousarray(y, np.float64)))


def ilu0_factor(ro, cols, vals):
    d = dcol(ro, cols)
    lu = np.array(vals, dtype=np.float64, copy=True)
    lib().orc_ilu0_factor(len(ro) - 1, _i(ro), _i(cols), _i(d), _d(lu))
    return lu


def ilu0_solve(ro, cols, lu, y):
    d = dcol(ro, cols)
    x = np.empty(len(ro) - 1)
    lib().orc_ilu0_solve(len(ro) - 1, _i(ro), _i(cols), _i(d), _d(lu), _d(np.ascontiguousarray(y, np.float64)), _d(x))
    return x


def levels(ro, cols, upper=False):
    d = dcol(ro, cols)
    lv = np.empty(len(ro) - 1, np.int32)
    f = lib().orc_levels_upper if upper else lib().orc_levels_lower
    nl = f(len(ro) - 1, _i(ro), _i(cols), _i(d), _i(lv))
    return lv, nl


def color(ro, cols):
    c = np.empty(len(ro) - 1, np.int32)
    nc = lib().orc_color_greedy(len(ro) - 1, _i(ro), _i(cols), _i(c))
    return c, nc


def solve(ro, cols, vals, b, x0=None, *, solver=ORC_CG, precond=PREC_CHEB, tol=1e-8, max_iter=1000,
          poly_factor=0.5, poly_order=3, poly_max_iter=10, nthreads=1, filu_factor_niter=0, filu_solver_niter=3,
          factor=None):
    """factor: (G, G^T) as scipy CSR matrices with sorted indices, for precond=PREC_FACTOR."""
    n = len(ro) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    p = SolveParams(solver, precond, tol, max_iter, poly_factor, poly_order, poly_max_iter, nthreads,
                    filu_factor_niter, filu_solver_niter)
    keep = []
    if factor is not None:
        arrs = []
        for M in factor:
            a = (np.ascontiguousarray(M.indptr, np.int32), np.ascontiguousarray(M.indices, np.int32),
                 np.ascontiguousarray(M.data, np.float64))
            keep.append(a)
            arrs += [_i(a[0]), _i(a[1]), _d(a[2])]
        p.g_ro, p.g_cols, p.g_vals, p.gt_ro, p.gt_cols, p.gt_vals = arrs
    r = SolveResult()
    lib().orc_solve(C.byref(p), n, _i(ro), _i(cols), _d(vals), _d(np.ascontiguousarray(b, np.float64)), _d(x),
                    C.byref(r))
    return x, r.as_dict()


def cheb_bounds(ro, cols, vals, factor=0.5, max_it=10):
    """(alpha, beta) of ChebyshevPreconditioner::init (power method, then alpha = beta / factor)."""
    a, b = C.c_double(), C.c_double()
    lib().orc_cheb_bounds(len(ro) - 1, _i(ro), _i(cols), _d(vals), factor, max_it, C.byref(a), C.byref(b))
    return a.value, b.value


def cheb_apply(ro, cols, vals, alpha, beta, order, x):
    """y = M^-1 x by ChebyshevPreconditioner::_algo2 with the given bounds."""
    y = np.empty(len(ro) - 1)
    lib().orc_cheb_apply(len(ro) - 1, _i(ro), _i(cols), _d(vals), alpha, beta, order,
                         _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def filu0_factor(ro, cols, vals, niter):
    """FILU0's fixed-point factorisation after niter sweeps (FILU0Preconditioner.h:146-170)."""
    d = dcol(ro, cols)
    lu = np.empty(len(vals))
    lib().orc_filu0_factor(len(ro) - 1, _i(ro), _i(cols), _i(d), _d(vals), niter, _d(lu))
    return lu


def time_cg_iters(ro, cols, vals, b, precond=PREC_DIAG, poly_order=3, n_iter=5, nthreads=1):
    return lib().orc_time_cg_iters(len(ro) - 1, _i(ro), _i(cols), _d(vals), _d(np.ascontiguousarray(b, np.float64)),
                                   precond, poly_order, n_iter, nthreads)


def time_spmv(ro, cols, vals, x, reps=3, nthreads=1):
    y = np.empty(len(ro) - 1)
    t = lib().orc_time_spmv(len(ro) - 1, _i(ro), _i(cols), _d(vals), _d(np.ascontiguousarray(x, np.float64)), _d(y),
                            reps, nthreads)
    return t, y
