"""ctypes binding of libalien_hip.so (the C ABI declared in include/alien_hip.h).

The library is built in-tree by ``python -m alien_amd.build`` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: there is
no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libalien_hip.so"

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_u8_p = C.POINTER(C.c_uint8)
c_void_p = C.c_void_p

# enums (include/alien_hip.h)
GEN_POISSON7, GEN_CONVDIFF7, GEN_LOGNORMAL7 = 0, 1, 2
CG, BICGS = 0, 1
PREC_DIAG, PREC_CHEB, PREC_ILU0, PREC_NONE, PREC_LEARNED, PREC_FILU0, PREC_NEUMANN = 0, 1, 2, 3, 4, 5, 6
LAYOUT_CSR, LAYOUT_BELLPACK, LAYOUT_RBCSR, LAYOUT_RBDIA = 0, 1, 2, 3
MODE_REFERENCE, MODE_FUSED, MODE_GRAPH, MODE_MIXED, MODE_REFERENCE_FUTURE = 0, 1, 2, 3, 4
EINVAL = -22
NOALLOC_MAX_ROWS = 16777216  # ALIEN_HIP_NOALLOC_MAX_ROWS


class SolveParams(C.Structure):
    _fields_ = [("solver", C.c_int), ("precond", C.c_int), ("layout", C.c_int), ("mode", C.c_int),
                ("tol", C.c_double), ("max_iter", C.c_int), ("poly_factor", C.c_double),
                ("poly_order", C.c_int), ("poly_max_iter", C.c_int), ("ilu_schedule", C.c_int),
                ("inner_tol", C.c_double), ("inner_max_iter", C.c_int), ("max_outer", C.c_int),
                ("poll_every", C.c_int), ("gnn_theta", c_void_p), ("stall_window", C.c_int),
                ("filu_factor_niter", C.c_int), ("filu_solver_niter", C.c_int), ("factor_storage", C.c_int)]


class SolveResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("error", C.c_int), ("outer", C.c_int),
                ("value", C.c_double), ("true_rel_res", C.c_double), ("cheb_alpha", C.c_double),
                ("cheb_beta", C.c_double), ("setup_s", C.c_double), ("solve_s", C.c_double),
                ("fallback", C.c_int), ("pad_", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class GnnTrainParams(C.Structure):
    _fields_ = [("steps", C.c_int), ("steps_per_system", C.c_int), ("n_min", C.c_int), ("n_max", C.c_int),
                ("sigma_min", C.c_double), ("sigma_max", C.c_double), ("aniso_max", C.c_double),
                ("lr", C.c_double), ("beta", C.c_double), ("seed", C.c_uint64), ("cg_steps", C.c_int),
                ("pad_", C.c_int)]


# (name, restype, argtypes)
_SIGS = [
    ("alien_hip_last_error", C.c_char_p, []),
    ("alien_hip_version", C.c_char_p, []),
    ("alien_hip_ctx_create", C.c_int, [c_void_p, C.POINTER(c_void_p)]),
    ("alien_hip_ctx_destroy", C.c_int, [c_void_p]),
    ("alien_hip_ctx_synchronize", C.c_int, [c_void_p]),
    ("alien_hip_ctx_stream", c_void_p, [c_void_p]),
    ("alien_hip_ctx_step_timing", C.c_int, [c_void_p, C.c_int, c_double_p, C.POINTER(C.c_longlong)]),
    ("alien_hip_gen_nnz", C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int]),
    ("alien_hip_gen_csr", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, c_double_p, C.c_uint64, c_int_p, c_int_p,
                                    c_double_p]),
    ("alien_hip_gen_bsr2", C.c_int, [C.c_int, C.c_int, C.c_int, c_double_p, C.c_uint64, c_int_p, c_int_p,
                                     c_double_p]),
    ("alien_hip_fill_uniform01", C.c_int, [C.c_uint64, C.c_int64, c_double_p]),
    ("alien_hip_host_dcol", C.c_int, [C.c_int, c_int_p, c_int_p, c_int_p]),
    ("alien_hip_host_bellpack_bro", C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p]),
    ("alien_hip_host_bellpack_fill", C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p, c_double_p, c_int_p, c_int_p,
                                               c_double_p, c_u8_p, c_u8_p]),
    ("alien_hip_host_levels", C.c_int, [C.c_int, c_int_p, c_int_p, c_int_p, C.c_int, c_int_p, c_int_p]),
    ("alien_hip_host_color", C.c_int, [C.c_int, c_int_p, c_int_p, c_int_p, c_int_p]),
    ("alien_hip_rbcsr_bundle_offsets", C.c_int, [C.c_int, c_int_p, c_int_p]),
    ("alien_hip_spmv_csr_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_csr_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_bellpack_f64", C.c_int, [c_void_p, C.c_int, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p]),
    ("alien_hip_spmv_bellpack_f32", C.c_int, [c_void_p, C.c_int, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p]),
    ("alien_hip_bellpack_fill_f64", C.c_int, [c_void_p, C.c_int, C.c_int] + [c_void_p] * 9),
    ("alien_hip_bellpack_fill_f32", C.c_int, [c_void_p, C.c_int, C.c_int] + [c_void_p] * 6),
    ("alien_hip_addmult_bellpack_f64", C.c_int, [c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, c_void_p,
                                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_addmult_csr_f64", C.c_int, [c_void_p, C.c_int, C.c_double, C.c_int, c_void_p, c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_bsr2_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_bsr2_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_vblock_f64", C.c_int, [c_void_p, C.c_int] + [c_void_p] * 8),
    ("alien_hip_vblock_scalar_nnz", C.c_int64, [C.c_int, c_int_p, c_int_p, c_int_p]),
    ("alien_hip_vblock_to_csr", C.c_int, [C.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_double_p,
                                          c_int_p, c_int_p, c_double_p]),
    ("alien_hip_rbbsr2_fill_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p]),
    ("alien_hip_rbbsr2_fill_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p]),
    ("alien_hip_spmv_rbbsr2_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_rbbsr2_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_invdiag_csr_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_invdiag_bellpack_f64", C.c_int, [c_void_p, C.c_int, C.c_int, c_void_p, c_void_p, c_void_p,
                                                 c_void_p]),
    ("alien_hip_rbcsr_fill_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    ("alien_hip_rbcsr_fill_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    ("alien_hip_rbdia_count", C.c_int64, [C.c_int, c_int_p, c_int_p, C.c_int, c_int_p]),
    ("alien_hip_rbdia_struct", C.c_int, [C.c_int, c_int_p, c_int_p, C.c_int, c_int_p, C.POINTER(C.c_uint64)]),
    ("alien_hip_rbdia_count_device", C.c_int64, [c_void_p, C.c_int, c_void_p, c_void_p, C.c_int, C.c_int, c_int_p]),
    ("alien_hip_rbdia_struct_device", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, C.c_int, c_void_p, c_void_p,
                                                c_void_p]),
    ("alien_hip_rbdia2_fill_f64", C.c_int, [c_void_p, C.c_int] + [c_void_p] * 6),
    ("alien_hip_rbdia2_fill_f32", C.c_int, [c_void_p, C.c_int] + [c_void_p] * 6),
    ("alien_hip_spmv_rbdia2_f64", C.c_int, [c_void_p, C.c_int] + [c_void_p] * 6),
    ("alien_hip_spmv_rbdia2_f32", C.c_int, [c_void_p, C.c_int] + [c_void_p] * 6),
    ("alien_hip_rbdia_fill_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    ("alien_hip_rbdia_fill_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    ("alien_hip_spmv_rbdia_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    ("alien_hip_spmv_rbdia_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    ("alien_hip_filu0_sweep_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p]),
    ("alien_hip_spmv_rbcsr_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_rbcsr_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_copy_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p]),
    ("alien_hip_axpy_f64", C.c_int, [c_void_p, C.c_int, C.c_double, c_void_p, c_void_p]),
    ("alien_hip_scal_f64", C.c_int, [c_void_p, C.c_int, C.c_double, c_void_p]),
    ("alien_hip_assign_f64", C.c_int, [c_void_p, C.c_int, C.c_double, c_void_p]),
    ("alien_hip_pointwise_mult_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_dot_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_dot_f32", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_axpy_dot_f64", C.c_int, [c_void_p, C.c_int, C.c_double, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_dot_rbdia_f64", C.c_int, [c_void_p, C.c_int] + [c_void_p] * 7),
    ("alien_hip_spmv_dot_rbdia_f32", C.c_int, [c_void_p, C.c_int] + [c_void_p] * 7),
    ("alien_hip_host_diagonals", C.c_int, [C.c_int, c_int_p, c_int_p, C.c_int, c_int_p, c_int_p]),
    ("alien_hip_rbdia_uniform_struct", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, C.c_int, c_int_p, c_void_p,
                                                 c_void_p, c_void_p]),
    ("alien_hip_spmv_udia_f64", C.c_int, [c_void_p, C.c_int, C.c_int, c_int_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_udia_f32", C.c_int, [c_void_p, C.c_int, C.c_int, c_int_p, c_void_p, c_void_p, c_void_p]),
    ("alien_hip_spmv_dot_udia_f64", C.c_int, [c_void_p, C.c_int, C.c_int, c_int_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p]),
    ("alien_hip_spmv_dot_udia_f32", C.c_int, [c_void_p, C.c_int, C.c_int, c_int_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p]),
    ("alien_hip_spmv_dot_csr_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                             c_void_p]),
    ("alien_hip_spmv_dot_csr_f64_noalloc", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                                     c_void_p, c_void_p]),
    ("alien_hip_ilu0_factor_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, C.c_int,
                                            c_int_p, c_void_p]),
    ("alien_hip_ilu0_solve_f64", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, C.c_int,
                                           c_int_p, c_void_p, C.c_int, c_int_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    ("alien_hip_solve_csr", C.c_int, [c_void_p, C.POINTER(SolveParams), C.c_int, c_int_p, c_int_p, c_double_p,
                                      c_double_p, c_double_p, C.POINTER(SolveResult)]),
    ("alien_hip_mtx_read_csr", C.c_int, [C.c_char_p, c_int_p, C.POINTER(C.c_int64), c_int_p, c_int_p, c_double_p]),
    ("alien_hip_mtx_read_vector", C.c_int, [C.c_char_p, c_int_p, c_double_p]),
    ("alien_hip_mtx_write_csr", C.c_int, [C.c_char_p, C.c_int, c_int_p, c_int_p, c_double_p, C.c_char_p]),
    ("alien_hip_mtx_write_vector", C.c_int, [C.c_char_p, C.c_int, c_double_p, C.c_char_p]),
    ("alien_hip_gnn_nparams", C.c_int, []),
    ("alien_hip_gnn_init", C.c_int, [C.c_uint64, c_float_p]),
    ("alien_hip_gnn_forward", C.c_int, [c_void_p, C.c_int, c_int_p, c_int_p, c_double_p, c_float_p, c_float_p]),
    ("alien_hip_gnn_loss_grad", C.c_int, [c_void_p, C.c_int, c_int_p, c_int_p, c_double_p, c_float_p, c_float_p,
                                          C.c_double, C.c_int, c_double_p, c_double_p, c_float_p]),
    ("alien_hip_gnn_train", C.c_int, [c_void_p, C.POINTER(GnnTrainParams), c_float_p, c_double_p]),
    ("alien_hip_matrix_create", C.c_int, [c_void_p, C.c_int, c_int_p, c_int_p, c_double_p, C.POINTER(c_void_p)]),
    ("alien_hip_matrix_update_values", C.c_int, [c_void_p, c_double_p]),
    ("alien_hip_matrix_destroy", C.c_int, [c_void_p]),
    ("alien_hip_solve", C.c_int, [c_void_p, C.POINTER(SolveParams), c_void_p, c_double_p, c_double_p,
                                  C.POINTER(SolveResult)]),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


class AlienHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libalien_hip.so (built in-tree). Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m alien_amd.build` first")
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        raise AlienHipError(rc, lib().alien_hip_last_error().decode())


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_int_p)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def u8ptr(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(c_u8_p)


# ---- host generators and builders (no GPU) -----------------------------------------
def gen_nnz(kind: int, nx: int, ny: int | None = None, nz: int | None = None) -> int:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return int(lib().alien_hip_gen_nnz(kind, nx, ny, nz))


def gen_csr(kind: int, nx: int, ny: int | None = None, nz: int | None = None, params=None, seed: int = 42):
    """Generate one of the synthetic 7-point systems as host CSR (int32/float64)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    nnz = gen_nnz(kind, nx, ny, nz)
    ro = np.empty(n + 1, np.int32)
    cols = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    par = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    check(lib().alien_hip_gen_csr(kind, nx, ny, nz, None if par is None else dptr(par), seed, iptr(ro), iptr(cols),
                                  dptr(vals)))
    return ro, cols, vals


def gen_bsr2(nx: int, ny: int | None = None, nz: int | None = None, params=None, seed: int = 42):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    nnzb = gen_nnz(GEN_POISSON7, nx, ny, nz)
    ro = np.empty(n + 1, np.int32)
    cols = np.empty(nnzb, np.int32)
    vals = np.empty(4 * nnzb, np.float64)
    par = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    check(lib().alien_hip_gen_bsr2(nx, ny, nz, None if par is None else dptr(par), seed, iptr(ro), iptr(cols),
                                   dptr(vals)))
    return ro, cols, vals


def uniform01(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    check(lib().alien_hip_fill_uniform01(seed, n, dptr(out)))
    return out


def host_dcol(ro, cols):
    n = len(ro) - 1
    d = np.empty(n, np.int32)
    check(lib().alien_hip_host_dcol(n, iptr(ro), iptr(cols), iptr(d)))
    return d


def vblock_to_csr(ro, cols, block_cols, block_size, block_offset, vals=None):
    """Scalar CSR (row_offset, cols, vals or None) of a variable-block CSR matrix, entries in the
    order of the variable-block product."""
    ro, cols, bcols, bsize, boff = (np.ascontiguousarray(a, dtype=np.int32)
                                    for a in (ro, cols, block_cols, block_size, block_offset))
    nb = len(ro) - 1
    nnz = lib().alien_hip_vblock_scalar_nnz(nb, iptr(ro), iptr(cols), iptr(bsize))
    check(0 if nnz >= 0 else -1)
    n = int(boff[-1] + bsize[-1]) if nb else 0
    sro = np.zeros(n + 1, np.int32)
    scols = np.empty(max(nnz, 1), np.int32)
    svals = None if vals is None else np.empty(max(nnz, 1))
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.float64)
    check(lib().alien_hip_vblock_to_csr(nb, iptr(ro), iptr(cols), iptr(bcols), iptr(bsize), iptr(boff),
                                        None if v is None else dptr(v), iptr(sro), iptr(scols),
                                        None if svals is None else dptr(svals)))
    return sro, scols[:nnz], (None if svals is None else svals[:nnz])


def bellpack_bro(ro, slice_=1024):
    """Slice offsets of the BEllPack layout (nslices + 1 entries)."""
    n = len(ro) - 1
    bro = np.empty((n + slice_ - 1) // slice_ + 1, np.int32)
    check(lib().alien_hip_host_bellpack_bro(n, slice_, iptr(np.ascontiguousarray(ro, np.int32)), iptr(bro)))
    return bro


def bellpack_build(ro, cols, vals=None, slice_=1024, masks=False):
    n = len(ro) - 1
    ns = (n + slice_ - 1) // slice_
    bro = np.empty(ns + 1, np.int32)
    check(lib().alien_hip_host_bellpack_bro(n, slice_, iptr(ro), iptr(bro)))
    total = int(bro[ns]) * slice_
    bc = np.empty(total, np.int32)
    bv = np.empty(total, np.float64) if vals is not None else None
    lm = np.empty(total, np.uint8) if masks else None
    um = np.empty(total, np.uint8) if masks else None
    check(lib().alien_hip_host_bellpack_fill(n, slice_, iptr(ro), iptr(cols),
                                             None if vals is None else dptr(vals), iptr(bro), iptr(bc),
                                             None if bv is None else dptr(bv), None if lm is None else u8ptr(lm),
                                             None if um is None else u8ptr(um)))
    return bro, bc, bv, lm, um


def host_levels(ro, cols, dcol, upper: bool):
    n = len(ro) - 1
    lv = np.empty(n, np.int32)
    nl = C.c_int(0)
    check(lib().alien_hip_host_levels(n, iptr(ro), iptr(cols), iptr(dcol), 1 if upper else 0, iptr(lv),
                                      C.byref(nl)))
    return lv, nl.value


def host_color(ro, cols):
    n = len(ro) - 1
    c = np.empty(n, np.int32)
    nc = C.c_int(0)
    check(lib().alien_hip_host_color(n, iptr(ro), iptr(cols), iptr(c), C.byref(nc)))
    return c, nc.value


def rbcsr_bundle_offsets(ro):
    n = len(ro) - 1
    bo = np.empty((n + 63) // 64 + 1, np.int32)
    check(lib().alien_hip_rbcsr_bundle_offsets(n, iptr(ro), iptr(bo)))
    return bo


def rbdia_build(ro, cols, max_width=64):
    """Host RB-DIA structure: (bundle_off, doff, dmask) or None when a bundle holds
    more than max_width distinct column offsets."""
    ro = np.ascontiguousarray(ro, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    n = len(ro) - 1
    bo = np.zeros((n + 63) // 64 + 1, dtype=np.int32)
    slots = lib().alien_hip_rbdia_count(n, iptr(ro), iptr(cols), max_width, iptr(bo))
    if slots < 0:
        return None
    doff = np.zeros(max(slots, 1), dtype=np.int32)
    dmask = np.zeros(max(slots, 1), dtype=np.uint64)
    check(lib().alien_hip_rbdia_struct(n, iptr(ro), iptr(cols), max_width, iptr(doff),
                                       dmask.ctypes.data_as(C.POINTER(C.c_uint64))))
    return bo, doff[:slots], dmask[:slots]


def host_diagonals(ro, cols, max_k=8):
    """The distinct diagonal offsets col - row of a CSR, ascending, or None if more than max_k."""
    ro = np.ascontiguousarray(ro, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    off = np.zeros(max_k, dtype=np.int32)
    k = C.c_int()
    check(lib().alien_hip_host_diagonals(len(ro) - 1, iptr(ro), iptr(cols), max_k, iptr(off), C.byref(k)))
    return None if k.value < 0 else off[:k.value].copy()


def group_by_level(level: np.ndarray, nlevels: int):
    """Rows grouped by level, ascending inside a level; returns (ptr, rows)."""
    order = np.argsort(level, kind="stable").astype(np.int32)
    counts = np.bincount(level, minlength=nlevels)
    ptr = np.zeros(nlevels + 1, np.int32)
    ptr[1:] = np.cumsum(counts)
    return ptr, order


# ---- device context ----------------------------------------------------------------------
class Context:
    """A HIP stream + reduction workspace of the backend (alien_hip_ctx)."""

    def __init__(self, stream: int | None = None):
        h = c_void_p()
        check(lib().alien_hip_ctx_create(c_void_p(stream) if stream else None, C.byref(h)))
        self.h = h

    def synchronize(self):
        check(lib().alien_hip_ctx_synchronize(self.h))

    @property
    def stream(self) -> int:
        return lib().alien_hip_ctx_stream(self.h)

    def step_timing(self, on: bool):
        """Switch the fused CG solvers' step-kernel timing (q = A p, HIP events on this stream)
        on or off; returns the (microseconds, launches) totals gathered before the switch."""
        us, k = C.c_double(), C.c_longlong()
        check(lib().alien_hip_ctx_step_timing(self.h, 1 if on else 0, C.byref(us), C.byref(k)))
        return us.value, k.value

    def close(self):
        if self.h:
            check(lib().alien_hip_ctx_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _solve_params(solver=CG, precond=PREC_CHEB, layout=LAYOUT_CSR, mode=MODE_REFERENCE, tol=1e-8, max_iter=1000,
                  poly_factor=0.5, poly_order=3, poly_max_iter=10, ilu_schedule=0, inner_tol=0.0, inner_max_iter=0,
                  max_outer=50, poll_every=8, gnn_theta=None, stall_window=0, filu_factor_niter=0,
                  filu_solver_niter=3, factor_storage=0):
    """(SolveParams, the float32 theta array it points to)."""
    theta = None if gnn_theta is None else np.ascontiguousarray(gnn_theta, dtype=np.float32)
    p = SolveParams(solver, precond, layout, mode, tol, max_iter, poly_factor, poly_order, poly_max_iter,
                    ilu_schedule, inner_tol, inner_max_iter, max_outer, poll_every,
                    None if theta is None else theta.ctypes.data, stall_window, filu_factor_niter, filu_solver_niter,
                    factor_storage)
    return p, theta


def solve_csr(ctx: Context, ro, cols, vals, b, x0=None, **kw):
    """Run one Krylov solve through the hip backend; returns (x, result dict). Keywords: the fields of
    alien_hip_solve_params (solver, precond, layout, mode, tol, max_iter, ...)."""
    n = len(ro) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    p, _theta = _solve_params(**kw)
    r = SolveResult()
    check(lib().alien_hip_solve_csr(ctx.h, C.byref(p), n, iptr(ro), iptr(cols), dptr(vals), dptr(b), dptr(x),
                                    C.byref(r)))
    return x, r.as_dict()


class Matrix:
    """A converted matrix kept between solves (alien_hip_matrix_create / _update_values / alien_hip_solve)."""

    def __init__(self, ctx: Context, ro, cols, vals):
        self.ctx = ctx
        self.n = len(ro) - 1
        self.nnz = len(cols)
        h = c_void_p()
        check(lib().alien_hip_matrix_create(ctx.h, self.n, iptr(np.ascontiguousarray(ro, np.int32)),
                                            iptr(np.ascontiguousarray(cols, np.int32)),
                                            dptr(np.ascontiguousarray(vals, np.float64)), C.byref(h)))
        self.h = h

    def update_values(self, vals):
        v = np.ascontiguousarray(vals, np.float64)
        assert len(v) == self.nnz
        check(lib().alien_hip_matrix_update_values(self.h, dptr(v)))

    def solve(self, b, x0=None, ctx_override: Context | None = None, **kw):
        """ctx_override: solve through another context (stream) than the one the matrix was created on."""
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        b = np.ascontiguousarray(b, dtype=np.float64)
        p, _theta = _solve_params(**kw)
        r = SolveResult()
        c = ctx_override if ctx_override is not None else self.ctx
        check(lib().alien_hip_solve(c.h, C.byref(p), self.h, dptr(b), dptr(x), C.byref(r)))
        return x, r.as_dict()

    def close(self):
        if self.h:
            lib().alien_hip_matrix_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# --------------------------------------------------------------------------
# learned preconditioner M = G G^T
# --------------------------------------------------------------------------
# --------------------------------------------------------------------------
# MatrixMarket system I/O
# --------------------------------------------------------------------------
def mtx_read_csr(path):
    """(row_offset, cols, vals) of a MatrixMarket 'coordinate real' matrix."""
    n, nnz = C.c_int(), C.c_int64()
    p = str(path).encode()
    check(lib().alien_hip_mtx_read_csr(p, C.byref(n), C.byref(nnz), None, None, None))
    ro = np.empty(n.value + 1, dtype=np.int32)
    cols = np.empty(max(nnz.value, 1), dtype=np.int32)
    vals = np.empty(max(nnz.value, 1))
    check(lib().alien_hip_mtx_read_csr(p, C.byref(n), C.byref(nnz), iptr(ro), iptr(cols), dptr(vals)))
    return ro, cols[:nnz.value], vals[:nnz.value]


def mtx_read_vector(path):
    n = C.c_int()
    p = str(path).encode()
    check(lib().alien_hip_mtx_read_vector(p, C.byref(n), None))
    v = np.empty(max(n.value, 1))
    check(lib().alien_hip_mtx_read_vector(p, C.byref(n), dptr(v)))
    return v[:n.value]


def mtx_write_csr(path, ro, cols, vals, description=""):
    ro = np.ascontiguousarray(ro, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    check(lib().alien_hip_mtx_write_csr(str(path).encode(), len(ro) - 1, iptr(ro), iptr(cols), dptr(vals),
                                        description.encode()))


def mtx_write_vector(path, v, description=""):
    v = np.ascontiguousarray(v, dtype=np.float64)
    check(lib().alien_hip_mtx_write_vector(str(path).encode(), len(v), dptr(v), description.encode()))


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_float_p)


def gnn_nparams() -> int:
    return lib().alien_hip_gnn_nparams()


def gnn_init(seed: int = 0) -> np.ndarray:
    """Initial parameters: random layers, heads at the first-order inverse factor."""
    th = np.zeros(gnn_nparams(), dtype=np.float32)
    check(lib().alien_hip_gnn_init(seed, fptr(th)))
    return th


def _csr_args(ro, cols, vals):
    ro = np.ascontiguousarray(ro, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    return ro, cols, vals


def gnn_forward(ctx: Context, ro, cols, vals, theta) -> np.ndarray:
    """Ghat per CSR entry of A (FP32; strictly-upper entries 0)."""
    ro, cols, vals = _csr_args(ro, cols, vals)
    theta = np.ascontiguousarray(theta, dtype=np.float32)
    out = np.zeros(len(cols), dtype=np.float32)
    check(lib().alien_hip_gnn_forward(ctx.h, len(ro) - 1, iptr(ro), iptr(cols), dptr(vals), fptr(theta),
                                      fptr(out)))
    return out


def gnn_loss_grad(ctx: Context, ro, cols, vals, theta, probes, beta=0.1, cg_steps=0):
    """(loss, (L1, probe), grad) of the training objective on 4 probes (N x 4); L1 = Kaporin
    (cg_steps = 0) or the residual reduction after cg_steps PCG steps."""
    ro, cols, vals = _csr_args(ro, cols, vals)
    theta = np.ascontiguousarray(theta, dtype=np.float32)
    probes = np.ascontiguousarray(probes, dtype=np.float32)
    assert probes.shape == (len(ro) - 1, 4)
    loss = C.c_double()
    parts = np.zeros(2)
    grad = np.zeros(gnn_nparams(), dtype=np.float32)
    check(lib().alien_hip_gnn_loss_grad(ctx.h, len(ro) - 1, iptr(ro), iptr(cols), dptr(vals), fptr(theta),
                                        fptr(probes), beta, cg_steps, C.byref(loss), dptr(parts), fptr(grad)))
    return loss.value, (parts[0], parts[1]), grad


def gnn_train(ctx: Context, theta, *, steps=200, steps_per_system=4, n_min=8, n_max=16, sigma_min=0.5,
              sigma_max=2.0, aniso_max=4.0, lr=2e-3, beta=0.1, seed=1, cg_steps=0):
    """Adam on generated log-normal diffusion systems; returns (theta, loss history)."""
    th = np.array(theta, dtype=np.float32, copy=True)
    hist = np.zeros(steps)
    p = GnnTrainParams(steps, steps_per_system, n_min, n_max, sigma_min, sigma_max, aniso_max, lr, beta, seed,
                       cg_steps, 0)
    check(lib().alien_hip_gnn_train(ctx.h, C.byref(p), fptr(th), dptr(hist)))
    return th, hist
