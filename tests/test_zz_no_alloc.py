"""Allocation-free fused SpMV + dot (alien_hip_spmv_dot_csr_f64_noalloc) against
the context entry point alien_hip_spmv_dot_csr_f64: two calls in a row on one
GPU, the first at the largest size the scratch array admits, the second on a
smaller system (the array still holds the first call's partials), each equal
to the context entry point bit for bit and to the oracle's CSR product."""
import ctypes as C

import numpy as np
import pytest

import alien_amd as aa
import oracle as orc
from alien_amd import _capi
from gpu_util import call, dev, empty, host, p


def say(*a):
    print("case", *a, flush=True)


def _both(ctx, ro, cols, vals, x):
    n = len(ro) - 1
    dro, dc, dv, dx = dev(ro), dev(cols), dev(vals), dev(x)
    y_ctx, r_ctx = empty(n), empty(1)
    call("alien_hip_spmv_dot_csr_f64", ctx.h, n, p(dro), p(dc), p(dv), p(dx), p(y_ctx), p(r_ctx))
    y_na, r_na = empty(n), empty(1)
    stream = aa.lib().alien_hip_ctx_stream(ctx.h)
    call("alien_hip_spmv_dot_csr_f64_noalloc", C.c_void_p(stream), n, p(dro), p(dc), p(dv), p(dx), p(y_na), p(r_na))
    return host(y_ctx, ctx)[:n], host(r_ctx, ctx)[0], host(y_na, ctx)[:n], host(r_na, ctx)[0]


@pytest.mark.gpu
def test_noalloc_spmv_dot_equals_context_entry_point(ctx):
    # call 1: a diagonal system of the maximum size (one partial per 256 rows fills the array)
    n = _capi.NOALLOC_MAX_ROWS
    say("noalloc spmv+dot diagonal", n)
    ro = np.arange(n + 1, dtype=np.int32)
    cols = np.arange(n, dtype=np.int32)
    rng = np.random.default_rng(11)
    vals = rng.uniform(1.0, 2.0, n)
    x = rng.standard_normal(n)
    y_ctx, r_ctx, y_na, r_na = _both(ctx, ro, cols, vals, x)
    np.testing.assert_array_equal(y_na, vals * x)  # one product per row: exact
    np.testing.assert_array_equal(y_na, y_ctx)
    assert r_na == r_ctx
    ref = float(np.dot(vals * x, x))
    assert abs(r_na - ref) <= 1e-13 * float(np.abs(vals * x * x).sum())
    del ro, cols, vals, x, y_ctx, y_na
    # call 2: a smaller 7-point system
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 24, params=[2.0, 1.0, 1.0, 1.0], seed=5)
    n = len(ro) - 1
    say("noalloc spmv+dot lognormal", n)
    x = np.random.default_rng(12).standard_normal(n)
    y_ctx, r_ctx, y_na, r_na = _both(ctx, ro, cols, vals, x)
    ref = orc.spmv(ro, cols, vals, x)
    np.testing.assert_array_equal(y_na, ref)
    np.testing.assert_array_equal(y_na, y_ctx)
    assert r_na == r_ctx
    assert abs(r_na - float(np.dot(ref, x))) <= 1e-13 * float(np.abs(ref * x).sum())


@pytest.mark.gpu
def test_noalloc_spmv_dot_empty_system(ctx):
    say("noalloc spmv+dot empty")
    res = empty(1)
    stream = aa.lib().alien_hip_ctx_stream(ctx.h)
    ro = dev(np.zeros(1, np.int32))
    call("alien_hip_spmv_dot_csr_f64_noalloc", C.c_void_p(stream), 0, p(ro), None, None, None, None, p(res))
    assert host(res, ctx)[0] == 0.0


def test_noalloc_rejects_sizes_beyond_the_scratch_array():
    """Host-side checks: no GPU call is made for a rejected size."""
    L = aa.lib()
    rc = L.alien_hip_spmv_dot_csr_f64_noalloc(None, _capi.NOALLOC_MAX_ROWS + 1, C.c_void_p(8), C.c_void_p(8),
                                              C.c_void_p(8), C.c_void_p(8), C.c_void_p(8), C.c_void_p(8))
    assert rc == _capi.EINVAL
    assert b"ALIEN_HIP_NOALLOC_MAX_ROWS" in L.alien_hip_last_error()
    rc = L.alien_hip_spmv_dot_csr_f64_noalloc(None, 10, None, None, None, None, None, None)
    assert rc == _capi.EINVAL
