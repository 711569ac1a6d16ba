"""GPU parity of the gfx950 kernels, called through the C ABI, against the CPU
oracle on the same seeded inputs. Integer/index work and every kernel that
keeps the reference's summation order must be bit-exact; dot products are
compared with a tolerance stated in each test."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import alien_amd as aa
import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from gpu_util import call, dev, empty, host, p, row_scale  # noqa: E402


def systems():
    yield "poisson16", aa.gen_csr(aa.GEN_POISSON7, 16)
    yield "convdiff 13x7x9", aa.gen_csr(aa.GEN_CONVDIFF7, 13, 7, 9)
    yield "lognormal 20", aa.gen_csr(aa.GEN_LOGNORMAL7, 20, params=[2.0, 1, 1, 1], seed=3)
    rng = np.random.default_rng(5)
    A = (sp.random(3001, 3001, density=0.003, random_state=rng, format="csr") + sp.identity(3001)).tocsr()
    A = A.tolil()
    for i in rng.choice(3001, 50, replace=False):  # some empty rows, some very long ones
        A.rows[i] = []
        A.data[i] = []
    for i in rng.choice(3001, 5, replace=False):
        A[i, rng.choice(3001, 200, replace=False)] = 1.5
    A = A.tocsr()
    A.sort_indices()
    yield "ragged 3001", (A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


SYSTEMS = list(systems())


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name,A", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_spmv_csr_and_rbcsr_bit_exact(ctx, name, A):
    ro, cols, vals = A
    n = len(ro) - 1
    x = np.random.default_rng(1).standard_normal(n)
    ref = orc.spmv(ro, cols, vals, x)
    say("case spmv csr/rbcsr", name, n)
    dro, dc, dv, dx = dev(ro), dev(cols), dev(vals), dev(x)
    y = empty(n)
    call("alien_hip_spmv_csr_f64", ctx.h, n, p(dro), p(dc), p(dv), p(dx), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:n], ref)
    bo = aa.rbcsr_bundle_offsets(ro)
    slots = int(bo[-1]) * 64
    dbo = dev(bo)
    bc = empty(slots, torch.int32)
    bv = empty(slots)
    call("alien_hip_rbcsr_fill_f64", ctx.h, n, p(dro), p(dc), p(dv), p(dbo), p(bc), p(bv))
    y2 = empty(n)
    call("alien_hip_spmv_rbcsr_f64", ctx.h, n, p(dbo), p(bc), p(bv), p(dx), p(y2))
    np.testing.assert_array_equal(host(y2, ctx)[:n], ref)
    # FP32 matrix and vectors: within FP32 rounding of the per-row scale
    bv32 = empty(slots, torch.float32)
    call("alien_hip_rbcsr_fill_f32", ctx.h, n, p(dro), p(dc), p(dv), p(dbo), p(bc), p(bv32))
    dx32 = dev(x.astype(np.float32))
    y3 = empty(n, torch.float32)
    call("alien_hip_spmv_rbcsr_f32", ctx.h, n, p(dbo), p(bc), p(bv32), p(dx32), p(y3))
    scale = row_scale(ro, cols, vals, x)
    err = np.abs(host(y3, ctx)[:n].astype(np.float64) - ref)
    assert np.all(err <= 1e-5 * scale + 1e-30)
    y4 = empty(n, torch.float32)
    dv32 = dev(vals.astype(np.float32))
    call("alien_hip_spmv_csr_f32", ctx.h, n, p(dro), p(dc), p(dv32), p(dx32), p(y4))
    np.testing.assert_array_equal(host(y4, ctx)[:n], host(y3, ctx)[:n])


@pytest.mark.parametrize("name,A", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_spmv_rbdia_bit_exact(ctx, name, A):
    """RB-DIA (no column stream): FP64 bit-exact with the sequential CSR product,
    FP32 bit-exact with the FP32 RB-CSR kernel (same per-row order)."""
    ro, cols, vals = A
    n = len(ro) - 1
    built = aa.rbdia_build(ro, cols, max_width=1 << 20)
    assert built is not None
    bo, doff, dmask = built
    slots = int(bo[-1]) * 64
    say("case spmv rbdia", name, n, "slots/row", slots / max(n, 1))
    x = np.random.default_rng(7).standard_normal(n)
    ref = orc.spmv(ro, cols, vals, x)
    dro, dc, dv, dx = dev(ro), dev(cols), dev(vals), dev(x)
    dbo, ddo, ddm = dev(bo), dev(doff), dev(dmask.view(np.int64))
    bv = empty(slots)
    call("alien_hip_rbdia_fill_f64", ctx.h, n, p(dro), p(dc), p(dv), p(dbo), p(ddo), p(bv))
    y = empty(n)
    call("alien_hip_spmv_rbdia_f64", ctx.h, n, p(dbo), p(ddo), p(ddm), p(bv), p(dx), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:n], ref)
    bv32 = empty(slots, torch.float32)
    call("alien_hip_rbdia_fill_f32", ctx.h, n, p(dro), p(dc), p(dv), p(dbo), p(ddo), p(bv32))
    dx32 = dev(x.astype(np.float32))
    y32 = empty(n, torch.float32)
    call("alien_hip_spmv_rbdia_f32", ctx.h, n, p(dbo), p(ddo), p(ddm), p(bv32), p(dx32), p(y32))
    cbo = aa.rbcsr_bundle_offsets(ro)
    cslots = int(cbo[-1]) * 64
    dcbo = dev(cbo)
    cbc = empty(cslots, torch.int32)
    cbv = empty(cslots, torch.float32)
    call("alien_hip_rbcsr_fill_f32", ctx.h, n, p(dro), p(dc), p(dv), p(dcbo), p(cbc), p(cbv))
    yc = empty(n, torch.float32)
    call("alien_hip_spmv_rbcsr_f32", ctx.h, n, p(dcbo), p(cbc), p(cbv), p(dx32), p(yc))
    np.testing.assert_array_equal(host(y32, ctx)[:n], host(yc, ctx)[:n])
    # the fused CG iteration's q = A p, p.q kernel: the same sums, and x . y
    res = empty(1)
    yq = empty(n)
    call("alien_hip_spmv_dot_rbdia_f64", ctx.h, n, p(dbo), p(ddo), p(ddm), p(bv), p(dx), p(yq), p(res))
    np.testing.assert_array_equal(host(yq, ctx)[:n], ref)
    assert abs(host(res, ctx)[0] - np.dot(ref, x)) <= 1e-13 * np.abs(ref * x).sum() + 1e-300
    yq32 = empty(n, torch.float32)
    call("alien_hip_spmv_dot_rbdia_f32", ctx.h, n, p(dbo), p(ddo), p(ddm), p(bv32), p(dx32), p(yq32), p(res))
    y32h = host(y32, ctx)[:n]
    np.testing.assert_array_equal(host(yq32, ctx)[:n], y32h)
    e32 = np.dot(y32h.astype(np.float64), x.astype(np.float32).astype(np.float64))
    assert abs(host(res, ctx)[0] - e32) <= 1e-13 * np.abs(y32h.astype(np.float64) * x).sum() + 1e-300


UDIA_SYSTEMS = SYSTEMS[:3] + [("poisson2d 37x29", None), ("one row", None), ("65 rows", None)]


def udia_system(name, A):
    if name.startswith("poisson2d"):
        B = sp.diags([-1, -1, 4, -1, -1], [-37, -1, 0, 1, 37], shape=(37 * 29, 37 * 29), format="csr")
        B.sort_indices()
        return B.indptr.astype(np.int32), B.indices.astype(np.int32), B.data.astype(np.float64)
    if name == "one row":
        return np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.5])
    if name == "65 rows":
        B = sp.diags([-1.5, 3.0, -0.5], [-1, 0, 1], shape=(65, 65), format="csr")
        B.sort_indices()
        return B.indptr.astype(np.int32), B.indices.astype(np.int32), B.data.astype(np.float64)
    return A


@pytest.mark.parametrize("name,A", UDIA_SYSTEMS, ids=[s[0] for s in UDIA_SYSTEMS])
def test_spmv_uniform_rbdia(ctx, name, A):
    """Uniform RB-DIA (every bundle padded to the matrix's k diagonals, offsets as kernel
    arguments): its structure is an ordinary RB-DIA of the matrix (the RB-DIA kernel on it is
    bit-exact with CSR); the uniform kernels give the CSR sum's values (FP64 against the oracle,
    FP32 against the FP32 RB-CSR kernel; absent entries add +0, so compared with ==), and the fused
    q = A p, p.q kernel its values and the dot within the FP64 reduction bound."""
    ro, cols, vals = udia_system(name, A)
    n = len(ro) - 1
    off = aa.host_diagonals(ro, cols)
    assert off is not None and np.all(np.diff(off) > 0)
    rows = np.repeat(np.arange(n), np.diff(ro))
    np.testing.assert_array_equal(off, np.unique(cols.astype(np.int64) - rows))
    k = len(off)
    nb = (n + 63) // 64
    say("case udia", name, n, "offsets", list(off))
    x = np.random.default_rng(11).standard_normal(n)
    ref = orc.spmv(ro, cols, vals, x)
    dro, dc, dv, dx = dev(ro), dev(cols), dev(vals), dev(x)
    dbo, ddo, ddm = empty(nb + 1, torch.int32), empty(nb * k, torch.int32), empty(nb * k, torch.int64)
    ioff = aa._capi.iptr(off)
    call("alien_hip_rbdia_uniform_struct", ctx.h, n, p(dro), p(dc), k, ioff, p(dbo), p(ddo), p(ddm))
    bo_h = host(dbo, ctx)
    np.testing.assert_array_equal(bo_h, np.arange(nb + 1) * k)
    np.testing.assert_array_equal(host(ddo, ctx), np.tile(off, nb))
    # the lane masks: bit l of slot j of bundle b <=> row 64 b + l has offset off[j]
    mh = host(ddm, ctx).view(np.uint64).reshape(nb, k)
    want = np.zeros((nb, k), np.uint64)
    jj = np.searchsorted(off, cols.astype(np.int64) - rows)
    np.bitwise_or.at(want, (rows // 64, jj), (np.uint64(1) << (rows % 64).astype(np.uint64)))
    np.testing.assert_array_equal(mh, want)
    bv = empty(nb * k * 64)
    call("alien_hip_rbdia_fill_f64", ctx.h, n, p(dro), p(dc), p(dv), p(dbo), p(ddo), p(bv))
    y = empty(n)
    call("alien_hip_spmv_rbdia_f64", ctx.h, n, p(dbo), p(ddo), p(ddm), p(bv), p(dx), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:n], ref)
    yu = empty(n)
    call("alien_hip_spmv_udia_f64", ctx.h, n, k, ioff, p(bv), p(dx), p(yu))
    np.testing.assert_array_equal(host(yu, ctx)[:n], ref)
    res = empty(1)
    yq = empty(n)
    call("alien_hip_spmv_dot_udia_f64", ctx.h, n, k, ioff, p(bv), p(dx), p(yq), p(res))
    np.testing.assert_array_equal(host(yq, ctx)[:n], ref)
    assert abs(host(res, ctx)[0] - np.dot(ref, x)) <= 1e-13 * np.abs(ref * x).sum() + 1e-300
    bv32 = empty(nb * k * 64, torch.float32)
    call("alien_hip_rbdia_fill_f32", ctx.h, n, p(dro), p(dc), p(dv), p(dbo), p(ddo), p(bv32))
    dx32 = dev(x.astype(np.float32))
    y32 = empty(n, torch.float32)
    call("alien_hip_spmv_udia_f32", ctx.h, n, k, ioff, p(bv32), p(dx32), p(y32))
    cbo = aa.rbcsr_bundle_offsets(ro)
    cslots = int(cbo[-1]) * 64
    dcbo = dev(cbo)
    cbc = empty(cslots, torch.int32)
    cbv = empty(cslots, torch.float32)
    call("alien_hip_rbcsr_fill_f32", ctx.h, n, p(dro), p(dc), p(dv), p(dcbo), p(cbc), p(cbv))
    yc = empty(n, torch.float32)
    call("alien_hip_spmv_rbcsr_f32", ctx.h, n, p(dcbo), p(cbc), p(cbv), p(dx32), p(yc))
    np.testing.assert_array_equal(host(y32, ctx)[:n], host(yc, ctx)[:n])
    yq32 = empty(n, torch.float32)
    call("alien_hip_spmv_dot_udia_f32", ctx.h, n, k, ioff, p(bv32), p(dx32), p(yq32), p(res))
    np.testing.assert_array_equal(host(yq32, ctx)[:n], host(yc, ctx)[:n])


def test_uniform_rbdia_rejects_bad_offsets(ctx):
    ro, cols, vals = aa.gen_csr(aa.GEN_POISSON7, 4)
    n = len(ro) - 1
    x, y, v = empty(n), empty(n), empty(64 * 7)
    for bad in (np.array([1, 0], np.int32), np.arange(9, dtype=np.int32)):
        ptr = [C.c_void_p(t.data_ptr()) for t in (v, x, y)]
        rc = aa.lib().alien_hip_spmv_udia_f64(ctx.h, n, len(bad), aa._capi.iptr(bad), *ptr)
        assert rc == aa._capi.EINVAL


def rbdia_device(ctx, dro, dc, n, max_width, pair=0):
    """RB-DIA structure built on the device: (bo, doff, dmask) on the host, or None."""
    nb = (n + 63) // 64
    bo = np.zeros(nb + 1, dtype=np.int32)
    slots = aa.lib().alien_hip_rbdia_count_device(ctx.h, n, C.c_void_p(dro.data_ptr()), C.c_void_p(dc.data_ptr()),
                                                  max_width, pair, bo.ctypes.data_as(C.POINTER(C.c_int)))
    if slots < 0:
        return None
    ddo = empty(max(slots, 1), torch.int32)
    ddm = empty(max(slots, 1), torch.int64)
    call("alien_hip_rbdia_struct_device", ctx.h, n, p(dro), p(dc), pair, p(dev(bo)), p(ddo), p(ddm))
    return bo, host(ddo, ctx)[:slots], host(ddm, ctx)[:slots].view(np.uint64)


@pytest.mark.parametrize("name,A", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_rbdia_structure_device_equals_host(ctx, name, A):
    """The device builder (wave-wide merge of the rows' offsets) gives the host
    builder's bundle offsets, offsets and lane masks exactly; both refuse a
    bundle wider than max_width."""
    ro, cols, _ = A
    n = len(ro) - 1
    say("case rbdia structure device/host", name, n)
    dro, dc = dev(ro), dev(cols)
    for mw in (64, 1 << 20):
        ref = aa.rbdia_build(ro, cols, max_width=mw)
        got = rbdia_device(ctx, dro, dc, n, mw)
        if ref is None:
            assert got is None
            continue
        for r, g in zip(ref, got):
            np.testing.assert_array_equal(g, r)


def test_rbdia_structure_device_rejects_unsorted_rows(ctx):
    """A row whose columns are not strictly ascending (descending pair, or a
    repeated column) has no RB-DIA structure, on the device as on the host."""
    ro, cols, _ = aa.gen_csr(aa.GEN_POISSON7, 6)
    n = len(ro) - 1
    for case in ("swap", "repeat"):
        c = cols.copy()
        k = ro[70]
        if case == "swap":
            c[k], c[k + 1] = c[k + 1], c[k]
        else:
            c[k + 1] = c[k]
        say("case rbdia structure unsorted", case)
        assert aa.rbdia_build(ro, c, max_width=1 << 20) is None
        assert rbdia_device(ctx, dev(ro), dev(c), n, 1 << 20) is None
    assert rbdia_device(ctx, dev(ro), dev(cols), n, 64) is not None


@pytest.mark.parametrize("name,A", SYSTEMS[:3], ids=[s[0] for s in SYSTEMS[:3]])
def test_spmv_bellpack_bit_exact(ctx, name, A):
    ro, cols, vals = A
    n = len(ro) - 1
    say("case spmv bellpack", name, n)
    bro, bc, bv, lm, um = aa.bellpack_build(ro, cols, vals, 1024, masks=True)
    x = np.random.default_rng(2).standard_normal(n)
    ref = orc.bellpack_spmv(n, bro, bc, bv, x)
    np.testing.assert_array_equal(ref, orc.spmv(ro, cols, vals, x))
    dbro, dbc, dbv, dx = dev(bro), dev(bc), dev(bv), dev(x)
    y = empty(n)
    call("alien_hip_spmv_bellpack_f64", ctx.h, n, 1024, p(dbro), p(dbc), p(dbv), p(dx), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:n], ref)
    # masked lower / upper products (addLMult / addUMult)
    y0 = np.random.default_rng(3).standard_normal(n)
    for upper, mask in ((0, lm), (1, um)):
        dy = dev(y0)
        call("alien_hip_addmult_bellpack_f64", ctx.h, upper, C.c_double(0.75), n, 1024, p(dbro), p(dbc), p(dbv),
             p(dev(mask)), p(dx), p(dy))
        got = host(dy, ctx)
        exp = orc.addmult(ro, cols, vals, 0.75, x, y0, bool(upper))
        np.testing.assert_allclose(got, exp, rtol=1e-14, atol=1e-14)
    # inverse diagonal, BEllPack rule (preset kept where a_ii is absent or zero)
    d = empty(n)
    call("alien_hip_assign_f64", ctx.h, n, C.c_double(1.0), p(d))
    call("alien_hip_invdiag_bellpack_f64", ctx.h, n, 1024, p(dbro), p(dbc), p(dbv), p(d))
    np.testing.assert_array_equal(host(d, ctx)[:n], orc.bellpack_invdiag(n, bro, bc, bv))


@pytest.mark.parametrize("slice_", [64, 100, 1024])
@pytest.mark.parametrize("name,A", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_bellpack_device_fill_and_spmv(ctx, name, A, slice_):
    """BEllPack built on the device equals the host builder (columns, values, masks, padding),
    and the SpMV (one wavefront per 64 rows of a slice when the slice height allows, else one
    lane per row) keeps the reference's per-row order: FP64 bit-exact with the oracle's
    MatrixInternal::mult restatement, FP32 with the sequential FP32 CSR sum."""
    ro, cols, vals = A
    n = len(ro) - 1
    say("case bellpack device fill + spmv", name, n, slice_)
    bro, bc, bv, lm, um = aa.bellpack_build(ro, cols, vals, slice_, masks=True)
    total = len(bc)
    dro, dc, dv, dbro = dev(ro), dev(cols), dev(vals), dev(bro)
    dd = dev(aa.host_dcol(ro, cols))
    gbc, gbv = empty(total, torch.int32), empty(total)
    glm, gum = empty(total, torch.uint8), empty(total, torch.uint8)
    call("alien_hip_bellpack_fill_f64", ctx.h, n, slice_, p(dro), p(dc), p(dv), p(dd), p(dbro), p(gbc), p(gbv),
         p(glm), p(gum))
    np.testing.assert_array_equal(host(gbc, ctx)[:total], bc)
    np.testing.assert_array_equal(host(gbv, ctx)[:total], bv)
    np.testing.assert_array_equal(host(glm, ctx)[:total], lm)
    np.testing.assert_array_equal(host(gum, ctx)[:total], um)
    x = np.random.default_rng(2).standard_normal(n)
    dx = dev(x)
    y = empty(n)
    call("alien_hip_spmv_bellpack_f64", ctx.h, n, slice_, p(dbro), p(gbc), p(gbv), p(dx), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:n], orc.bellpack_spmv(n, bro, bc, bv, x, slice_))
    gbv32 = empty(total, torch.float32)
    call("alien_hip_bellpack_fill_f32", ctx.h, n, slice_, p(dro), p(dc), p(dv), p(dbro), None, p(gbv32))
    np.testing.assert_array_equal(host(gbv32, ctx)[:total], bv.astype(np.float32))
    dx32 = dev(x.astype(np.float32))
    y32, yc32 = empty(n, torch.float32), empty(n, torch.float32)
    call("alien_hip_spmv_bellpack_f32", ctx.h, n, slice_, p(dbro), p(gbc), p(gbv32), p(dx32), p(y32))
    call("alien_hip_spmv_csr_f32", ctx.h, n, p(dro), p(dc), p(dev(vals.astype(np.float32))), p(dx32), p(yc32))
    np.testing.assert_array_equal(host(y32, ctx)[:n], host(yc32, ctx)[:n])


@pytest.mark.parametrize("name,A", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_addmult_and_invdiag_csr_bit_exact(ctx, name, A):
    ro, cols, vals = A
    n = len(ro) - 1
    say("case addmult/invdiag csr", name, n)
    d = aa.host_dcol(ro, cols)
    x = np.random.default_rng(4).standard_normal(n)
    y0 = np.random.default_rng(5).standard_normal(n)
    dro, dc, dd, dv, dx = dev(ro), dev(cols), dev(d), dev(vals), dev(x)
    for upper in (0, 1):
        dy = dev(y0)
        call("alien_hip_addmult_csr_f64", ctx.h, upper, C.c_double(-1.25), n, p(dro), p(dc), p(dd), p(dv), p(dx),
             p(dy))
        np.testing.assert_array_equal(host(dy, ctx), orc.addmult(ro, cols, vals, -1.25, x, y0, bool(upper)))
    has_diag = np.array([i in cols[ro[i]:ro[i + 1]] for i in range(n)])
    if has_diag.all():
        y = empty(n)
        call("alien_hip_invdiag_csr_f64", ctx.h, n, p(dd), p(dv), p(y))
        np.testing.assert_array_equal(host(y, ctx)[:n], orc.invdiag_csr(ro, cols, vals))


def test_spmv_bsr2_bit_exact(ctx):
    ro, cols, vals = aa.gen_bsr2(11, 9, 7, params=[2.0, 0.3], seed=9)
    nb = len(ro) - 1
    say("case bsr2", nb)
    x = np.random.default_rng(6).standard_normal(2 * nb)
    ref = orc.spmv_bsr(ro, cols, vals, x)
    y = empty(2 * nb)
    call("alien_hip_spmv_bsr2_f64", ctx.h, nb, p(dev(ro)), p(dev(cols)), p(dev(vals)), p(dev(x)), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:2 * nb], ref)
    y32 = empty(2 * nb, torch.float32)
    call("alien_hip_spmv_bsr2_f32", ctx.h, nb, p(dev(ro)), p(dev(cols)), p(dev(vals.astype(np.float32))),
         p(dev(x.astype(np.float32))), p(y32))
    np.testing.assert_allclose(host(y32, ctx)[:2 * nb], ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())


def ragged_bsr2(nb=700, seed=4):
    """Random 2x2 block pattern: empty rows, rows of up to 40 blocks, ascending columns."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, nb)
    lens[rng.choice(nb, 30, replace=False)] = 0
    lens[rng.choice(nb, 5, replace=False)] = 40
    ro = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(nb, l, replace=False)) for l in lens]).astype(np.int32)
    vals = rng.standard_normal(4 * len(cols))
    return ro, cols, vals


@pytest.mark.parametrize("case", ["grid 11x9x7", "ragged 700", "one row"])
def test_spmv_rbbsr2_bit_exact(ctx, case):
    """Row-bundled 2x2 blocks: FP64 bit-exact with the oracle's _seqMultBlock
    restatement, FP32 bit-exact with the block-CSR FP32 kernel (same order)."""
    if case == "grid 11x9x7":
        ro, cols, vals = aa.gen_bsr2(11, 9, 7, params=[2.0, 0.3], seed=9)
    elif case == "ragged 700":
        ro, cols, vals = ragged_bsr2()
    else:
        ro, cols, vals = aa.gen_bsr2(1, 1, 1, params=[2.0, 0.3], seed=1)
    nb = len(ro) - 1
    say("case rbbsr2", case, nb)
    x = np.random.default_rng(6).standard_normal(2 * nb)
    ref = orc.spmv_bsr(ro, cols, vals, x)
    bo = aa.rbcsr_bundle_offsets(ro)
    slots = int(bo[-1]) * 64
    dro, dc, dv, dbo = dev(ro), dev(cols), dev(vals), dev(bo)
    bc = empty(slots, torch.int32)
    bv = empty(4 * slots)
    call("alien_hip_rbbsr2_fill_f64", ctx.h, nb, p(dro), p(dc), p(dv), p(dbo), p(bc), p(bv))
    y = empty(2 * nb)
    call("alien_hip_spmv_rbbsr2_f64", ctx.h, nb, p(dbo), p(bc), p(bv), p(dev(x)), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:2 * nb], ref)
    bv32 = empty(4 * slots, torch.float32)
    call("alien_hip_rbbsr2_fill_f32", ctx.h, nb, p(dro), p(dc), p(dv), p(dbo), p(bc), p(bv32))
    x32 = dev(x.astype(np.float32))
    y32 = empty(2 * nb, torch.float32)
    call("alien_hip_spmv_rbbsr2_f32", ctx.h, nb, p(dbo), p(bc), p(bv32), p(x32), p(y32))
    yb = empty(2 * nb, torch.float32)
    call("alien_hip_spmv_bsr2_f32", ctx.h, nb, p(dro), p(dc), p(dev(vals.astype(np.float32))), p(x32), p(yb))
    np.testing.assert_array_equal(host(y32, ctx)[:2 * nb], host(yb, ctx)[:2 * nb])


@pytest.mark.parametrize("case", ["sizes 1..4", "sizes 1..7", "all 3"])
def test_spmv_vblock_bit_exact(ctx, case):
    """Variable-block product (_seqMultVariableBlock order) on the device, bit-exact
    with the oracle; the same matrix converted to scalar CSR on the host and run
    through the RB-CSR kernel gives the same vector bit for bit."""
    from test_oracle import random_vblock
    if case == "sizes 1..4":
        ro, cols, bcols, bsize, boff, vals = random_vblock(nb=3000, max_bs=4, seed=11)
    elif case == "sizes 1..7":
        ro, cols, bcols, bsize, boff, vals = random_vblock(nb=500, max_bs=7, seed=12)
    else:
        ro, cols, bcols, bsize, boff, vals = random_vblock(nb=1000, sizes=3, seed=13)
    nb = len(ro) - 1
    n = int(boff[-1] + bsize[-1])
    say("case vblock", case, nb, n)
    x = np.random.default_rng(3).standard_normal(n)
    ref = orc.spmv_vblock(ro, cols, bcols, bsize, boff, vals, x)
    dx = dev(x)
    y = empty(n)
    call("alien_hip_spmv_vblock_f64", ctx.h, nb, p(dev(ro)), p(dev(cols)), p(dev(bcols)), p(dev(bsize)),
         p(dev(boff)), p(dev(vals)), p(dx), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:n], ref)
    sro, scols, svals = aa.vblock_to_csr(ro, cols, bcols, bsize, boff, vals)
    bo = aa.rbcsr_bundle_offsets(sro)
    slots = int(bo[-1]) * 64
    dbo = dev(bo)
    bc = empty(slots, torch.int32)
    bv = empty(slots)
    call("alien_hip_rbcsr_fill_f64", ctx.h, n, p(dev(sro)), p(dev(scols)), p(dev(svals)), p(dbo), p(bc), p(bv))
    y2 = empty(n)
    call("alien_hip_spmv_rbcsr_f64", ctx.h, n, p(dbo), p(bc), p(bv), p(dx), p(y2))
    np.testing.assert_array_equal(host(y2, ctx)[:n], ref)


@pytest.mark.parametrize("n", [0, 1, 63, 1000, 1 << 20])
def test_blas1_bit_exact_and_dot_tolerance(ctx, n):
    say("case blas1", n)
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    dx, dy = dev(x), dev(y)
    out = empty(n)
    call("alien_hip_copy_f64", ctx.h, n, p(dx), p(out))
    np.testing.assert_array_equal(host(out, ctx)[:n], x)
    call("alien_hip_axpy_f64", ctx.h, n, C.c_double(-0.3), p(dx), p(dy))
    np.testing.assert_array_equal(host(dy, ctx)[:n], y + (-0.3) * x)
    y1 = y + (-0.3) * x
    call("alien_hip_scal_f64", ctx.h, n, C.c_double(1.7), p(dy))
    np.testing.assert_array_equal(host(dy, ctx)[:n], 1.7 * y1)
    call("alien_hip_pointwise_mult_f64", ctx.h, n, p(dx), p(dy), p(out))
    np.testing.assert_array_equal(host(out, ctx)[:n], x * (1.7 * y1))
    call("alien_hip_assign_f64", ctx.h, n, C.c_double(2.5), p(out))
    np.testing.assert_array_equal(host(out, ctx)[:n], np.full(n, 2.5))
    res = empty(1)
    call("alien_hip_dot_f64", ctx.h, n, p(dx), p(dy), p(res))
    exact = float(np.dot(x.astype(np.longdouble), (1.7 * y1).astype(np.longdouble)))
    tol = 1e-15 * max(1, n) ** 0.5 * float(np.abs(x * (1.7 * y1)).sum()) + 1e-300
    assert abs(host(res, ctx)[0] - exact) <= tol
    # determinism: same inputs, same bits
    res2 = empty(1)
    call("alien_hip_dot_f64", ctx.h, n, p(dx), p(dy), p(res2))
    assert host(res2, ctx)[0] == host(res, ctx)[0]
    # FP32 vectors, FP64 accumulation
    dx32, dy32 = dev(x.astype(np.float32)), dev(y.astype(np.float32))
    call("alien_hip_dot_f32", ctx.h, n, p(dx32), p(dy32), p(res))
    e32 = float(np.dot(x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)))
    assert abs(host(res, ctx)[0] - e32) <= 1e-14 * float(np.abs(x * y).sum()) + 1e-300
    # fused axpy + dot
    dy = dev(y)
    call("alien_hip_axpy_dot_f64", ctx.h, n, C.c_double(0.5), p(dx), p(dy), p(res))
    yn = y + 0.5 * x
    np.testing.assert_array_equal(host(dy, ctx)[:n], yn)
    assert abs(host(res, ctx)[0] - float(np.dot(yn, yn))) <= 1e-13 * float(np.dot(yn, yn)) + 1e-300


def test_spmv_dot_fused(ctx):
    ro, cols, vals = aa.gen_csr(aa.GEN_POISSON7, 24)
    n = len(ro) - 1
    say("case spmv+dot", n)
    x = np.random.default_rng(8).standard_normal(n)
    y = empty(n)
    res = empty(1)
    call("alien_hip_spmv_dot_csr_f64", ctx.h, n, p(dev(ro)), p(dev(cols)), p(dev(vals)), p(dev(x)), p(y), p(res))
    ref = orc.spmv(ro, cols, vals, x)
    np.testing.assert_array_equal(host(y, ctx)[:n], ref)
    assert abs(host(res, ctx)[0] - np.dot(ref, x)) <= 1e-13 * np.abs(ref * x).sum()


def level_schedule(ro, cols, d, upper):
    lv, nl = aa.host_levels(ro, cols, d, upper)
    ptr, rows = aa.group_by_level(lv, nl)
    return nl, ptr, rows


@pytest.mark.parametrize("name,A", SYSTEMS[:3], ids=[s[0] for s in SYSTEMS[:3]])
def test_ilu0_factor_and_solve_bit_exact(ctx, name, A):
    ro, cols, vals = A
    n = len(ro) - 1
    say("case ilu0", name, n)
    d = aa.host_dcol(ro, cols)
    nl, lptr, lrows = level_schedule(ro, cols, d, False)
    nu, uptr, urows = level_schedule(ro, cols, d, True)
    dro, dc, dd = dev(ro), dev(cols), dev(d)
    lu = dev(vals)
    call("alien_hip_ilu0_factor_f64", ctx.h, n, p(dro), p(dc), p(dd), p(lu), nl, aa._capi.iptr(lptr),
         p(dev(lrows)))
    ref_lu = orc.ilu0_factor(ro, cols, vals)
    np.testing.assert_array_equal(host(lu, ctx), ref_lu)
    yv = np.random.default_rng(9).standard_normal(n)
    t = empty(n)
    x = empty(n)
    call("alien_hip_ilu0_solve_f64", ctx.h, n, p(dro), p(dc), p(dd), p(lu), nl, aa._capi.iptr(lptr), p(dev(lrows)),
         nu, aa._capi.iptr(uptr), p(dev(urows)), p(dev(yv)), p(t), p(x))
    np.testing.assert_array_equal(host(x, ctx)[:n], orc.ilu0_solve(ro, cols, ref_lu, yv))


def test_ilu0_multicolour_permuted_bit_exact(ctx):
    ro, cols, vals = aa.gen_csr(aa.GEN_CONVDIFF7, 12)
    n = len(ro) - 1
    say("case ilu0 multicolour", n)
    c, nc = aa.host_color(ro, cols)
    perm = np.argsort(c, kind="stable")
    A = sp.csr_matrix((vals, cols, ro), shape=(n, n))[perm][:, perm].tocsr()
    A.sort_indices()
    pro, pcols, pvals = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
    d = aa.host_dcol(pro, pcols)
    nl, lptr, lrows = level_schedule(pro, pcols, d, False)
    nu, uptr, urows = level_schedule(pro, pcols, d, True)
    assert nl <= nc and nu <= nc
    lu = dev(pvals)
    call("alien_hip_ilu0_factor_f64", ctx.h, n, p(dev(pro)), p(dev(pcols)), p(dev(d)), p(lu), nl,
         aa._capi.iptr(lptr), p(dev(lrows)))
    np.testing.assert_array_equal(host(lu, ctx), orc.ilu0_factor(pro, pcols, pvals))


def test_full_size_spmv_properties_256(ctx):
    """256^3 log-normal system (N = 16.8M, nnz = 117M): RB-CSR and CSR kernels
    agree bit for bit, and A.1 equals the row sums computed on the host."""
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 256, params=[2.0, 1, 1, 1], seed=42)
    n = len(ro) - 1
    say("case full-size 256^3 spmv", n, len(vals))
    dro, dc, dv = dev(ro), dev(cols), dev(vals)
    ones = torch.ones(n, dtype=torch.float64, device="cuda")
    y = empty(n)
    call("alien_hip_spmv_csr_f64", ctx.h, n, p(dro), p(dc), p(dv), p(ones), p(y))
    rs = np.add.reduceat(vals, ro[:-1])  # sequential left-to-right per row == _seqMult order on ones
    yh = host(y, ctx)
    np.testing.assert_allclose(yh, rs, rtol=1e-12, atol=1e-12 * np.abs(vals).max())
    bo = aa.rbcsr_bundle_offsets(ro)
    dbo = dev(bo)
    bc = empty(int(bo[-1]) * 64, torch.int32)
    bv = empty(int(bo[-1]) * 64)
    call("alien_hip_rbcsr_fill_f64", ctx.h, n, p(dro), p(dc), p(dv), p(dbo), p(bc), p(bv))
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y1 = empty(n)
    y2 = empty(n)
    call("alien_hip_spmv_csr_f64", ctx.h, n, p(dro), p(dc), p(dv), p(x), p(y1))
    call("alien_hip_spmv_rbcsr_f64", ctx.h, n, p(dbo), p(bc), p(bv), p(x), p(y2))
    ctx.synchronize()
    assert torch.equal(y1, y2)
    del bc, bv
    # RB-DIA at full size: same sums bit for bit, FP64 and (against RB-CSR) FP32;
    # the device-built structure equals the host-built one
    dbo_d, doff, dmask = aa.rbdia_build(ro, cols)
    for r, g in zip((dbo_d, doff, dmask), rbdia_device(ctx, dro, dc, n, 64)):
        np.testing.assert_array_equal(g, r)
    assert int(dbo_d[-1]) * 64 <= 1.01 * len(vals) + 64 * len(dbo_d)
    ddbo, ddoff, ddm = dev(dbo_d), dev(doff), dev(dmask.view(np.int64))
    dvv = empty(int(dbo_d[-1]) * 64)
    call("alien_hip_rbdia_fill_f64", ctx.h, n, p(dro), p(dc), p(dv), p(ddbo), p(ddoff), p(dvv))
    y3 = empty(n)
    call("alien_hip_spmv_rbdia_f64", ctx.h, n, p(ddbo), p(ddoff), p(ddm), p(dvv), p(x), p(y3))
    ctx.synchronize()
    assert torch.equal(y1, y3)


def test_full_size_learned_factor_closed_form_256(ctx):
    """256^3 log-normal system: with the initial parameters the network's output
    is exactly the first-order factor, Ghat_ii = exp(0) = 1 and Ghat_ij = e_ij =
    -a_ij / sqrt(d_i d_j) (rounded to FP32 as the kernel does), for all 60M lower
    entries: a size-independent check of the full forward at its real shape."""
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 256, params=[2.0, 1, 1, 1], seed=42)
    n = len(ro) - 1
    say("case full-size learned factor closed form", n)
    gh = aa.gnn_forward(ctx, ro, cols, vals, aa.gnn_init(0))
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ro))
    d = vals[orc.dcol(ro, cols)]
    dis = (1.0 / np.sqrt(d)).astype(np.float32).astype(np.float64)
    e = ((-vals * dis[rows]) * dis[cols]).astype(np.float32)
    ref = np.where(cols < rows, e, np.float32(0.0))
    ref[cols == rows] = 1.0
    np.testing.assert_array_equal(gh, ref)


def test_filu0_factor_sweeps_bit_exact(ctx):
    """The GPU's fixed-point sweeps give the oracle's factor bit for bit (same
    per-row operations), and reach the exact ILU(0) factor after as many sweeps as
    the lower DAG has levels."""
    ro, cols, vals = aa.gen_csr(aa.GEN_CONVDIFF7, 9, 7, 5)
    n = len(ro) - 1
    d = orc.dcol(ro, cols)
    _, nl = orc.levels(ro, cols, upper=False)
    for niter in (1, 3, nl):
        say("case filu0 sweeps", niter)
        dro, dc, dd, da = dev(ro), dev(cols), dev(d), dev(vals)
        cur = dev(vals)
        for _ in range(niter):
            prev = cur.clone()
            torch.cuda.synchronize()
            call("alien_hip_filu0_sweep_f64", ctx.h, n, p(dro), p(dc), p(dd), p(da), p(prev), p(cur))
        np.testing.assert_array_equal(host(cur, ctx), orc.filu0_factor(ro, cols, vals, niter))
    np.testing.assert_array_equal(orc.filu0_factor(ro, cols, vals, nl), orc.ilu0_factor(ro, cols, vals))


def pair_structure_brute(ro, cols):
    """Pair-mode RB-DIA structure by brute force: per bundle of 64 rows the sorted
    distinct offsets c - (r & ~1) of the rows' even columns, one mask bit per row."""
    n = len(ro) - 1
    bo, doff, dmask = [0], [], []
    for b in range((n + 63) // 64):
        offs = {}
        for r in range(b * 64, min(n, b * 64 + 64)):
            for k in range(ro[r], ro[r + 1], 2):
                offs.setdefault(int(cols[k]) - (r & ~1), 0)
                offs[int(cols[k]) - (r & ~1)] |= 1 << (r - b * 64)
        for d in sorted(offs):
            doff.append(d)
            dmask.append(offs[d])
        bo.append(len(doff))
    return np.array(bo), np.array(doff), np.array(dmask, dtype=np.uint64)


@pytest.mark.parametrize("n", [5, 12])
def test_rbdia_pair_mode_bit_exact(ctx, n):
    """Pair-mode RB-DIA (two unknowns per cell, interleaved): structure equal to
    brute force, FP64 SpMV bit-exact with the sequential CSR product, FP32 with
    the FP32 RB-CSR kernel; a matrix whose rows are not column pairs has none."""
    import scipy.sparse as sp
    bro, bcols, bvals = aa.gen_bsr2(n, params=[2.0, 0.3], seed=5)
    nb = len(bro) - 1
    blocks = bvals.reshape(-1, 2, 2).transpose(0, 2, 1)
    A = sp.bsr_matrix((blocks, bcols, bro), shape=(2 * nb, 2 * nb)).tocsr()
    A.sort_indices()
    ro, cols, vals = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    N = len(ro) - 1
    say("case rbdia pair", n, N)
    dro, dc, dv = dev(ro), dev(cols), dev(vals)
    got = rbdia_device(ctx, dro, dc, N, 64, pair=1)
    assert got is not None
    for g, e in zip(got, pair_structure_brute(ro, cols)):
        np.testing.assert_array_equal(g.astype(np.int64) if g.dtype != np.uint64 else g, e)
    bo, doff, dmask = got
    slots = int(bo[-1]) * 64
    dbo, ddo, ddm = dev(bo), dev(doff), dev(dmask.view(np.int64))
    bv = empty(2 * slots)
    call("alien_hip_rbdia2_fill_f64", ctx.h, N, p(dro), p(dc), p(dv), p(dbo), p(ddo), p(bv))
    x = np.random.default_rng(3).standard_normal(N)
    y = empty(N)
    call("alien_hip_spmv_rbdia2_f64", ctx.h, N, p(dbo), p(ddo), p(ddm), p(bv), p(dev(x)), p(y))
    np.testing.assert_array_equal(host(y, ctx)[:N], orc.spmv(ro, cols, vals, x))
    bv32 = empty(2 * slots, torch.float32)
    call("alien_hip_rbdia2_fill_f32", ctx.h, N, p(dro), p(dc), p(dv), p(dbo), p(ddo), p(bv32))
    x32 = dev(x.astype(np.float32))
    y32 = empty(N, torch.float32)
    call("alien_hip_spmv_rbdia2_f32", ctx.h, N, p(dbo), p(ddo), p(ddm), p(bv32), p(x32), p(y32))
    cbo = aa.rbcsr_bundle_offsets(ro)
    cslots = int(cbo[-1]) * 64
    dcbo = dev(cbo)
    cbc = empty(cslots, torch.int32)
    cbv = empty(cslots, torch.float32)
    call("alien_hip_rbcsr_fill_f32", ctx.h, N, p(dro), p(dc), p(dv), p(dcbo), p(cbc), p(cbv))
    yc = empty(N, torch.float32)
    call("alien_hip_spmv_rbcsr_f32", ctx.h, N, p(dcbo), p(cbc), p(cbv), p(x32), p(yc))
    np.testing.assert_array_equal(host(y32, ctx)[:N], host(yc, ctx)[:N])
    pro, pcols, _ = aa.gen_csr(aa.GEN_POISSON7, 6)
    assert rbdia_device(ctx, dev(pro), dev(pcols), len(pro) - 1, 64, pair=1) is None
