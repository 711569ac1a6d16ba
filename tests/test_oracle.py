"""Pins the CPU oracle (oracle/alien_oracle.cpp) against what it must reproduce:

* golden runs of the reference's own Krylov templates on the 32^3 Poisson
  system (recorded values from compiling the reference's templates: Chebyshev-CG 38 iterations,
  relative residual 7.03e-9, ALPHA 3.6392 / BETA 1.8196; Diag-CG 108
  iterations; Chebyshev factor 30 not converged after 1000 iterations,
  relative residual 3.9e-2);
* the known-answer tests of src/refsemantic/tests/TestMVExpr.cc:105-171;
* closed forms and brute force (dense matvec, dense ILU(0) by Saad's IKJ
  algorithm on a pattern mask, the ILU(0) identity (LU)_ij = a_ij on the
  pattern, exact LU on tridiagonal matrices, scipy CSR products, level and
  colouring invariants).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import alien_amd as aa
import oracle as orc


def poisson(n):
    return aa.gen_csr(aa.GEN_POISSON7, n)


def c0_system():
    ro, cols, vals = poisson(32)
    xt = aa.uniform01(42, 32 ** 3)
    b = orc.spmv(ro, cols, vals, xt)
    return ro, cols, vals, b


def random_csr(n, density, seed, diag=True, empty_rows=False):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format="csr")
    if diag:
        A = A + sp.diags(rng.uniform(4, 8, n))
    A = A.tocsr()
    if empty_rows:
        A = A.tolil()
        for i in rng.choice(n, size=max(1, n // 10), replace=False):
            A.rows[i] = []
            A.data[i] = []
        A = A.tocsr()
    A.sort_indices()
    return (A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


# ---- golden runs of the reference templates ------------------------------------------------
def test_golden_c0_chebyshev_cg():
    ro, cols, vals, b = c0_system()
    x, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_CG, precond=orc.PREC_CHEB, tol=1e-8, max_iter=1000,
                     poly_factor=0.5, poly_order=3, poly_max_iter=10)
    assert r["iterations"] == 38 and r["status"] == 1
    assert r["true_rel_res"] == pytest.approx(7.03e-9, rel=2e-3)
    assert r["cheb_alpha"] == pytest.approx(3.6392, abs=1e-4)
    assert r["cheb_beta"] == pytest.approx(1.8196, abs=1e-4)


def test_golden_c0_diag_cg():
    ro, cols, vals, b = c0_system()
    _, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_CG, precond=orc.PREC_DIAG)
    assert r["iterations"] == 108 and r["status"] == 1


def test_golden_c0_chebyshev_factor30_stalls():
    ro, cols, vals, b = c0_system()
    _, r = orc.solve(ro, cols, vals, b, precond=orc.PREC_CHEB, poly_factor=30.0)
    assert r["iterations"] == 1000 and r["status"] == 0
    assert r["true_rel_res"] == pytest.approx(3.9e-2, rel=2e-2)


# ---- TestMVExpr.cc known answers (10x10 tridiag(-1,2,-1)) -----------------------------------
def tridiag10():
    n = 10
    A = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def test_testmvexpr_known_answers():
    ro, cols, vals = tridiag10()
    x = np.ones(10)
    assert np.sqrt(orc.dot(x, x)) == np.sqrt(10)                        # :105
    y = orc.spmv(ro, cols, vals, x)
    assert np.sqrt(orc.dot(y, y)) == np.sqrt(2)                         # :123
    r = y - orc.spmv(ro, cols, vals, x)
    assert np.sqrt(orc.dot(r, r)) == 0.0                                # :128
    yi = np.arange(10, dtype=np.float64)
    assert orc.dot(x, yi) == 45                                         # :148
    assert orc.dot(0.5 * x, yi) == 22.5                                 # :153
    assert orc.dot(x, orc.spmv(ro, cols, vals, yi)) == 9                # :160
    y2 = orc.spmv(ro, cols, vals, 0.5 * x)
    assert np.sqrt(orc.dot(y2, y2)) == np.sqrt(0.5)                     # :141


# ---- SpMV --------------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spmv_matches_scipy_and_dense(seed):
    ro, cols, vals = random_csr(300, 0.02, seed, empty_rows=True)
    A = sp.csr_matrix((vals, cols, ro), shape=(300, 300))
    x = np.random.default_rng(seed).standard_normal(300)
    y = orc.spmv(ro, cols, vals, x)
    np.testing.assert_allclose(y, A @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(y, A.toarray() @ x, rtol=1e-13, atol=1e-13)


def test_spmv_poisson_row_sums_closed_form():
    n = 9
    ro, cols, vals = poisson(n)
    y = orc.spmv(ro, cols, vals, np.ones(n ** 3))
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    faces = sum((c == e).astype(int) for c in (i, j, k) for e in (0, n - 1)).transpose(2, 1, 0).ravel()
    np.testing.assert_array_equal(y, faces.astype(np.float64))


def test_bsr_matches_dense_blocks():
    ro, cols, vals = aa.gen_bsr2(4, 3, 2, params=[1.0, 0.3], seed=5)
    nb = len(ro) - 1
    D = np.zeros((2 * nb, 2 * nb))
    for i in range(nb):
        for k in range(ro[i], ro[i + 1]):
            m = vals[4 * k:4 * k + 4].reshape(2, 2, order="F")  # m[ieq + 2 iu]
            D[2 * i:2 * i + 2, 2 * cols[k]:2 * cols[k] + 2] = m
    x = np.random.default_rng(1).standard_normal(2 * nb)
    np.testing.assert_allclose(orc.spmv_bsr(ro, cols, vals, x), D @ x, rtol=1e-13, atol=1e-12)


def test_addlmult_addumult_brute_force():
    ro, cols, vals = random_csr(200, 0.03, 7)
    A = sp.csr_matrix((vals, cols, ro), shape=(200, 200)).toarray()
    x = np.random.default_rng(3).standard_normal(200)
    y0 = np.random.default_rng(4).standard_normal(200)
    np.testing.assert_allclose(orc.addmult(ro, cols, vals, 0.7, x, y0, False), y0 + 0.7 * np.tril(A, -1) @ x,
                               rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(orc.addmult(ro, cols, vals, -1.3, x, y0, True), y0 - 1.3 * np.triu(A, 1) @ x,
                               rtol=1e-13, atol=1e-13)


def test_dcol_brute_force():
    ro, cols, vals = random_csr(150, 0.05, 9, diag=False, empty_rows=True)
    d = orc.dcol(ro, cols)
    for i in range(150):
        row = cols[ro[i]:ro[i + 1]]
        assert d[i] == ro[i] + int(np.sum(row < i))


# ---- BEllPack (SYCLBEllPackMatrix.cc:69-88,127-170,235-315,422-460) --------------------------------
def brute_bellpack(ro, cols, vals, S):
    n = len(ro) - 1
    ns = (n + S - 1) // S
    widths = [max([ro[i + 1] - ro[i] for i in range(s * S, min(n, s * S + S))] + [0]) for s in range(ns)]
    bro = np.concatenate([[0], np.cumsum(widths)]).astype(np.int32)
    bc = np.zeros(bro[-1] * S, np.int32)
    bv = np.zeros(bro[-1] * S)
    for i in range(n):
        s, l = divmod(i, S)
        for j in range(ro[i + 1] - ro[i]):
            bc[bro[s] * S + j * S + l] = cols[ro[i] + j]
            bv[bro[s] * S + j * S + l] = vals[ro[i] + j]
    return bro, bc, bv


@pytest.mark.parametrize("S", [4, 16, 1024])
def test_bellpack_layout_brute_force(S):
    ro, cols, vals = random_csr(257, 0.03, 11, empty_rows=True)
    bro, bc, bv = orc.bellpack(ro, cols, vals, S)
    ebro, ebc, ebv = brute_bellpack(ro, cols, vals, S)
    np.testing.assert_array_equal(bro, ebro)
    np.testing.assert_array_equal(bc, ebc)
    np.testing.assert_array_equal(bv, ebv)
    x = np.random.default_rng(0).standard_normal(257)
    np.testing.assert_array_equal(orc.bellpack_spmv(257, bro, bc, bv, x, S), orc.spmv(ro, cols, vals, x))
    lm, um = orc.bellpack_masks(ro, cols, bro, S)
    d = orc.dcol(ro, cols)
    for i in range(257):
        s, l = divmod(i, S)
        for j in range(ro[i + 1] - ro[i]):
            e = bro[s] * S + j * S + l
            assert lm[e] == (ro[i] + j < d[i]) and um[e] == (ro[i] + j > d[i])


def test_bellpack_invdiag_rule():
    ro, cols, vals = random_csr(100, 0.05, 12)
    vals = vals.copy()
    d = orc.dcol(ro, cols)
    vals[d[5]] = 0.0  # zero diagonal: keeps the preset
    bro, bc, bv = orc.bellpack(ro, cols, vals, 16)
    y = orc.bellpack_invdiag(100, bro, bc, bv, preset=1.0, slice_=16)
    assert y[5] == 1.0
    mask = np.arange(100) != 5
    np.testing.assert_array_equal(y[mask], (1.0 / vals[d])[mask])
    np.testing.assert_array_equal(orc.invdiag_csr(ro, cols, vals)[mask], (1.0 / vals[d])[mask])


# ---- ILU(0) ----------------------------------------------------------------------------------------
def dense_ilu0_ikj(A, pattern):
    """Saad, Iterative Methods for Sparse Linear Systems, Alg. 10.4 (IKJ) on a dense copy."""
    a = A.copy()
    n = a.shape[0]
    for i in range(1, n):
        for k in range(i):
            if not pattern[i, k]:
                continue
            a[i, k] = a[i, k] / a[k, k]
            for j in range(k + 1, n):
                if pattern[i, j]:
                    a[i, j] = a[i, j] - a[i, k] * a[k, j]
    return a


@pytest.mark.parametrize("seed", [0, 1])
def test_ilu0_matches_textbook_and_pattern_identity(seed):
    ro, cols, vals = random_csr(60, 0.08, seed)
    A = sp.csr_matrix((vals, cols, ro), shape=(60, 60)).toarray()
    pattern = A != 0
    lu = orc.ilu0_factor(ro, cols, vals)
    F = sp.csr_matrix((lu, cols, ro), shape=(60, 60)).toarray()
    np.testing.assert_allclose(F[pattern], dense_ilu0_ikj(A, pattern)[pattern], rtol=1e-12, atol=1e-12)
    L = np.tril(F, -1) + np.eye(60)
    U = np.triu(F)
    np.testing.assert_allclose((L @ U)[pattern], A[pattern], rtol=1e-11, atol=1e-11)
    y = np.random.default_rng(seed).standard_normal(60)
    np.testing.assert_allclose(orc.ilu0_solve(ro, cols, lu, y), np.linalg.solve(L @ U, y), rtol=1e-10, atol=1e-10)


def test_ilu0_tridiagonal_is_exact_lu():
    ro, cols, vals = tridiag10()
    lu = orc.ilu0_factor(ro, cols, vals)
    y = np.arange(1.0, 11.0)
    A = sp.csr_matrix((vals, cols, ro), shape=(10, 10)).toarray()
    np.testing.assert_allclose(orc.ilu0_solve(ro, cols, lu, y), np.linalg.solve(A, y), rtol=1e-13)


# ---- schedules ---------------------------------------------------------------------------------------
def test_levels_poisson_closed_form():
    n = 7
    ro, cols, _ = poisson(n)
    lv, nl = orc.levels(ro, cols, upper=False)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    np.testing.assert_array_equal(lv, (i + j + k).transpose(2, 1, 0).ravel())
    assert nl == 3 * n - 2
    lu, nu = orc.levels(ro, cols, upper=True)
    np.testing.assert_array_equal(lu, (3 * (n - 1) - (i + j + k)).transpose(2, 1, 0).ravel())
    assert nu == 3 * n - 2


def test_levels_are_longest_paths():
    ro, cols, _ = random_csr(200, 0.03, 5)
    d = orc.dcol(ro, cols)
    lv, nl = orc.levels(ro, cols)
    for i in range(200):
        deps = cols[ro[i]:d[i]]
        assert lv[i] == (0 if len(deps) == 0 else 1 + lv[deps].max())
    assert nl == lv.max() + 1


def test_color_poisson_is_red_black_and_valid():
    n = 6
    ro, cols, _ = poisson(n)
    c, nc = orc.color(ro, cols)
    assert nc == 2
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    np.testing.assert_array_equal(c, ((i + j + k) % 2).transpose(2, 1, 0).ravel())
    ro, cols, _ = random_csr(300, 0.02, 8)
    c, nc = orc.color(ro, cols)
    for r in range(300):
        for q in cols[ro[r]:ro[r + 1]]:
            if q != r:
                assert c[q] != c[r]


# ---- solvers: textbook behaviour ---------------------------------------------------------------------------
def test_cg_and_bicgstab_reach_numpy_solution():
    ro, cols, vals = poisson(6)
    A = sp.csr_matrix((vals, cols, ro), shape=(216, 216)).toarray()
    b = np.random.default_rng(2).uniform(size=216)
    for prec in (orc.PREC_DIAG, orc.PREC_CHEB, orc.PREC_ILU0, orc.PREC_NONE):
        for solver in (orc.ORC_CG, orc.ORC_BICGS):
            x, r = orc.solve(ro, cols, vals, b, solver=solver, precond=prec, tol=1e-10)
            assert r["status"] == 1, (solver, prec)
            np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-9)


def test_bicgstab_nonsymmetric_convdiff():
    ro, cols, vals = aa.gen_csr(aa.GEN_CONVDIFF7, 8)
    A = sp.csr_matrix((vals, cols, ro), shape=(512, 512)).toarray()
    assert not np.allclose(A, A.T)
    b = A @ aa.uniform01(42, 512)
    x, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_BICGS, precond=orc.PREC_ILU0, tol=1e-10)
    assert r["status"] == 1
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-7)


def test_null_rhs_and_max_iter():
    ro, cols, vals = poisson(5)
    x, r = orc.solve(ro, cols, vals, np.zeros(125))
    assert r["iterations"] == 0 and r["status"] == 1
    b = np.ones(125)
    _, r = orc.solve(ro, cols, vals, b, max_iter=1, precond=orc.PREC_DIAG)
    assert r["iterations"] == 1 and r["status"] == 0


def test_cg_residual_orthogonality_small_spd():
    # CG's Krylov subspace property: on an n x n SPD system CG is exact after n steps
    ro, cols, vals = tridiag10()
    b = np.arange(1.0, 11.0)
    x, r = orc.solve(ro, cols, vals, b, precond=orc.PREC_NONE, tol=1e-14, max_iter=10)
    A = sp.csr_matrix((vals, cols, ro), shape=(10, 10)).toarray()
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10)
    assert r["iterations"] <= 10


def test_threaded_oracle_matches_sequential_iterations():
    ro, cols, vals, b = c0_system()
    _, r1 = orc.solve(ro, cols, vals, b, precond=orc.PREC_CHEB)
    _, r4 = orc.solve(ro, cols, vals, b, precond=orc.PREC_CHEB, nthreads=4)
    assert abs(r1["iterations"] - r4["iterations"]) <= 2 and r4["status"] == 1


def test_filu0_factor_fixed_point_reaches_ilu0():
    """FILU0's fixed-point sweeps (FILU0Preconditioner.h:146-170): after as many
    sweeps as the lower DAG has levels every row has read final rows, so the factor
    equals the sequential ILU(0) factor bit for bit (Chow & Patel 2015); before
    that it differs."""
    ro, cols, vals = aa.gen_csr(aa.GEN_CONVDIFF7, 6, 5, 4)
    exact = orc.ilu0_factor(ro, cols, vals)
    d = orc.dcol(ro, cols)
    lv, nl = orc.levels(ro, cols, upper=False)
    np.testing.assert_array_equal(orc.filu0_factor(ro, cols, vals, nl), exact)
    np.testing.assert_array_equal(orc.filu0_factor(ro, cols, vals, nl + 3), exact)
    assert not np.array_equal(orc.filu0_factor(ro, cols, vals, 1), exact)
    # one sweep: rows read A's own values for the other rows (closed form on row 1)
    lu1 = orc.filu0_factor(ro, cols, vals, 1)
    np.testing.assert_array_equal(lu1[ro[0]:ro[1]], vals[ro[0]:ro[1]])


def test_filu0_solver_sweeps_converge_to_ilu0():
    """With as many Jacobi sweeps as levels, FILU0's triangular solves equal the
    exact ILU(0) solves (L part bit for bit; U within rounding of 1/u_ii), so a
    BiCGStab run with FILU0 matches the ILU0 run's iteration count."""
    ro, cols, vals = aa.gen_csr(aa.GEN_CONVDIFF7, 8)
    n = len(ro) - 1
    b = orc.spmv(ro, cols, vals, aa.uniform01(42, n))
    d = orc.dcol(ro, cols)
    _, nl = orc.levels(ro, cols, upper=False)
    _, ri = orc.solve(ro, cols, vals, b, solver=orc.ORC_BICGS, precond=orc.PREC_ILU0, tol=1e-8)
    _, rf = orc.solve(ro, cols, vals, b, solver=orc.ORC_BICGS, precond=orc.PREC_FILU0, tol=1e-8,
                      filu_factor_niter=0, filu_solver_niter=nl)
    assert rf["status"] == ri["status"] == 1
    assert abs(rf["iterations"] - ri["iterations"]) <= 1
    _, r3 = orc.solve(ro, cols, vals, b, solver=orc.ORC_BICGS, precond=orc.PREC_FILU0, tol=1e-8,
                      filu_factor_niter=2, filu_solver_niter=3)
    assert r3["status"] == 1 and r3["iterations"] >= ri["iterations"]


def test_neumann_polynomial_closed_form():
    """NeumannPolyPreconditioner (NeumannPolyPreconditioner.h:85-175): the apply
    equals factor * sum_{k<=order} (I - factor A)^k x (dense check), and CG with
    it converges on the 7-point Laplacian for a factor below 2 / lambda_max."""
    ro, cols, vals = aa.gen_csr(aa.GEN_POISSON7, 4)
    n = len(ro) - 1
    A = np.zeros((n, n))
    for i in range(n):
        A[i, cols[ro[i]:ro[i + 1]]] = vals[ro[i]:ro[i + 1]]
    f, order = 0.125, 3
    S = sum(np.linalg.matrix_power(np.eye(n) - f * A, k) for k in range(order + 1))
    b = aa.uniform01(3, n)
    # one CG step with M = f S from x0 = 0: x1 = (r.z / z.A z) z, z = f S b
    x1, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_CG, precond=orc.PREC_NEUMANN, tol=1e-30, max_iter=1,
                      poly_factor=f, poly_order=order)
    z = f * S @ b
    np.testing.assert_allclose(x1, (b @ z) / (z @ A @ z) * z, rtol=1e-12)
    _, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_CG, precond=orc.PREC_NEUMANN, tol=1e-10, max_iter=500,
                     poly_factor=f, poly_order=order)
    assert r["status"] == 1


def random_vblock(nb=40, max_bs=4, seed=3, sizes=None):
    """Random variable-block matrix (ascending block columns, diagonal block in every
    row) with start offsets for the blocks' values (CSRStructInfo::getBlockCols)."""
    rng = np.random.default_rng(seed)
    bsize = (rng.integers(1, max_bs + 1, nb) if sizes is None else np.full(nb, sizes)).astype(np.int32)
    boff = np.concatenate([[0], np.cumsum(bsize)[:-1]]).astype(np.int32)
    rows = [np.union1d([i], rng.choice(nb, rng.integers(0, 6), replace=False)) for i in range(nb)]
    ro = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    cols = np.concatenate(rows).astype(np.int32)
    sizes_k = np.array([bsize[i] * bsize[c] for i, r in enumerate(rows) for c in r])
    bcols = np.concatenate([[0], np.cumsum(sizes_k)[:-1]]).astype(np.int32)
    vals = rng.standard_normal(int(sizes_k.sum()))
    return ro, cols, bcols, bsize, boff, vals


def dense_vblock(ro, cols, bcols, bsize, boff, vals):
    """The scalar matrix, assembled block by block with numpy (row-major blocks)."""
    n = int(boff[-1] + bsize[-1])
    A = np.zeros((n, n))
    for i in range(len(ro) - 1):
        for j in range(ro[i], ro[i + 1]):
            c = cols[j]
            blk = vals[bcols[j]:bcols[j] + bsize[i] * bsize[c]].reshape(bsize[i], bsize[c])
            A[boff[i]:boff[i] + bsize[i], boff[c]:boff[c] + bsize[c]] = blk
    return A


def test_vblock_spmv_pins():
    """_seqMultVariableBlock restatement: against the block-assembled dense product
    (random sizes 1..4, non-symmetric blocks, so a transposed block fails), against
    the scalar CSR product bit for bit when every block is 1x1, and against the
    fixed-block product (_seqMultBlock, column-major) bit for bit when every block
    is 2x2 and stored transposed."""
    ro, cols, bcols, bsize, boff, vals = random_vblock()
    A = dense_vblock(ro, cols, bcols, bsize, boff, vals)
    x = np.random.default_rng(9).standard_normal(A.shape[0])
    y = orc.spmv_vblock(ro, cols, bcols, bsize, boff, vals, x)
    np.testing.assert_allclose(y, A @ x, rtol=1e-13, atol=1e-13 * np.abs(A).sum(1).max())
    # every block 1x1: the CSR product
    ro1, cols1, bcols1, bs1, bo1, v1 = random_vblock(sizes=1, seed=4)
    x1 = np.random.default_rng(1).standard_normal(len(bs1))
    np.testing.assert_array_equal(orc.spmv_vblock(ro1, cols1, bcols1, bs1, bo1, v1, x1), orc.spmv(ro1, cols1, v1, x1))
    # every block 2x2: the fixed-block product with column-major blocks
    ro2, cols2, bcols2, bs2, bo2, v2 = random_vblock(sizes=2, seed=5)
    x2 = np.random.default_rng(2).standard_normal(2 * len(bs2))
    colmajor = v2.reshape(-1, 2, 2).transpose(0, 2, 1).reshape(-1)
    np.testing.assert_array_equal(orc.spmv_vblock(ro2, cols2, bcols2, bs2, bo2, v2, x2),
                                  orc.spmv_bsr(ro2, cols2, colmajor, x2))


def test_factor_preconditioner_pins():
    """The factored preconditioner y = G (G^T x) in the reference's CG / BiCGStab call
    sequence: the identity factor is the unpreconditioned solver (same iterations,
    same x bit for bit), G = D^-1/2 is Jacobi up to rounding (iterations within 1),
    and the exact inverse factor (G G^T = A^-1, dense Cholesky) solves in one step."""
    import scipy.sparse as sp
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 8, params=[1.5, 1, 1, 1], seed=3)
    n = len(ro) - 1
    b = aa.uniform01(5, n)
    A = sp.csr_matrix((vals, cols, ro), shape=(n, n))
    I = sp.identity(n, format="csr")
    for solver in (orc.ORC_CG, orc.ORC_BICGS):
        x1, r1 = orc.solve(ro, cols, vals, b, solver=solver, precond=orc.PREC_FACTOR, factor=(I, I), max_iter=2000)
        x0, r0 = orc.solve(ro, cols, vals, b, solver=solver, precond=orc.PREC_NONE, max_iter=2000)
        assert r1["iterations"] == r0["iterations"] and r1["status"] == 1
        np.testing.assert_array_equal(x1, x0)
        D = sp.diags(1.0 / np.sqrt(A.diagonal())).tocsr()
        _, rd = orc.solve(ro, cols, vals, b, solver=solver, precond=orc.PREC_FACTOR, factor=(D, D), max_iter=2000)
        _, rj = orc.solve(ro, cols, vals, b, solver=solver, precond=orc.PREC_DIAG, max_iter=2000)
        assert abs(rd["iterations"] - rj["iterations"]) <= 1
    C = np.linalg.cholesky(np.linalg.inv(A.toarray()))
    G = sp.csr_matrix(C)
    G.sort_indices()
    GT = G.T.tocsr()
    GT.sort_indices()
    _, re = orc.solve(ro, cols, vals, b, solver=orc.ORC_CG, precond=orc.PREC_FACTOR, factor=(G, GT), max_iter=10)
    assert re["iterations"] <= 2 and re["status"] == 1


# ---- Chebyshev: power method and _algo2 against what the mathematics fixes ------------------------
def _dense(ro, cols, vals):
    n = len(ro) - 1
    return sp.csr_matrix((vals, cols, ro), shape=(n, n)).toarray()


@pytest.mark.parametrize("n", [4, 5, 6])
def test_power_method_reaches_laplacian_extreme_eigenvalues(n):
    """computeBeta/computeAlpha (ChebyshevPreconditioner.h:118-202) with many iterations:
    on the n^3 7-point Laplacian D^-1 A = I - adj/6 has lambda = 1 -+ cos(pi/(n+1))
    at its ends. factor = 0 keeps computeAlpha's result (the shifted power method,
    converging to lambda_min - beta, plus beta); factor 0.5 overwrites it with beta/0.5."""
    ro, cols, vals = poisson(n)
    a, b = orc.cheb_bounds(ro, cols, vals, factor=0.0, max_it=3000)
    assert b == pytest.approx(1 + np.cos(np.pi / (n + 1)), abs=1e-12)
    assert a == pytest.approx(1 - np.cos(np.pi / (n + 1)), abs=1e-12)
    a2, b2 = orc.cheb_bounds(ro, cols, vals, factor=0.5, max_it=3000)
    assert b2 == b and a2 == b / 0.5  # Q1: alpha = beta / factor (:99-100)


def test_power_method_uses_the_inverse_diagonal():
    """Non-constant diagonal (log-normal coefficients): the bounds are the extreme
    eigenvalues of D^-1 A (= those of D^-1/2 A D^-1/2), not of A."""
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 5, params=[1.0, 1.0, 1.0, 1.0], seed=3)
    A = _dense(ro, cols, vals)
    dh = np.sqrt(np.diag(A))
    ev = np.linalg.eigvalsh(A / np.outer(dh, dh))
    a, b = orc.cheb_bounds(ro, cols, vals, factor=0.0, max_it=1000)
    assert b == pytest.approx(ev[-1], abs=1e-12)
    assert a == pytest.approx(ev[0], abs=1e-12)
    assert abs(b - np.linalg.eigvalsh(A)[-1]) > 1.0


def test_cheb_apply_order1_and_order2_closed_forms():
    """_algo2 starts from y = x (:399), first step alpha = 2/d (:409):
    order 1: y = x + (2/d) D^-1 (x - A x).
    order 2, expanded by hand: R_2(t) = 1 - (a0 + a1 + a1 b0) t + a0 a1 t^2 with
    a0 = 2/d, b0 = a0 c^2/4, a1 = 1/(d - b0), and y = A^-1 x - R_2(D^-1 A)(A^-1 x - x)."""
    import pin_systems as ps
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 4, params=[1.0, 1.0, 1.0, 1.0], seed=1)
    A = _dense(ro, cols, vals)
    x = np.random.default_rng(5).standard_normal(len(ro) - 1)
    for al, be in ((0.05, 1.9), (3.64, 1.82)):
        d, c = (be + al) / 2, (be - al) / 2
        y1 = orc.cheb_apply(ro, cols, vals, al, be, 1, x)
        np.testing.assert_allclose(y1, x + (2 / d) * (x - A @ x) / np.diag(A), rtol=1e-13, atol=1e-13)
        a0 = 2 / d
        b0 = a0 * c * c / 4
        a1 = 1 / (d - b0)
        R2 = np.array([1.0, -(a0 + a1 + a1 * b0), a0 * a1])
        np.testing.assert_allclose(ps.cheb_residual_poly(al, be, 2), R2, rtol=1e-15)
        ainv = np.linalg.solve(A, x)
        y2 = orc.cheb_apply(ro, cols, vals, al, be, 2, x)
        pred = ainv - ps.apply_poly_of_jacobi_operator(A, R2, ainv - x)
        assert np.abs(y2 - pred).max() <= 1e-12 * np.abs(y2).max()


@pytest.mark.parametrize("bounds", [(0.05, 1.9), (3.6392, 1.8196)], ids=["sane", "factor0.5"])
def test_cheb_apply_is_its_residual_polynomial(bounds):
    """Orders 1-6 on a log-normal system: y_k = A^-1 x - R_k(D^-1 A)(A^-1 x - x), R_k
    evaluated on the eigen-decomposition of D^-1/2 A D^-1/2 (no vector recurrence). A
    start from y = 0, or d and c exchanged, misses by O(1) (checked, so the test can fail)."""
    import pin_systems as ps
    al, be = bounds
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 5, params=[1.0, 1.0, 1.0, 1.0], seed=3)
    A = _dense(ro, cols, vals)
    x = np.random.default_rng(1).standard_normal(len(ro) - 1)
    for k in range(1, 7):
        y = orc.cheb_apply(ro, cols, vals, al, be, k, x)
        scale = np.abs(y).max()
        assert np.abs(y - ps.cheb_apply_predicted(A, x, al, be, k)).max() <= 1e-12 * scale, k
        wrong_start = ps.cheb_apply_predicted(A, x, al, be, k, y0=np.zeros_like(x))
        assert np.abs(y - wrong_start).max() > 1e-2 * scale
        # d and c exchanged (the recurrence run with -alpha for alpha)
        wrong_dc = ps.cheb_apply_predicted(A, x, -al, be, k)
        assert np.abs(y - wrong_dc).max() > 1e-2 * scale


def test_cheb_algo2_is_not_the_textbook_chebyshev_polynomial():
    """Finding: _algo2's recurrence (first step 2/d, b = a (c/2)^2, :409,:424-425) is not the
    Chebyshev semi-iteration on [alpha, beta], whose residual polynomial is
    T_k((d - t)/c) / T_k(d/c) (first step 1/d). Already at order 1 they differ:
    1 - 2t/d against 1 - t/d."""
    import pin_systems as ps
    from numpy.polynomial import chebyshev as Ch
    al, be = 0.05, 1.9
    d, c = (be + al) / 2, (be - al) / 2
    t = np.linspace(al, be, 7)
    for k in (1, 2, 3):
        ref = np.polynomial.polynomial.polyval(t, ps.cheb_residual_poly(al, be, k))
        textbook = Ch.chebval((d - t) / c, [0] * k + [1]) / Ch.chebval(d / c, [0] * k + [1])
        assert np.abs(ref - textbook).max() > 0.1, k


# ---- Krylov exit and count rules against finite termination ------------------------------------
@pytest.mark.parametrize("m", [1, 2, 3, 4])
def test_krylov_counts_on_m_eigenvalue_systems(m):
    """D^-1 A with m distinct eigenvalues: PCG needs exactly m iterations; BiCGStab exits on
    the half-step vector s of its m-th step, which counts the iteration only in the first
    block (BiCGStab.h:194-199, m = 1) and not inside the loop (:259-263): m - 1 otherwise."""
    import pin_systems as ps
    ro, cols, vals = ps.m_eig_system(m)
    A = _dense(ro, cols, vals)
    dh = np.sqrt(np.diag(A))
    ev = np.linalg.eigvalsh(A / np.outer(dh, dh))
    np.testing.assert_allclose(np.unique(np.round(ev, 12)), ps.m_eig_spectrum(m), atol=1e-12)
    b = np.random.default_rng(7).uniform(0, 1, len(ro) - 1)
    xs = np.linalg.solve(A, b)
    for solver, want in ((orc.ORC_CG, ps.CG_ITERATIONS[m]), (orc.ORC_BICGS, ps.BICGSTAB_ITERATIONS[m])):
        x, r = orc.solve(ro, cols, vals, b, solver=solver, precond=orc.PREC_DIAG, tol=1e-8, max_iter=100)
        assert (r["iterations"], r["status"], r["error"]) == (want, 1, 0), (solver, r)
        assert np.abs(x - xs).max() <= 1e-13 * np.abs(xs).max()


def test_bicgstab_nullvalue_rules():
    """The NullValueException exits, on systems worked out by hand (pin_systems)."""
    import pin_systems as ps
    (ro, cols, vals), b = ps.bicgstab_tt_zero_system()
    for prec in (orc.PREC_NONE, orc.PREC_DIAG):
        x, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_BICGS, precond=prec, max_iter=100)
        assert (r["iterations"], r["error"]) == (0, 1)
        np.testing.assert_array_equal(x, [0.0, 0.0])
    (ro, cols, vals), b, x_after = ps.bicgstab_vr0_zero_system()
    x, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_BICGS, precond=orc.PREC_NONE, max_iter=100)
    assert (r["iterations"], r["error"]) == (1, 1)
    np.testing.assert_array_equal(x, x_after)


def test_max_iter_zero_still_takes_the_first_step():
    """CG.h:89-109 and BiCGStab.h:169-232 run their first block before any stop test, and
    Iteration::stop returns true at it >= max_it without a dot (Iteration.h:83-90): with
    max_iter = 0 CG takes one step, x = (r.z / z.A z) z with z = D^-1 b, and BiCGStab leaves
    at its first half step with ++iter and x = alpha phat, alpha = (b.b) / (A phat . b)."""
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 4, params=[1.0, 1.0, 1.0, 1.0], seed=2)
    A = _dense(ro, cols, vals)
    b = aa.uniform01(42, len(ro) - 1)
    z = b / np.diag(A)
    x, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_CG, precond=orc.PREC_DIAG, max_iter=0)
    assert (r["iterations"], r["status"]) == (1, 0)
    np.testing.assert_allclose(x, (b @ z) / (z @ (A @ z)) * z, rtol=1e-13)
    x, r = orc.solve(ro, cols, vals, b, solver=orc.ORC_BICGS, precond=orc.PREC_DIAG, max_iter=0)
    assert (r["iterations"], r["status"]) == (1, 0)
    np.testing.assert_allclose(x, (b @ b) / ((A @ z) @ b) * z, rtol=1e-13)
