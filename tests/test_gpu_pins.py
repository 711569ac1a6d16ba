"""The systems that pin the oracle's Krylov rules (tests/pin_systems.py) through the
GPU solvers, in every mode that runs them: exact iteration counts fixed by finite
termination, the NullValueException exits, the max_iter = 0 first step, and the
Chebyshev preconditioner's power-method bounds and _algo2 apply against their
closed forms. These compare with the mathematics directly, not with the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

import alien_amd as aa
import pin_systems as ps

pytestmark = pytest.mark.gpu

ALL_MODES = [aa.MODE_REFERENCE, aa.MODE_REFERENCE_FUTURE, aa.MODE_FUSED, aa.MODE_GRAPH]
MODE_IDS = ["reference", "future", "fused", "graph"]


def say(*a):
    print("case", *a, flush=True)


def dense(ro, cols, vals):
    n = len(ro) - 1
    return sp.csr_matrix((vals, cols, ro), shape=(n, n)).toarray()


@pytest.mark.parametrize("mode", ALL_MODES, ids=MODE_IDS)
@pytest.mark.parametrize("m", [1, 2, 3, 4])
def test_counts_on_m_eigenvalue_systems(ctx, m, mode):
    ro, cols, vals = ps.m_eig_system(m)
    A = dense(ro, cols, vals)
    b = np.random.default_rng(7).uniform(0, 1, len(ro) - 1)
    xs = np.linalg.solve(A, b)
    for solver, want in ((aa.CG, ps.CG_ITERATIONS[m]), (aa.BICGS, ps.BICGSTAB_ITERATIONS[m])):
        say("m-eigenvalue counts", m, mode, solver)
        x, r = aa.solve_csr(ctx, ro, cols, vals, b, solver=solver, precond=aa.PREC_DIAG, mode=mode, tol=1e-8,
                            max_iter=100, poll_every=1)
        assert (r["iterations"], r["status"], r["error"]) == (want, 1, 0), (solver, r)
        assert np.abs(x - xs).max() <= 1e-13 * np.abs(xs).max()


@pytest.mark.parametrize("mode", ALL_MODES, ids=MODE_IDS)
def test_bicgstab_tt_zero_throws_in_first_block(ctx, mode):
    (ro, cols, vals), b = ps.bicgstab_tt_zero_system()
    say("bicgstab t.t = 0", mode)
    x, r = aa.solve_csr(ctx, ro, cols, vals, b, solver=aa.BICGS, precond=aa.PREC_DIAG, mode=mode, max_iter=100,
                        poll_every=1)
    assert (r["iterations"], r["error"]) == (0, 1), r
    np.testing.assert_array_equal(x, [0.0, 0.0])


@pytest.mark.parametrize("mode", [aa.MODE_REFERENCE, aa.MODE_REFERENCE_FUTURE], ids=MODE_IDS[:2])
def test_bicgstab_vr0_zero_throws_in_loop(ctx, mode):
    """No preconditioner (the zero diagonal entry has no inverse): the REFERENCE modes."""
    (ro, cols, vals), b, x_after = ps.bicgstab_vr0_zero_system()
    say("bicgstab v.r0 = 0", mode)
    x, r = aa.solve_csr(ctx, ro, cols, vals, b, solver=aa.BICGS, precond=aa.PREC_NONE, mode=mode, max_iter=100)
    assert (r["iterations"], r["error"]) == (1, 1), r
    np.testing.assert_array_equal(x, x_after)


@pytest.mark.parametrize("mode", ALL_MODES, ids=MODE_IDS)
def test_max_iter_zero_takes_the_first_step(ctx, mode):
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 4, params=[1.0, 1.0, 1.0, 1.0], seed=2)
    A = dense(ro, cols, vals)
    b = aa.uniform01(42, len(ro) - 1)
    z = b / np.diag(A)
    say("max_iter 0", mode)
    x, r = aa.solve_csr(ctx, ro, cols, vals, b, solver=aa.CG, precond=aa.PREC_DIAG, mode=mode, max_iter=0,
                        poll_every=1)
    assert (r["iterations"], r["status"], r["error"]) == (1, 0, 0), r
    np.testing.assert_allclose(x, (b @ z) / (z @ (A @ z)) * z, rtol=1e-13)
    x, r = aa.solve_csr(ctx, ro, cols, vals, b, solver=aa.BICGS, precond=aa.PREC_DIAG, mode=mode, max_iter=0,
                        poll_every=1)
    assert (r["iterations"], r["status"], r["error"]) == (1, 0, 0), r
    np.testing.assert_allclose(x, (b @ b) / ((A @ z) @ b) * z, rtol=1e-13)


@pytest.mark.parametrize("mode", [aa.MODE_REFERENCE, aa.MODE_FUSED], ids=["reference", "fused"])
def test_chebyshev_power_method_closed_form(ctx, mode):
    """poly_factor = 0 keeps computeAlpha's estimate: the bounds converge to
    1 -+ cos(pi/(n+1)), the ends of the spectrum of D^-1 A on the n^3 Laplacian."""
    n = 5
    ro, cols, vals = aa.gen_csr(aa.GEN_POISSON7, n)
    b = aa.uniform01(42, n ** 3)
    say("chebyshev power method", mode)
    _, r = aa.solve_csr(ctx, ro, cols, vals, b, solver=aa.CG, precond=aa.PREC_CHEB, mode=mode, poly_factor=0.0,
                        poly_max_iter=3000, max_iter=0)
    assert r["cheb_beta"] == pytest.approx(1 + np.cos(np.pi / (n + 1)), abs=1e-12)
    assert r["cheb_alpha"] == pytest.approx(1 - np.cos(np.pi / (n + 1)), abs=1e-12)


@pytest.mark.parametrize("mode", ALL_MODES, ids=MODE_IDS)
@pytest.mark.parametrize("order", [1, 2, 3, 5])
def test_chebyshev_apply_is_its_residual_polynomial(ctx, mode, order):
    """One CG step (max_iter = 0) from x = 0 gives x = (b.z / z.A z) z with z = M^-1 b, the
    _algo2 apply of b; z is predicted from the residual polynomial of the recurrence on the
    eigen-decomposition of D^-1/2 A D^-1/2, with the bounds the solve reports."""
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 5, params=[1.0, 1.0, 1.0, 1.0], seed=3)
    A = dense(ro, cols, vals)
    b = aa.uniform01(42, len(ro) - 1)
    say("chebyshev apply", mode, order)
    x, r = aa.solve_csr(ctx, ro, cols, vals, b, solver=aa.CG, precond=aa.PREC_CHEB, mode=mode, poly_factor=0.5,
                        poly_order=order, poly_max_iter=10, max_iter=0, poll_every=1)
    assert (r["iterations"], r["error"]) == (1, 0), r
    z = ps.cheb_apply_predicted(A, b, r["cheb_alpha"], r["cheb_beta"], order)
    pred = (b @ z) / (z @ (A @ z)) * z
    assert np.abs(x - pred).max() <= 1e-12 * np.abs(pred).max()
