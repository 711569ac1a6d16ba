"""Pins the FP64 CPU oracle of the learned preconditioner (oracle/gnn_oracle.py):
its analytic gradient against central finite differences, the SPD property of
M = G G^T, the initial factor G = D^-1/2 (I + E_lower), and the bf16 rounding
against numpy's own float32 round-to-nearest-even on exact cases; the k-step
CG residual loss against finite differences, an exact factor and plain CG."""
import numpy as np
import pytest
import scipy.sparse as sp

import alien_amd as aa
from oracle import gnn_oracle as go


def small_graph(seed=3, dims=(5, 4, 3)):
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, *dims, params=[1.5, 1.0, 4.0, 0.5], seed=seed)
    return go.Graph(ro, cols, vals)


def perturbed_theta(seed=0):
    th = go.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    return th + 0.05 * rng.standard_normal(th.shape)


def test_gradient_matches_finite_differences():
    g = small_graph()
    th = perturbed_theta()
    V = np.random.default_rng(7).choice([-1.0, 1.0], size=(g.n, 3))
    L, grad, _ = go.loss_and_grad(g, th, V)
    rng = np.random.default_rng(11)
    idx = rng.choice(go.NPARAM, size=60, replace=False)
    h = 1e-6
    for k in idx:
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd = (go.loss(g, go.forward(g, tp)[0], V)[0] - go.loss(g, go.forward(g, tm)[0], V)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, grad[k])


def test_initial_factor_and_spd():
    g = small_graph()
    G = go.preconditioner(g, go.init_params(0).astype(np.float64))
    Dm = sp.diags(1.0 / np.sqrt(g.d))
    expect = (Dm @ (sp.identity(g.n) + g.EL)).tocsr()
    np.testing.assert_allclose(G.toarray(), expect.toarray(), rtol=1e-12, atol=1e-14)
    th = perturbed_theta(1)
    G = go.preconditioner(g, th)
    M = (G @ G.T).toarray()
    np.testing.assert_allclose(M, M.T, rtol=1e-13, atol=1e-15)
    assert np.linalg.eigvalsh(M).min() > 0
    assert np.all(np.diag(G.toarray()) > 0)
    assert sp.triu(G, 1).nnz == 0


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 3.0e38])
    r = go.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    np.testing.assert_array_equal(r[:6], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0])
    assert np.isfinite(r[6])


def test_bf16_gradient_consistent_with_fp64():
    g = small_graph(5)
    th = perturbed_theta(2)
    V = np.random.default_rng(3).choice([-1.0, 1.0], size=(g.n, 2))
    L64, g64, _ = go.loss_and_grad(g, th, V, bf=False)
    L16, g16, _ = go.loss_and_grad(g, th, V, bf=True)
    assert abs(L16 - L64) <= 2e-2 * L64
    cos = g64 @ g16 / (np.linalg.norm(g64) * np.linalg.norm(g16))
    assert cos > 0.99


def test_cg_loss_gradient_matches_finite_differences():
    """The reverse-mode gradient of the k-step CG residual loss agrees with
    central differences on the parameters (FP64, no bf16 rounding)."""
    g = small_graph()
    rng = np.random.default_rng(3)
    th = perturbed_theta(1)
    B = rng.choice([-1.0, 1.0], size=(g.n, 4))
    L, grad, _ = go.cg_total_loss_and_grad(g, th, B, k=5, beta=0.1)
    idx = rng.choice(go.NPARAM, 40, replace=False)
    idx = np.union1d(idx, np.arange(go.NPARAM - 98, go.NPARAM, 7))  # heads
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (go.cg_total_loss_and_grad(g, tp, B, k=5, beta=0.1)[0] -
              go.cg_total_loss_and_grad(g, tm, B, k=5, beta=0.1)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_cg_loss_measures_convergence():
    """With M = A^-1 (exact factor) one CG step solves every system; with the
    identity the k-step loss equals plain CG's residual reduction."""
    g = small_graph()
    rng = np.random.default_rng(4)
    B = rng.standard_normal((g.n, 2))
    # exact: Ghat = chol(Ahat^-1) lower-triangular (dense, small system)
    Ai = np.linalg.inv(g.Ahat.toarray())
    C = np.linalg.cholesky(Ai)
    L1, _ = go.cg_loss(g, sp.csr_matrix(C), B, 1)
    assert L1 < -20
    # identity: compare with a plain CG loop
    Li, _ = go.cg_loss(g, sp.identity(g.n, format="csr"), B, 3)
    A = g.Ahat.toarray()
    ref = []
    for c in range(2):
        x = np.zeros(g.n); r = B[:, c].copy(); p = r.copy(); rr = r @ r
        for _ in range(3):
            q = A @ p; a = rr / (p @ q); x += a * p; r -= a * q; rn = r @ r; p = r + (rn / rr) * p; rr = rn
        ref.append(np.log(rr / (B[:, c] @ B[:, c])))
    assert abs(Li - np.mean(ref)) < 1e-10


def bsr2_graph(n=3, seed=5):
    """Scalarised 2-unknown block system (nonsymmetric values, symmetric pattern)."""
    ro, cols, vals = aa.gen_bsr2(n, params=[1.5, 0.3], seed=seed)
    nb = len(ro) - 1
    blocks = vals.reshape(-1, 2, 2).transpose(0, 2, 1)
    A = sp.bsr_matrix((blocks, cols, ro), shape=(2 * nb, 2 * nb)).tocsr()
    A.sort_indices()
    return go.Graph(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)


def test_total_loss_gradient_nonsymmetric_system():
    """Kaporin + probe objective and its gradient on a nonsymmetric matrix (the
    2-unknown block system): central differences through the whole network."""
    g = bsr2_graph()
    assert abs(g.Ahat - g.Ahat.T).max() > 1e-3
    th = perturbed_theta(2)
    V = np.random.default_rng(8).choice([-1.0, 1.0], size=(g.n, 4))
    L, grad, _ = go.total_loss_and_grad(g, th, V, beta=0.1)
    rng = np.random.default_rng(12)
    idx = np.union1d(rng.choice(go.NPARAM, 40, replace=False), np.arange(go.NPARAM - 98, go.NPARAM, 5))
    h = 1e-6
    for k in idx:
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd = (go.total_loss_and_grad(g, tp, V, beta=0.1)[0] - go.total_loss_and_grad(g, tm, V, beta=0.1)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grad[k])
