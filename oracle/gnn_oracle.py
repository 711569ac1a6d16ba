"""TEST INFRASTRUCTURE ONLY: FP64 CPU oracle of the learned preconditioner.

The network has no counterpart in the reference; this is the CPU statement of
its mathematics that the HIP forward/backward kernels are checked against
(tests/), itself pinned by finite differences (tests/test_gnn_oracle.py).

Model (parameters in one flat FP32 vector, layout `LAYOUT` below):
  d_i = a_ii,  e_ij = -a_ij / sqrt(d_i d_j)  (j != i, on A's pattern)
  f_i = [1, s_i, 1 - s_i, max_j e_ij, u_i, l_i],  s_i = sum_j e_ij,
        u_i = sum_{j > i} e_ij^2,  l_i = sum_{j < i} e_ij^2
  h0  = relu(f W_in + b_in)                                   (N x 32)
  h_l = h_{l-1} + relu([h_{l-1}, S h_{l-1}] W_l + b_l), l=1..3, S_ij = e_ij
  Ghat_ii = exp(h3_i . w_d + c_d)
  Ghat_ij = e_ij (h3_i . w_a + h3_j . w_b + c_e)   for j < i on A's pattern
  G = D^-1/2 Ghat,  M = G G^T  (symmetric positive definite: Ghat_ii > 0)
Loss on K probe vectors V (N x K):
  L = || V - Ghat Ghat^T Ahat V ||_F^2 / (N K),  Ahat = D^-1/2 A D^-1/2
With bf16=True the weights and the inputs of every dense layer are rounded to
bfloat16 (round to nearest even) before use, as on the GPU's MFMA path.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

H = 32
F = 6
LAYOUT = [("W_in", (F, H)), ("b_in", (H,)),
          ("W1", (2 * H, H)), ("b1", (H,)),
          ("W2", (2 * H, H)), ("b2", (H,)),
          ("W3", (2 * H, H)), ("b3", (H,)),
          ("w_d", (H,)), ("c_d", (1,)), ("w_a", (H,)), ("w_b", (H,)), ("c_e", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(theta):
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in LAYOUT])


def bf16(x):
    """Round float64 values to bfloat16 (via float32, round to nearest even)."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0):
    """Deterministic initialisation: small random layers, output heads at the
    first-order factor G = D^-1/2 (I + E_lower) (Ghat_ii = 1, Ghat_ij = e_ij)."""
    rng = np.random.default_rng(seed)
    p = {}
    p["W_in"] = rng.standard_normal((F, H)) * (1.0 / np.sqrt(F))
    p["b_in"] = np.zeros(H)
    for l in (1, 2, 3):
        p[f"W{l}"] = rng.standard_normal((2 * H, H)) * (0.1 / np.sqrt(2 * H))
        p[f"b{l}"] = np.zeros(H)
    p["w_d"] = np.zeros(H)
    p["c_d"] = np.zeros(1)
    p["w_a"] = np.zeros(H)
    p["w_b"] = np.zeros(H)
    p["c_e"] = np.ones(1)
    return pack(p).astype(np.float32)


class Graph:
    """Structure and scaled coefficients of an SPD matrix with a symmetric pattern."""

    def __init__(self, ro, cols, vals):
        n = len(ro) - 1
        self.n = n
        A = sp.csr_matrix((vals, cols, ro), shape=(n, n))
        d = A.diagonal()
        self.d = d
        Dm = sp.diags(1.0 / np.sqrt(d))
        Ah = (Dm @ A @ Dm).tocsr()
        Ah.sort_indices()
        self.Ahat = Ah
        E = -(Ah - sp.diags(Ah.diagonal()))
        E.eliminate_zeros()
        self.E = E.tocsr()
        self.EL = sp.tril(self.E, -1).tocsr()
        s = np.asarray(self.E.sum(axis=1)).ravel()
        mx = np.zeros(n)
        Ec = self.E.tocsr()
        for i in range(n):
            row = Ec.data[Ec.indptr[i]:Ec.indptr[i + 1]]
            mx[i] = row.max() if len(row) else 0.0
        EU = sp.triu(self.E, 1).tocsr()
        ELs = sp.tril(self.E, -1).tocsr()
        u = np.asarray(EU.multiply(EU).sum(axis=1)).ravel()
        lo = np.asarray(ELs.multiply(ELs).sum(axis=1)).ravel()
        self.f = np.stack([np.ones(n), s, 1.0 - s, mx, u, lo], axis=1)
        L = self.EL.tocoo()
        self.li, self.lj, self.le = L.row, L.col, L.data  # strictly lower edges (i > j)


def forward(g: Graph, theta, bf=False):
    p = unpack(theta)
    rw = (lambda w: bf16(w)) if bf else (lambda w: w)
    rx = rw
    cache = {}
    z0 = rx(g.f) @ rw(p["W_in"]) + p["b_in"]
    h = np.maximum(z0, 0.0)
    cache["z0"] = z0
    hs = [h]
    us, zs = [], []
    for l in (1, 2, 3):
        m = g.E @ h
        u = np.concatenate([h, m], axis=1)
        z = rx(u) @ rw(p[f"W{l}"]) + p[f"b{l}"]
        h = h + np.maximum(z, 0.0)
        us.append(u)
        zs.append(z)
        hs.append(h)
    h3 = hs[-1]
    od = h3 @ p["w_d"] + p["c_d"][0]
    gd = np.exp(od)
    c = h3[g.li] @ p["w_a"] + h3[g.lj] @ p["w_b"] + p["c_e"][0]
    gl = g.le * c
    Ghat = (sp.csr_matrix((gl, (g.li, g.lj)), shape=(g.n, g.n)) + sp.diags(gd)).tocsr()
    cache.update(hs=hs, us=us, zs=zs, gd=gd, c=c, p=p)
    return Ghat, cache


def loss(g: Graph, Ghat, V):
    W = g.Ahat @ V
    U = Ghat.T @ W
    Y = Ghat @ U
    R = V - Y
    NK = V.size
    return float(np.sum(R * R) / NK), (W, U, R)


def grad_ghat(g: Graph, Ghat, V, aux):
    """dL/dGhat on the lower pattern: diagonal values and strictly-lower edge values."""
    W, U, R = aux
    dY = -2.0 * R / V.size
    Q = Ghat.T @ dY
    # grad_ij = sum_c dY_ic U_jc + W_ic Q_jc
    gd = np.sum(dY * U, axis=1) + np.sum(W * Q, axis=1)
    ge = np.sum(dY[g.li] * U[g.lj], axis=1) + np.sum(W[g.li] * Q[g.lj], axis=1)
    return gd, ge


def backward(g: Graph, theta, cache, dgd, dge, bf=False):
    """Gradient of the loss w.r.t. the flat parameter vector, given dL/dGhat."""
    p = cache["p"]
    rw = (lambda w: bf16(w)) if bf else (lambda w: w)
    hs, us, zs = cache["hs"], cache["us"], cache["zs"]
    h3 = hs[-1]
    grads = {k: np.zeros_like(v) for k, v in p.items()}
    do = dgd * cache["gd"]                     # d exp
    dc = dge * g.le                            # Ghat_ij = e_ij c_ij
    grads["w_d"] = h3.T @ do
    grads["c_d"] = np.array([do.sum()])
    grads["w_a"] = h3[g.li].T @ dc
    grads["w_b"] = h3[g.lj].T @ dc
    grads["c_e"] = np.array([dc.sum()])
    dh = np.outer(do, p["w_d"])
    np.add.at(dh, g.li, np.outer(dc, p["w_a"]))
    np.add.at(dh, g.lj, np.outer(dc, p["w_b"]))
    for l in (3, 2, 1):
        u, z = us[l - 1], zs[l - 1]
        dz = dh * (z > 0)
        ub = bf16(u) if bf else u
        grads[f"W{l}"] = ub.T @ dz
        grads[f"b{l}"] = dz.sum(axis=0)
        du = dz @ rw(p[f"W{l}"]).T
        dh = dh + du[:, :H] + g.E.T @ du[:, H:]
    dz0 = dh * (cache["z0"] > 0)
    fb = bf16(g.f) if bf else g.f
    grads["W_in"] = fb.T @ dz0
    grads["b_in"] = dz0.sum(axis=0)
    return pack(grads)


def loss_and_grad(g: Graph, theta, V, bf=False):
    Ghat, cache = forward(g, theta, bf)
    L, aux = loss(g, Ghat, V)
    dgd, dge = grad_ghat(g, Ghat, V, aux)
    return L, backward(g, theta, cache, dgd, dge, bf), Ghat


def preconditioner(g: Graph, theta, bf=False):
    """G = D^-1/2 Ghat (lower triangular); M = G G^T."""
    Ghat, _ = forward(g, theta, bf)
    return (sp.diags(1.0 / np.sqrt(g.d)) @ Ghat).tocsr()


def kcond_loss(g: Graph, Ghat):
    """Kaporin's condition number of B = Ghat^T Ahat Ghat, per row and up to the
    constant log det(Ahat)/N:  L = log(tr(B)/N) - (2/N) sum_j log Ghat_jj.
    It bounds the CG iteration count (Kaporin 1994) and its minimiser over the
    pattern is the factorised sparse approximate inverse (FSAI). For a
    nonsymmetric Ahat the trace sees its symmetric part only."""
    AG = (g.Ahat @ Ghat).tocsr()
    tr = float(Ghat.multiply(AG).sum())
    gd = Ghat.diagonal()
    L = np.log(tr / g.n) - 2.0 * np.sum(np.log(gd)) / g.n
    AGs = (0.5 * ((g.Ahat + g.Ahat.T) @ Ghat)).tocsr()
    return L, (AGs, tr)


def grad_kcond(g: Graph, Ghat, aux):
    """dL/dGhat of kcond_loss on the lower pattern (diagonal, strictly-lower
    edges): d tr(G^T A G)/dG = (A + A^T) G."""
    AGs, tr = aux
    AGd = np.asarray(AGs[np.arange(g.n), np.arange(g.n)]).ravel()
    gd = 2.0 * AGd / tr - 2.0 / (g.n * Ghat.diagonal())
    ge = 2.0 * np.asarray(AGs[g.li, g.lj]).ravel() / tr
    return gd, ge


def total_loss_and_grad(g: Graph, theta, V, beta=0.1, bf=False):
    """Training objective: Kaporin term + beta * probe term, and its gradient."""
    Ghat, cache = forward(g, theta, bf)
    LK, auxK = kcond_loss(g, Ghat)
    LP, auxP = loss(g, Ghat, V)
    dK = grad_kcond(g, Ghat, auxK)
    dP = grad_ghat(g, Ghat, V, auxP)
    dgd = dK[0] + beta * dP[0]
    dge = dK[1] + beta * dP[1]
    return LK + beta * LP, backward(g, theta, cache, dgd, dge, bf), Ghat


def cg_loss(g: Graph, Ghat, B, k):
    """Residual reduction after k preconditioned CG steps: CG on Ahat with
    M = Ghat Ghat^T, x0 = 0, one right-hand side per column of B;
    L = mean_c log(|r_k,c|^2 / |r_0,c|^2). rho = (Ghat^T r).(Ghat^T r) = r.M r."""
    A, G = g.Ahat, Ghat.tocsr()
    GT = G.T.tocsr()
    K = B.shape[1]
    r = [B.copy()]
    t = [GT @ B]
    z = G @ t[0]
    p = [z]
    rho = [np.sum(t[0] * t[0], axis=0)]
    q, sig, alpha, beta = [], [], [], []
    for j in range(k):
        qj = A @ p[j]
        q.append(qj)
        sig.append(np.sum(p[j] * qj, axis=0))
        alpha.append(rho[j] / sig[j])
        r.append(r[j] - alpha[j] * qj)
        if j < k - 1:
            tj = GT @ r[j + 1]
            t.append(tj)
            zj = G @ tj
            rho.append(np.sum(tj * tj, axis=0))
            beta.append(rho[j + 1] / rho[j])
            p.append(zj + beta[j] * p[j])
    rr0 = np.sum(B * B, axis=0)
    rrk = np.sum(r[k] * r[k], axis=0)
    L = float(np.mean(np.log(rrk / rr0)))
    return L, dict(r=r, t=t, p=p, q=q, rho=rho, sig=sig, alpha=alpha, beta=beta, rrk=rrk, K=K, k=k)


def grad_cg_loss(g: Graph, Ghat, aux):
    """dL/dGhat of cg_loss on the lower pattern (diagonal, strictly-lower edges),
    by reverse-mode differentiation of the k CG steps."""
    A, G = g.Ahat, Ghat.tocsr()
    GT = G.T.tocsr()
    r, t, p, q = aux["r"], aux["t"], aux["p"], aux["q"]
    rho, sig, alpha, beta = aux["rho"], aux["sig"], aux["alpha"], aux["beta"]
    K, k = aux["K"], aux["k"]
    n = g.n
    gd = np.zeros(n)
    ge = np.zeros(len(g.li))

    def acc(U, V):  # dGhat_il += sum_c U_ic V_lc on the lower pattern
        gd[:] += np.sum(U * V, axis=1)
        ge[:] += np.sum(U[g.li] * V[g.lj], axis=1)

    rbar = 2.0 * r[k] / (K * aux["rrk"])
    pbar_next = None       # dL/dp_{j+1}
    rhobar_next = None     # dL/drho_{j+1} accumulated so far
    for j in range(k - 1, -1, -1):
        rhobar_j = np.zeros(K)
        if j < k - 1:
            zbar = pbar_next
            betabar = np.sum(pbar_next * p[j], axis=0)
            pbar = beta[j] * pbar_next
            rhobar_jp1 = rhobar_next + betabar / rho[j]
            rhobar_j += -betabar * rho[j + 1] / rho[j] ** 2
            tbar = 2.0 * rhobar_jp1 * t[j + 1]
            acc(zbar, t[j + 1])
            tbar = tbar + GT @ zbar
            acc(r[j + 1], tbar)
            rbar = rbar + G @ tbar
        else:
            pbar = np.zeros_like(p[j])
        alphabar = -np.sum(q[j] * rbar, axis=0)
        qbar = -alpha[j] * rbar
        rhobar_j += alphabar / sig[j]
        sigbar = -alphabar * rho[j] / sig[j] ** 2
        pbar = pbar + sigbar * q[j]
        qbar = qbar + sigbar * p[j]
        pbar = pbar + A @ qbar
        pbar_next, rhobar_next = pbar, rhobar_j
    # p_0 = z_0 = Ghat t_0, rho_0 = t_0.t_0, t_0 = Ghat^T r_0 (r_0 = B fixed)
    zbar = pbar_next
    tbar = 2.0 * rhobar_next * t[0]
    acc(zbar, t[0])
    tbar = tbar + GT @ zbar
    acc(r[0], tbar)
    return gd, ge


def cg_total_loss_and_grad(g: Graph, theta, B, k=8, beta=0.1, bf=False):
    """Training objective with the CG loss: cg_loss(k) + beta * probe term, and its gradient."""
    Ghat, cache = forward(g, theta, bf)
    LC, auxC = cg_loss(g, Ghat, B, k)
    LP, auxP = loss(g, Ghat, B)
    dC = grad_cg_loss(g, Ghat, auxC)
    dP = grad_ghat(g, Ghat, B, auxP)
    dgd = dC[0] + beta * dP[0]
    dge = dC[1] + beta * dP[1]
    return LC + beta * LP, backward(g, theta, cache, dgd, dge, bf), Ghat
