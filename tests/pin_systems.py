"""Small systems whose Krylov behaviour the mathematics fixes, shared by the
oracle pins (test_oracle.py) and the GPU runs of the same systems
(test_gpu_pins.py).

* m_eig_system(m): SPD, block diagonal, with exactly m distinct eigenvalues of
  D^-1 A (so of the Jacobi-preconditioned operators A D^-1 and D^-1 A). In
  exact arithmetic PCG then terminates after m iterations, and BiCG's residual
  vanishes at its m-th step, which BiCGStab computes as the half-step vector s
  of its m-th step (s_k = Q_{k-1}(A M) r_k^BiCG).
* The NullValueException systems of BiCGStab, worked out by hand in exact
  binary arithmetic.
* cheb_residual_poly: the residual polynomial R_k of the three-term recurrence
  of ChebyshevPreconditioner::_algo2 (ChebyshevPreconditioner.h:392-434):
  for y_k the k-th iterate, A^-1 x - y_k = R_k(D^-1 A) (A^-1 x - y_0).
"""
import numpy as np
import scipy.sparse as sp
from numpy.polynomial import polynomial as P


def _csr(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def m_eig_system(m, nblocks=12, seed=0):
    """Block-diagonal SPD matrix, block k = a_k * (I + rho T) with a_k ~ U[1, 4):
    D^-1 A is the same for every block of one kind, so its spectrum is
    m=1: {1}; m=2: {1 +- 0.5}; m=3: {1, 1 +- 0.4 sqrt 2} (3x3 path);
    m=4: {1 +- 0.3, 1 +- 0.6} (two kinds of 2x2 blocks)."""
    rng = np.random.default_rng(seed)
    blocks = []
    for k in range(nblocks):
        a = rng.uniform(1.0, 4.0)
        if m == 1:
            blocks.append(np.array([[a]]))
        elif m == 2:
            blocks.append(a * np.array([[1.0, 0.5], [0.5, 1.0]]))
        elif m == 3:
            blocks.append(a * np.array([[1.0, 0.4, 0.0], [0.4, 1.0, 0.4], [0.0, 0.4, 1.0]]))
        elif m == 4:
            r = 0.3 if k % 2 else 0.6
            blocks.append(a * np.array([[1.0, r], [r, 1.0]]))
        else:
            raise ValueError(m)
    return _csr(sp.block_diag(blocks))


def m_eig_spectrum(m):
    return {1: [1.0], 2: [0.5, 1.5], 3: [1 - 0.4 * 2 ** 0.5, 1.0, 1 + 0.4 * 2 ** 0.5],
            4: [0.4, 0.7, 1.3, 1.6]}[m]


# PCG: m iterations. BiCGStab: the first block's half-step exit (counted) when m = 1,
# else the m-th step's half-step exit inside the loop (not counted): m - 1.
CG_ITERATIONS = {1: 1, 2: 2, 3: 3, 4: 4}
BICGSTAB_ITERATIONS = {1: 1, 2: 1, 3: 2, 4: 3}


def bicgstab_tt_zero_system():
    """A = [[1, 0.5], [2, 1]] (unit diagonal, so Jacobi is the identity), b = (2, 1).
    A = 2 u w^T with u = (1, 2), w = (1/2, 1/4), w.u = 1, so A^2 = 2A. First block:
    v = A b = (2.5, 5), alpha = b.b / v.b = 5/10 = 1/2, s = b - v/2 = (0.75, -1.5),
    t = A s = (0, 0): t.t = 0 while r = b is far from converged, so
    BiCGStab.h:210-218 throws NullValueException("beta") before any iteration."""
    return _csr(np.array([[1.0, 0.5], [2.0, 1.0]])), np.array([2.0, 1.0])


def bicgstab_vr0_zero_system():
    """A = diag(1, 0) (the zero diagonal entry stored), b = (1, 1), no preconditioner.
    First block: v = (1, 0), alpha = 2, s = (-1, 1), t = (-1, 0), omega = 1,
    x = omega s + alpha b = (1, 3), r = s - t = (0, 1), one iteration. Loop:
    rho1 = 1, beta = (1/2)(2/1) = 1, p = (p - omega v) beta + r = (0, 2), v = A p = 0:
    v.r0 = 0 throws NullValueException("alpha") (BiCGStab.h:249-250) after 1 iteration."""
    return _csr(sp.csr_matrix((np.array([1.0, 0.0]), (np.array([0, 1]), np.array([0, 1]))), shape=(2, 2))), \
        np.array([1.0, 1.0]), np.array([1.0, 3.0])


def cheb_residual_poly(alpha, beta, order):
    """Coefficients (low to high) of R_k for _algo2 with bounds alpha, beta:
    d = (beta + alpha)/2, c = (beta - alpha)/2; step 0: p_0 = z_0, y_1 = y_0 + (2/d) p_0;
    step i: b = a (c/2)^2, a = 1/(d - b), p_i = z_i + b p_{i-1}, y_{i+1} = y_i + a p_i,
    with z_i = D^-1 (x - A y_i) = K e_i, K = D^-1 A, e_i = A^-1 x - y_i."""
    d = (beta + alpha) / 2
    c = (beta - alpha) / 2
    a = 2.0 / d
    S = np.array([0.0, 1.0])               # p_0 = K e_0
    R = P.polysub([1.0], a * S)            # e_1 = e_0 - a p_0
    for _ in range(1, order):
        Z = P.polymul([0.0, 1.0], R)       # z_i = K e_i
        b = a * (c / 2.0) ** 2
        a = 1.0 / (d - b)
        S = P.polyadd(Z, b * S)
        R = P.polysub(R, a * S)
    return R


def apply_poly_of_jacobi_operator(A, R, v):
    """R(D^-1 A) v for a dense SPD A, through S = D^-1/2 A D^-1/2 = V diag(lam) V^T:
    R(D^-1 A) = D^-1/2 R(S) D^1/2."""
    dh = np.sqrt(np.diag(A))
    lam, V = np.linalg.eigh(A / np.outer(dh, dh))
    return (V @ (P.polyval(lam, R) * (V.T @ (dh * v)))) / dh


def cheb_apply_predicted(A, x, alpha, beta, order, y0=None):
    """y_k of _algo2 from its residual polynomial: y_k = A^-1 x - R_k(K) (A^-1 x - y0), y0 = x."""
    y0 = x if y0 is None else y0
    ainv_x = np.linalg.solve(A, x)
    return ainv_x - apply_poly_of_jacobi_operator(A, cheb_residual_poly(alpha, beta, order), ainv_x - y0)
