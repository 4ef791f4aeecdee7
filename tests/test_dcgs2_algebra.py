"""The delayed-CGS2 Arnoldi recurrences of the device FGMRES, replayed in numpy.

The device solver (markovkit_amd/csrc/api.hip, fgmres) applies the preconditioner and the
operator to the once-orthogonalised candidate u_j scaled by 1 / beta_j, then in one reduction
gets Q^T w and u_j^T w, finalises q_j = (u_j - Q s_j) / beta_j, forms eta = q_j^T w from
(u_j^T w - s_j^T Q^T w) / beta_j and the next candidate u_{j+1} = w - Q Q^T w - eta q_j, and takes
column j of H as (Q^T w + s_{j+1}, eta + s_{j+1}[j], beta_{j+1}) with the Pythagorean
beta_{j+1}^2 = |u_{j+1}|^2 - |s_{j+1}|^2 (or an explicit extra projection when that cancels).
This replay pins the algebra against what must hold for any Arnoldi process: the flexible
Arnoldi relation A Z = Q_{m+1} H, orthonormal Q, and the FGMRES iterate equal to the
least-squares minimiser of |b - A Z y| computed directly.
"""
import numpy as np
import pytest


def dcgs2_arnoldi(A, M, r, m, explicit_threshold=0.25, force_explicit=False):
    n = r.size
    V = np.zeros((m + 1, n))
    Z = np.zeros((m, n))
    H = np.zeros((m + 1, m))
    V[0] = r
    bj = np.linalg.norm(r)
    s = np.zeros(0)
    explicit = 0
    for j in range(m):
        Z[j] = M @ (V[j] * (1.0 / bj))
        w = A @ Z[j]
        tg = V[: j + 1] @ w
        t, gamma = tg[:j], tg[j]
        eta = (gamma - s @ t) / bj
        q = (V[j] - V[:j].T @ s) * (1.0 / bj)
        V[j] = q
        u = (w - V[:j].T @ t) - eta * q
        V[j + 1] = u
        sn = V[: j + 1] @ u
        alpha = u @ u
        col = np.r_[t, eta] + sn
        ss = sn @ sn
        if force_explicit or not ss <= explicit_threshold * alpha:
            explicit += 1
            u = u - V[: j + 1].T @ sn
            V[j + 1] = u
            sn = V[: j + 1] @ u
            alpha = u @ u
            col = col + sn
            ss = sn @ sn
        hn = np.sqrt(max(alpha - ss, 0.0))
        H[: j + 1, j] = col
        H[j + 1, j] = hn
        s, bj = sn, hn
        if hn == 0.0:
            return V, Z, H[: j + 2, : j + 1], s, bj, explicit, j + 1
    return V, Z, H, s, bj, explicit, m


def final_basis(V, H, s, bj, k):
    """Q_{k+1}: the k finalised columns plus the last candidate finalised with (s, beta)."""
    Q = V[: k + 1].copy()
    Q[k] = (V[k] - V[:k].T @ s) / bj
    return Q


def problem(n, seed, spectrum=None):
    rng = np.random.default_rng(seed)
    if spectrum is None:
        A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) / np.sqrt(n)
    else:
        X = np.linalg.qr(rng.standard_normal((n, n)))[0]
        A = X @ np.diag(rng.choice(spectrum, n)) @ X.T
    # an inexact, nonsymmetric preconditioner
    M = np.linalg.inv(A + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n))
    b = rng.random(n)
    return A, M, b


@pytest.mark.parametrize("force", [False, True], ids=["pythagorean", "explicit"])
def test_flexible_arnoldi_relation(force):
    A, M, b = problem(120, 1)
    m = 30
    V, Z, H, s, bj, explicit, k = dcgs2_arnoldi(A, M, b, m, force_explicit=force)
    assert k == m and (explicit == m if force else explicit == 0)
    Q = final_basis(V, H, s, bj, k)
    # orthonormal basis
    assert np.abs(Q @ Q.T - np.eye(k + 1)).max() <= 1e-13
    # A Z = Q_{m+1} H, relative to |A Z|
    AZ = (A @ Z.T)
    assert np.abs(AZ - Q.T @ H).max() <= 1e-13 * np.abs(AZ).max()
    # the first basis vector is b / |b|
    assert np.abs(Q[0] - b / np.linalg.norm(b)).max() <= 1e-15 * 4


def test_fgmres_iterate_is_the_least_squares_minimiser():
    A, M, b = problem(150, 2)
    m = 12
    V, Z, H, s, bj, _, k = dcgs2_arnoldi(A, M, b, m)
    beta = np.linalg.norm(b)
    g = np.zeros(k + 1)
    g[0] = beta
    y = np.linalg.lstsq(H, g, rcond=None)[0]
    x = Z.T @ y
    y_direct = np.linalg.lstsq((A @ Z.T), b, rcond=None)[0]
    x_direct = Z.T @ y_direct
    assert np.linalg.norm(x - x_direct) <= 1e-11 * np.linalg.norm(x_direct)
    # the residual estimate |g - H y| equals the true residual
    est = np.linalg.norm(g - H @ y)
    assert abs(est - np.linalg.norm(b - A @ x)) <= 1e-11 * beta


def test_invariant_subspace_gives_the_solution():
    """With three distinct eigenvalues and M = I the Krylov space closes after three steps: the
    fourth candidate is rounding noise (its subdiagonal entry is at rounding level) and the
    iterate after three steps solves the system."""
    n = 60
    A, _, b = problem(n, 3, spectrum=(1.0, 2.0, 5.0))
    V, Z, H, s, bj, explicit, k = dcgs2_arnoldi(A, np.eye(n), b, 3)
    assert H[3, 2] <= 1e-13 * np.linalg.norm(b)
    g = np.zeros(4)
    g[0] = np.linalg.norm(b)
    y = np.linalg.lstsq(H, g, rcond=None)[0]
    x = Z.T @ y
    assert np.linalg.norm(b - A @ x) <= 1e-12 * np.linalg.norm(b)
