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


def dcgs2_nonflexible(B, v, m, force_explicit=False, leak=0.0, seed=0):
    """The batched exponential's Arnoldi (markovkit_amd/csrc/expv.hip, expv_engine_step): delayed
    CGS2 with M = I, q_0 = v / |v| (beta_0 = 1, s_0 empty), B applied to u_j / beta_j, and per step
    the recorded (t, eta, first and second projections, flag, beta_{j+1}) from which the host
    assembles the Hessenberg matrix. leak > 0 perturbs the first projections t and eta by a
    relative amount leak (a poor first pass: the candidate then keeps a part in span(Q)), which
    the delayed second projection removes and the Hessenberg correction must account for."""
    rng = np.random.default_rng(seed)
    n = v.size
    V = np.zeros((m + 1, n))
    V[0] = v / np.linalg.norm(v)
    s, bj = np.zeros(0), 1.0
    rec = []
    for j in range(m):
        w = B @ (V[j] * (1.0 / bj))
        tg = V[: j + 1] @ w
        t, gamma = tg[:j], tg[j]
        eta = (gamma - s @ t) / bj
        q = (V[j] - V[:j].T @ s) * (1.0 / bj)
        V[j] = q
        if leak > 0.0:
            d = leak * np.linalg.norm(np.r_[t, eta]) * rng.standard_normal(j + 1) / np.sqrt(j + 1)
            t, eta = t + d[:j], eta + d[j]
        u = (w - V[:j].T @ t) - eta * q
        V[j + 1] = u
        s1 = np.r_[V[: j + 1] @ u, u @ u]
        flag = force_explicit or not (s1[:-1] @ s1[:-1] <= 0.25 * s1[-1])
        s2 = np.zeros_like(s1)
        if flag:
            u = u - V[: j + 1].T @ s1[:-1]
            V[j + 1] = u
            s2 = np.r_[V[: j + 1] @ u, u @ u]
        sn = s2 if flag else s1
        bn = np.sqrt(max(sn[-1] - sn[:-1] @ sn[:-1], 0.0))
        rec.append(dict(t=t, eta=eta, s1=s1, s2=s2, flag=flag, beta=bn, s_used=sn[:-1]))
        s, bj = sn[:-1], bn
    V[m] = (V[m] - V[:m].T @ s) / bj
    return V, rec


def host_hessenberg(rec, m):
    """The host's assembly in expv_engine_step: column k = (t + s [+ s1 where flagged], eta + ...,
    beta_{k+1}), minus sum_i (s_k[i] / beta_k) h_i for the basis vectors u_k / beta_k carries."""
    H = np.zeros((m + 1, m))
    for k in range(m):
        r = rec[k]
        sn = r["s2"] if r["flag"] else r["s1"]
        c = np.zeros(m + 1)
        c[:k] = r["t"] + sn[:k]
        c[k] = r["eta"] + sn[k]
        if r["flag"]:
            c[: k + 1] += r["s1"][: k + 1]
        c[k + 1] = r["beta"]
        if k > 0:
            bk = rec[k - 1]["beta"]
            s_k = rec[k - 1]["s_used"]
            for q in range(k):
                c[: q + 2] -= (s_k[q] / bk) * H[: q + 2, q]
        H[:, k] = c
    return H


def mgs_arnoldi(B, v, m):
    n = v.size
    Q = np.zeros((m + 1, n))
    H = np.zeros((m + 1, m))
    Q[0] = v / np.linalg.norm(v)
    for j in range(m):
        w = B @ Q[j]
        for i in range(j + 1):
            H[i, j] = Q[i] @ w
            w = w - H[i, j] * Q[i]
        H[j + 1, j] = np.linalg.norm(w)
        Q[j + 1] = w / H[j + 1, j]
    return Q, H


@pytest.mark.parametrize("force,leak", [(False, 0.0), (True, 0.0), (False, 0.05), (True, 0.05)],
                         ids=["pythagorean", "explicit", "poor-first-pass", "poor-first-pass-explicit"])
def test_nonflexible_lagged_normalisation(force, leak):
    """The corrected columns satisfy B Q_m = Q_{m+1} H for the finalised (orthonormal) basis, their
    first columns equal a modified Gram-Schmidt Arnoldi's from the same vector, and the Krylov
    exponential from them equals expm; with candidates that keep a part in span(Q) (the case the
    correction is for), leaving the correction out breaks the relation."""
    from scipy.linalg import expm
    rng = np.random.default_rng(7)
    n, m = 140, 30
    B = np.eye(n) * 3.0 + rng.standard_normal((n, n)) / np.sqrt(n)
    v = rng.random(n)
    V, rec = dcgs2_nonflexible(B, v, m, force_explicit=force, leak=leak)
    assert sum(r["flag"] for r in rec) == (m if force else 0)
    H = host_hessenberg(rec, m)
    assert np.abs(V @ V.T - np.eye(m + 1)).max() <= 1e-13
    BQ = B @ V[:m].T
    assert np.abs(BQ - V.T @ H).max() <= 1e-13 * np.abs(BQ).max()
    # the first columns, before the Krylov basis turns ill-conditioned, equal MGS Arnoldi's
    _, Hm = mgs_arnoldi(B, v, 6)
    assert np.abs(H[:7, :6] - Hm).max() <= 1e-12 * np.abs(Hm).max()
    # exp(-B / 2) v from the Krylov basis
    ref = expm(-0.5 * B) @ v
    kry = np.linalg.norm(v) * V[:m].T @ expm(-0.5 * H[:m, :m])[:, 0]
    assert np.abs(kry - ref).max() <= 1e-12 * np.abs(ref).max()
    # without the correction the relation fails where candidates kept a part in span(Q) (an
    # explicit second projection removes it before the candidate is used)
    raw = H.copy()
    for k in range(1, m):
        raw[:, k] += sum((rec[k - 1]["s_used"][q] / rec[k - 1]["beta"]) * H[:, q] for q in range(k))
    err_raw = np.abs(BQ - V.T @ raw).max() / np.abs(BQ).max()
    if leak > 0.0 and not force:
        assert err_raw > 1e-3
