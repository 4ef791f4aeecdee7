"""GPU tests of batched Krylov exponentials (mk_expv_batch and the server's expv engines):
many vectors over one generator advance together, each with Expokit's step control of its own.

- closed form: the two-state chain, several times at once (t = 0 and a zero vector included);
- scipy's expm_multiply and the CPU oracle's expv on the toggle switch (absorbing and
  reflecting boxes; tol 1e-10, compared at 1e-8 in L1 as the single expv is);
- a vector's result does not depend on the vectors beside it: alone and in a mix of other
  vectors and times, bitwise, with the same step counts; the same for the server's engine;
- more vectors than lanes (continuous admission) against the single-vector expv.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

import markovkit_amd as mk
from markovkit_amd.fsp import box_generator, toggle_switch
from oracle.oracle import expv_jumps

pytestmark = pytest.mark.gpu


def q_matrix(first, to, rate, exit_):
    n = len(first) - 1
    cols = np.repeat(np.arange(n), np.diff(first))
    rows = np.concatenate([np.arange(n), to])
    cc = np.concatenate([np.arange(n), cols])
    return sp.csr_matrix((np.concatenate([-exit_, rate]), (rows, cc)), shape=(n, n))


def test_two_state_closed_form():
    a, b = 3.0, 0.5
    Q = mk.Generator.from_jumps(np.array([0, 1, 2]), np.array([1, 0]), np.array([a, b]),
                                np.array([a, b]), np.zeros(2))
    pi = np.array([b, a]) / (a + b)
    ts = [0.01, 0.7, 5.0, 0.0, 2.0]
    V = np.array([[1.0, 0.0], [0.3, 0.7], [0.0, 1.0], [0.2, 0.8], [0.0, 0.0]])
    W, res = mk.expv_batch(Q, ts, V, tolerance=1e-12)
    for t, v, w, r in zip(ts, V, W, res):
        exact = v.sum() * pi + np.exp(-(a + b) * t) * (v - v.sum() * pi)
        assert np.abs(w - exact).max() <= 1e-12, (t, v, w, exact)
    assert res[3].steps == 0 and res[4].steps == 0  # t = 0, zero vector: w = v


@pytest.mark.parametrize("absorbing", [True, False])
def test_toggle_switch_against_scipy_and_oracle(absorbing):
    J = box_generator(toggle_switch(), [30, 30], absorbing=absorbing)
    Q = mk.Generator.from_jumps(*J)
    n = Q.rows()
    A = q_matrix(*J[:4])
    rng = np.random.default_rng(5)
    ts = [0.5, 2.0, 1.3, 0.1, 3.0]
    V = np.zeros((5, n))
    V[0, 0] = 1.0
    V[1, 17] = 1.0
    V[2] = rng.random(n)
    V[2] /= V[2].sum()
    V[3, n - 1] = 1.0
    V[4, n // 2] = 0.5
    V[4, n // 3] = 0.5
    W, res = mk.expv_batch(Q, ts, V, tolerance=1e-10)
    for k, (t, v, w) in enumerate(zip(ts, V, W)):
        ref = expm_multiply(t * A, v)
        assert np.abs(w - ref).sum() <= 1e-8, (k, t, np.abs(w - ref).sum())
        if not absorbing:
            assert abs(w.sum() - v.sum()) <= 1e-10
    wo, steps, _ = expv_jumps(*J[:4], ts[0], V[0], tol=1e-10)
    assert np.abs(W[0] - wo).sum() <= 1e-8
    # the single-vector expv: same algorithm, other kernels
    w1, r1 = mk.expv(Q, ts[1], V[1], tolerance=1e-10)
    assert np.abs(W[1] - w1).sum() <= 1e-8
    print(f"[expv batch absorbing={absorbing}] steps {[r.steps for r in res]} single {r1.steps}")


def test_alone_and_in_the_mix_are_bitwise():
    J = box_generator(toggle_switch(), [40, 40], absorbing=True)
    Q = mk.Generator.from_jumps(*J)
    n = Q.rows()
    rng = np.random.default_rng(9)
    V = rng.random((7, n))
    V /= V.sum(axis=1, keepdims=True)
    ts = rng.uniform(0.2, 4.0, 7)
    Wmix, rmix = mk.expv_batch(Q, ts, V, tolerance=1e-8)
    for k in (0, 3, 6):
        Wa, ra = mk.expv_batch(Q, ts[k:k + 1], V[k:k + 1], tolerance=1e-8)
        assert np.array_equal(Wa[0], Wmix[k]), k
        assert (ra[0].steps, ra[0].rejected) == (rmix[k].steps, rmix[k].rejected)
        assert ra[0].error == rmix[k].error


def test_more_vectors_than_lanes():
    """300 vectors (256 lanes at a time, the rest admitted as lanes free) against the
    single-vector expv"""
    J = box_generator(toggle_switch(), [12, 12], absorbing=False)
    Q = mk.Generator.from_jumps(*J)
    n = Q.rows()
    rng = np.random.default_rng(2)
    V = rng.random((300, n))
    V /= V.sum(axis=1, keepdims=True)
    ts = rng.uniform(0.1, 3.0, 300)
    W, res = mk.expv_batch(Q, ts, V, tolerance=1e-10)
    for k in range(0, 300, 37):
        w1, _ = mk.expv(Q, ts[k], V[k], tolerance=1e-10)
        assert np.abs(W[k] - w1).sum() <= 1e-8, k
    assert np.abs(W.sum(axis=1) - 1.0).max() <= 1e-10


def test_server_engine_matches_the_batch_alone():
    """Exponentials submitted to a server with batching on run in its expv engine: each result
    is bitwise the batched expv of that vector alone, and within tolerance of the single expv."""
    J = box_generator(toggle_switch(), [40, 40], absorbing=True)
    F = mk.Generator.from_jumps(*J)
    n = F.rows()
    rng = np.random.default_rng(4)
    srv = mk.Server(workers=2)
    reqs = []
    for k in range(6):
        v = np.zeros(n)
        v[int(rng.integers(0, n))] = 1.0
        t = float(rng.uniform(0.5, 3.0))
        reqs.append((t, v, srv.submit_expv(F, t, v, tolerance=1e-8)))
    for t, v, rid in reqs:
        w, info = srv.wait(rid)
        assert info.kind == "expv"
        wa, ra = mk.expv_batch(F, [t], v[None, :], tolerance=1e-8)
        assert np.array_equal(w, wa[0])
        assert info.expv.steps == ra[0].steps
        w1, _ = mk.expv(F, t, v, tolerance=1e-8)
        assert np.abs(w - w1).sum() <= 1e-7
    srv.close()
