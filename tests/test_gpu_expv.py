"""GPU parity of the Krylov exponential (mk_expv): w = exp(t Q) v against the CPU oracle's expv
(the same algorithm, restated), scipy's expm_multiply (a library routine) and closed forms.

Tolerance: both Krylov results carry local errors below the requested tolerance per unit time,
so they are compared with the exact answer at 1e-8 in L1 for tol = 1e-10, and with each other
at 1e-7 for the 40,401-state toggle switch at tol = 1e-8.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

import markovkit_amd as mk
from markovkit_amd.fsp import birth_death, box_generator, toggle_switch
from oracle.oracle import expv_jumps

pytestmark = pytest.mark.gpu


def generator(first, to, rate, exit_, leak):
    return mk.Generator.from_jumps(first, to, rate, exit_, leak)


def q_matrix(first, to, rate, exit_):
    n = len(first) - 1
    cols = np.repeat(np.arange(n), np.diff(first))
    rows = np.concatenate([np.arange(n), to])
    cc = np.concatenate([np.arange(n), cols])
    return sp.csr_matrix((np.concatenate([-exit_, rate]), (rows, cc)), shape=(n, n))


def test_two_state_closed_form():
    a, b = 3.0, 0.5
    Q = generator(np.array([0, 1, 2]), np.array([1, 0]), np.array([a, b]), np.array([a, b]),
                  np.zeros(2))
    pi = np.array([b, a]) / (a + b)
    p0 = np.array([1.0, 0.0])
    for t in (0.01, 0.7, 5.0):
        w, r = mk.expv(Q, t, p0, tolerance=1e-12)
        exact = pi + np.exp(-(a + b) * t) * (p0 - pi)
        assert np.abs(w - exact).max() <= 1e-12, (t, w, exact)


@pytest.mark.parametrize("absorbing", [True, False])
def test_toggle_switch_small_box(absorbing):
    J = box_generator(toggle_switch(), [30, 30], absorbing=absorbing)
    Q = generator(*J)
    n = Q.rows()
    p0 = np.zeros(n)
    p0[0] = 1.0
    A = q_matrix(*J[:4])
    for t in (0.5, 2.0):
        w, r = mk.expv(Q, t, p0, tolerance=1e-10)
        ref = expm_multiply(t * A, p0)
        assert np.abs(w - ref).sum() <= 1e-8, (t, np.abs(w - ref).sum())
        wo, steps, _ = expv_jumps(*J[:4], t, p0, tol=1e-10)
        assert np.abs(w - wo).sum() <= 1e-8
        print(f"[toggle 31x31 absorbing={absorbing}] t={t} gpu steps={r.steps} "
              f"rejected={r.rejected} oracle steps={steps}")
        if not absorbing:
            assert abs(w.sum() - 1.0) <= 1e-10


def test_rdme_generator_expv():
    """exp(t Q) on an RDME box generator (mk_generator_create_rdme) against the dense exponential
    of the oracle's matrix."""
    from markovkit_amd.workloads import birth_death_1d
    from oracle.oracle import Problem
    w = birth_death_1d(8, 2, 0.5, "empty")
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    P = Problem(w.model, w.max_totals)
    p0 = np.zeros(P.n)
    p0[0] = 1.0
    A = P.matrix(0.0, 1.0)  # Q itself
    out, r = mk.expv(Q, 0.8, p0, tolerance=1e-10)
    ref = expm_multiply(0.8 * A, p0)
    assert np.abs(out - ref).sum() <= 1e-8


def test_toggle_switch_40401_states():
    """The serving mix's FSP request: toggle switch on 0..200 x 0..200 (40,401 states),
    absorbing box, p0 at (10, 10), exp(5 Q) p0 at tol 1e-8, against the oracle's expv."""
    J = box_generator(toggle_switch(), [200, 200], absorbing=True)
    Q = generator(*J)
    n = Q.rows()
    assert n == 40401
    p0 = np.zeros(n)
    p0[10 * 201 + 10] = 1.0
    w, r = mk.expv(Q, 5.0, p0, tolerance=1e-8)
    wo, steps, _ = expv_jumps(*J[:4], 5.0, p0, tol=1e-8)
    print(f"[toggle 201x201] gpu steps={r.steps} rejected={r.rejected} oracle steps={steps} "
          f"mass={w.sum():.12f} l1={np.abs(w - wo).sum():.2e}")
    assert np.abs(w - wo).sum() <= 1e-7
    assert 0.0 < w.sum() <= 1.0 + 1e-12
    assert w.min() >= -1e-9
