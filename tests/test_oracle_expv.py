"""Pins of the oracle's Krylov expv (Sidje's Expokit expv restated; num::expv is absent, call
sites include/markovkit/fsp.hpp:152 and :225) and of the box generator of reaction networks.

* a two-state chain against its closed form p(t) = pi + exp(-(a + b) t) (p0 - pi);
* the toggle switch on a small box against scipy's expm_multiply (a library routine);
* conservation: with the reflecting boundary the probability stays 1; with the absorbing one
  the lost mass equals the integrated leak flux (checked against expm_multiply as well);
* the semigroup law exp((t1 + t2) Q) v = exp(t2 Q) exp(t1 Q) v;
* the box generator against a loop-by-loop restatement of reaction_network::transitions.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

from markovkit_amd.fsp import (ReactionNetwork, Reaction, birth_death, box_generator, box_states,
                               toggle_switch)
from oracle.oracle import expv_jumps


def q_matrix(first, to, rate, exit_):
    n = len(first) - 1
    cols = np.repeat(np.arange(n), np.diff(first))
    rows = np.concatenate([np.arange(n), to])
    cc = np.concatenate([np.arange(n), cols])
    vals = np.concatenate([-exit_, rate])
    return sp.csr_matrix((vals, (rows, cc)), shape=(n, n))


def test_two_state_closed_form():
    a, b = 3.0, 0.5
    first = np.array([0, 1, 2])
    to = np.array([1, 0])
    rate = np.array([a, b])
    exit_ = np.array([a, b])
    p0 = np.array([1.0, 0.0])
    pi = np.array([b, a]) / (a + b)
    for t in (0.01, 0.7, 5.0):
        w, steps, _ = expv_jumps(first, to, rate, exit_, t, p0, krylov_dim=30, tol=1e-12)
        exact = pi + np.exp(-(a + b) * t) * (p0 - pi)
        assert np.abs(w - exact).max() <= 1e-12, (t, w, exact)


@pytest.mark.parametrize("absorbing", [True, False])
def test_toggle_switch_against_expm_multiply(absorbing):
    first, to, rate, exit_, leak = box_generator(toggle_switch(), [30, 30], absorbing=absorbing)
    n = len(first) - 1
    assert n == 31 * 31
    p0 = np.zeros(n)
    p0[0] = 1.0
    Q = q_matrix(first, to, rate, exit_)
    for t in (0.5, 2.0):
        w, steps, rej = expv_jumps(first, to, rate, exit_, t, p0, tol=1e-10)
        ref = expm_multiply(t * Q, p0)
        assert np.abs(w - ref).sum() <= 1e-8, (t, np.abs(w - ref).sum())
        if not absorbing:
            assert abs(w.sum() - 1.0) <= 1e-10
        else:
            assert w.sum() < 1.0


def test_semigroup():
    first, to, rate, exit_, _ = box_generator(birth_death(20.0, 1.0), [60])
    n = len(first) - 1
    p0 = np.zeros(n)
    p0[5] = 1.0
    w12, _, _ = expv_jumps(first, to, rate, exit_, 3.0, p0, tol=1e-12)
    w1, _, _ = expv_jumps(first, to, rate, exit_, 1.0, p0, tol=1e-12)
    w2, _, _ = expv_jumps(first, to, rate, exit_, 2.0, w1, tol=1e-12)
    assert np.abs(w12 - w2).sum() <= 1e-9
    # a plausible mistake fails: the transposed generator gives another answer
    Q = q_matrix(first, to, rate, exit_)
    wrong = expm_multiply(3.0 * Q.T.tocsr(), p0)
    assert np.abs(wrong - w12).sum() > 1e-3


def test_box_generator_matches_transition_loop():
    net = ReactionNetwork([
        Reaction([1, 0], lambda x: 2.0 + 0.5 * x[:, 1]),
        Reaction([-1, 1], lambda x: 1.5 * x[:, 0]),
        Reaction([0, -1], lambda x: np.where(x[:, 1] > 2, 0.0, 0.3 * x[:, 1])),
    ])
    bounds = [3, 4]
    first, to, rate, exit_, leak = box_generator(net, bounds, absorbing=True)
    X = box_states(bounds)
    index = {tuple(x): i for i, x in enumerate(X)}
    for i, x in enumerate(X):
        jumps, lk = [], 0.0
        for r in net.reactions:
            rt = float(r.rate(x[None, :])[0])
            y = x + np.asarray(r.change)
            if not rt > 0 or (y < 0).any():
                continue
            if tuple(y) in index:
                jumps.append((index[tuple(y)], rt))
            else:
                lk += rt
        got = list(zip(to[first[i]:first[i + 1]], rate[first[i]:first[i + 1]]))
        assert [(int(a), float(b)) for a, b in got] == jumps
        assert leak[i] == lk
        assert abs(exit_[i] - (sum(r for _, r in jumps) + lk)) <= 1e-15
