"""The FSP solvers (fsp.hpp:124-254) with every exponential on the device, against the
reference's own checks (tests/test_fsp.cpp:68-91) and stronger closed forms: a pure death from
5 molecules is Binomial(5, e^-t) at time t, and a birth-death process started empty is
Poisson(b/d (1 - e^-dt)). The Krylov FSP on the toggle switch is compared with the device expv
of the absorbing box generator that contains its support."""
import math

import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.fsp import (AdaptiveFspOptions, KrylovFspOptions, Reaction, ReactionNetwork,
                               birth_death, box, box_generator, solve_adaptive_fsp,
                               solve_krylov_fsp, toggle_switch)

pytestmark = pytest.mark.gpu


def test_krylov_fsp_pure_death():
    death = ReactionNetwork([Reaction([-1], lambda x: 1.0 * x[:, 0])])
    sol = solve_krylov_fsp(death, (5,), 1.0,
                           KrylovFspOptions(tolerance=1e-10, max_steps=100, krylov_tol=1e-12))
    assert sol.times[-1] == 1.0
    mean = sol.means()[-1, 0]
    assert abs(mean - 5.0 * math.exp(-1.0)) < 1e-8
    last = sol.snapshots[-1]
    assert abs(last.marginal(0)[0] - (1.0 - math.exp(-1.0)) ** 5) < 1e-8
    q = math.exp(-1.0)
    exact = {k: math.comb(5, k) * q ** k * (1 - q) ** (5 - k) for k in range(6)}
    got = last.marginal(0)
    assert sum(abs(got.get(k, 0.0) - v) for k, v in exact.items()) < 1e-8
    print(f"[krylov fsp] {len(sol.times)} snapshots, mean {mean:.12f}")


def test_adaptive_fsp_birth_death():
    sol = solve_adaptive_fsp(birth_death(5.0, 0.1), (0,), 2.0,
                             AdaptiveFspOptions(inside=box([100]), step_scale=0.1))
    assert len(sol.times) > 1
    mean = sol.means()[-1, 0]
    lam = (5.0 / 0.1) * (1.0 - math.exp(-0.1 * 2.0))
    assert abs(mean - lam) < 0.1
    got = sol.snapshots[-1].marginal(0)
    pois = {k: math.exp(-lam + k * math.log(lam) - math.lgamma(k + 1)) for k in range(101)}
    l1 = sum(abs(got.get(k, 0.0) - v) for k, v in pois.items()) + sum(
        v for k, v in got.items() if k not in pois)
    # the reflecting frontier grows one ring per step: the method's own truncation error is
    # about 1e-3 here; a pruning or normalisation mistake shows as an error of order 0.1
    assert l1 < 5e-3, l1
    print(f"[adaptive fsp] mean {mean:.6f} exact {lam:.6f}, L1 to Poisson {l1:.2e}, "
          f"support {sol.snapshots[-1].size()}")


def test_krylov_fsp_toggle_switch_against_box():
    net = toggle_switch()
    T = 0.5
    sol = solve_krylov_fsp(net, (0, 0), T, KrylovFspOptions(tolerance=1e-8, krylov_tol=1e-10))
    last = sol.snapshots[-1]
    X = np.array(last.states())
    bound = int(X.max()) + 5
    J = box_generator(net, [bound, bound], absorbing=True)
    F = mk.Generator.from_jumps(*J)
    v = np.zeros(F.rows())
    v[0] = 1.0
    w, _ = mk.expv(F, T, v, tolerance=1e-10)
    ref = {(i // (bound + 1), i % (bound + 1)): w[i] for i in range(len(w)) if w[i] > 0}
    p = last.probabilities()
    l1 = sum(abs(pv - ref.get(x, 0.0)) for x, pv in zip(last.states(), p))
    l1 += sum(rv for x, rv in ref.items() if last.find(x) is None)
    assert l1 < 1e-6, l1
    print(f"[krylov fsp toggle] support {last.size()} states, L1 to the box solution {l1:.2e}")
