"""HIP-graph replay of solver iterations and expv steps: a configuration runs eagerly on its first
use, is captured on the second and replayed afterwards. Every replay must give bitwise the
result of eager launches (graphs off), with eager work and synchronisations between replays.
"""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.fsp import box_generator, toggle_switch
from markovkit_amd.workloads import annihilation_2d, birth_death_1d

pytestmark = pytest.mark.gpu


def solve(Q, M, s, b, restart=50):
    x = np.zeros(Q.rows())
    r = mk.gmres(mk.ShiftedGenerator(Q, s), M, b, x, tolerance=1e-10 * np.linalg.norm(b),
                 restart=restart)
    assert r.converged
    return x, r.iterations


@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("wl", [birth_death_1d(16, 3, 0.5, "random"), annihilation_2d(4, (2, 2), 1.0)],
                         ids=["1d", "2d"])
def test_solver_replays_match_eager(wl, precision):
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    Q = mk.Generator(box, wl.model)
    s = wl.shifts[0]
    M = mk.SectorMultigrid(wl.model.domain, box, s, mk.MultigridOptions(precision=precision))
    b = wl.rhs(0, box.size())
    b2 = np.random.default_rng(5).random(box.size())
    try:
        # graphs capture the host-driven loop's steps, not the one-workgroup solve
        mk.tiny_gmres_enable(False)
        mk.graphs_enable(False)
        x_eager, it = solve(Q, M, s, b)
        x2_eager, _ = solve(Q, M, s, b2)
        mk.graphs_enable(True)
        for rep in range(4):  # eager, capture, replay, replay
            x, it_g = solve(Q, M, s, b)
            assert it_g == it
            assert np.array_equal(x, x_eager), rep
            # other work between replays: an eager SpMV and a solve of another rhs
            Q.apply_shifted(s, b2)
            x2, _ = solve(Q, M, s, b2)
            assert np.array_equal(x2, x2_eager), rep
        # a shift change re-targets the hierarchy; the old shift's graphs stay valid afterwards
        M.set_shift(2.0 * s)
        solve(Q, M, 2.0 * s, b)
        M.set_shift(s)
        x, _ = solve(Q, M, s, b)
        assert np.array_equal(x, x_eager)
    finally:
        mk.graphs_enable(False)
        mk.tiny_gmres_enable(True)


def test_expv_replays_match_eager():
    J = box_generator(toggle_switch(), [40, 40], absorbing=True)
    F = mk.Generator.from_jumps(*J)
    v = np.zeros(F.rows())
    v[5 * 41 + 7] = 1.0
    try:
        mk.graphs_enable(False)
        w_eager, r = mk.expv(F, 2.0, v)
        mk.graphs_enable(True)
        for rep in range(3):
            w, rg = mk.expv(F, 2.0, v)
            assert rg.steps == r.steps
            assert np.array_equal(w, w_eager), rep
    finally:
        mk.graphs_enable(False)
