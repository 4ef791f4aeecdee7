"""The autotuner (markovkit_amd.autotune): it times every candidate shape, returns the fastest
converged one, leaves the XCD chunk it picked set, and the solve with the tuned options reaches
the tolerance; the batch tuner does the same for a batched multi-shift solve."""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd import autotune
from markovkit_amd.workloads import birth_death_1d

pytestmark = pytest.mark.gpu


def test_tune_solve_picks_the_fastest_trial():
    w = birth_death_1d(32, 4, 0.5, "random")
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    b = w.rhs(0, box.size())
    start = autotune.xcd_chunk()
    res = autotune.tune_solve(Q, w.model.domain, box, w.shifts[0], b, sweeps=((2, 2), (1, 1)),
                              chunks=(16, 0))
    try:
        assert len(res.trials) == 4
        assert all(np.isfinite(t.seconds) and t.iterations > 0 for t in res.trials)
        assert res.seconds == min(t.seconds for t in res.trials)
        assert autotune.xcd_chunk() == res.xcd_chunk
        M = mk.SectorMultigrid(w.model.domain, box, w.shifts[0],
                               res.options(precision=mk.FP32, smoother=mk.RED_BLACK))
        x = np.zeros(box.size())
        r = mk.gmres(mk.ShiftedGenerator(Q, w.shifts[0]), M, b, x,
                     tolerance=1e-10 * np.linalg.norm(b), restart=50)
        assert r.converged
        print("[autotune]", [(t.pre, t.post, t.xcd_chunk, round(t.seconds * 1e3, 2), t.iterations)
                             for t in res.trials])
    finally:
        autotune.set_xcd_chunk(start)


def test_tune_batch():
    w = birth_death_1d(16, 3, 0.5, "random")
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    shifts = [0.1, 1.0, 10.0]
    B = np.random.default_rng(2).random((3, box.size()))
    res = autotune.tune_batch(Q, w.model.domain, box, shifts, B, sweeps=((2, 2), (1, 1)))
    # both smoothers (Jacobi, red-black) times both sweep pairs
    assert len(res.trials) == 4 and (res.pre, res.post) in ((2, 2), (1, 1))
    assert {t.smoother for t in res.trials} == {mk.JACOBI, mk.RED_BLACK}
    assert res.smoother in (mk.JACOBI, mk.RED_BLACK)
    assert all(np.isfinite(t.seconds) for t in res.trials)
    only = autotune.tune_batch(Q, w.model.domain, box, shifts, B, smoothers=(mk.JACOBI,),
                               sweeps=((2, 2), (1, 1)))
    assert len(only.trials) == 2 and only.smoother == mk.JACOBI
