"""The fused multigrid tail (levels of at most 1024 states in one workgroup-wide kernel), with
its levels in HBM or staged in LDS, against the per-level kernels: same formulas and summation
orders, so the cycle and the solves must be bitwise identical, in FP32 and FP64, with the Jacobi and red-black Gauss-Seidel smoothers, on
1D and 2D boxes, for a whole cycle in the tail (969 states) and for a tail below larger levels
(58,905 states)."""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.workloads import annihilation_2d, birth_death_1d, config_small

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("smoother", [mk.JACOBI, mk.RED_BLACK], ids=["jacobi", "red_black"])
@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("wl", [config_small(), annihilation_2d(4, (2, 2), 1.0),
                                birth_death_1d(32, 4, 0.7, "random")], ids=["969", "2d", "58905"])
def test_tail_matches_per_level_kernels(wl, precision, smoother):
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    Q = mk.Generator(box, wl.model)
    s = wl.shifts[0]
    M = mk.SectorMultigrid(wl.model.domain, box, s,
                           mk.MultigridOptions(precision=precision, smoother=smoother))
    rng = np.random.default_rng(3)
    b = rng.random(box.size())
    try:
        out = {}
        for name, tail, lds in (("levels", False, False), ("tail", True, False), ("lds", True, True)):
            mk.fused_tail_enable(tail)
            mk.tail_lds_enable(lds)
            z = M.apply(b)
            x = np.zeros(box.size())
            r = mk.gmres(mk.ShiftedGenerator(Q, s), M, b, x, tolerance=1e-10 * np.linalg.norm(b))
            assert r.converged
            out[name] = (z, x, r.iterations)
        for name in ("tail", "lds"):
            assert np.array_equal(out["levels"][0], out[name][0]), name
            assert np.array_equal(out["levels"][1], out[name][1]), name
            assert out["levels"][2] == out[name][2], name
    finally:
        mk.fused_tail_enable(True)
        mk.tail_lds_enable(True)
