"""The fused multigrid tail (levels of at most 2048 states in one workgroup-wide kernel), with
its levels in HBM or staged in LDS, against the per-level kernels: same formulas and summation
orders, so the cycle and the solves must be bitwise identical, in FP32 and FP64, with the
Jacobi and red-black Gauss-Seidel smoothers, on 1D and 2D boxes, for a whole cycle in the tail
(969 states) and for a tail below larger levels (58,905 states), with V-cycles, with
W-cycles on the top levels over dense bottom blocks (the bench's hierarchy shape, whose
bottom level is solved directly) and with three coarse cycles per visit; and the LDS image
after a shift change. The same for a red-black level's first half-sweep fused into the
restriction or the coarse-cycle residual against its own launch."""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.workloads import annihilation_2d, birth_death_1d, config_small

pytestmark = pytest.mark.gpu


SHAPES = {"V": {}, "W3-c100": dict(pre_sweeps=1, post_sweeps=2, coarse_cycles=2, cycle_levels=3,
                                   coarsest_size=100),
          # three coarse cycles per visit: the tail's middle cycles (accumulate, then a fresh
          # residual) run only with coarse_cycles >= 3
          "C3-L3": dict(pre_sweeps=1, post_sweeps=2, coarse_cycles=3, cycle_levels=3,
                        coarsest_size=100)}


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("smoother", [mk.JACOBI, mk.RED_BLACK], ids=["jacobi", "red_black"])
@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("wl", [config_small(), annihilation_2d(4, (2, 2), 1.0),
                                birth_death_1d(32, 4, 0.7, "random")], ids=["969", "2d", "58905"])
def test_tail_matches_per_level_kernels(wl, precision, smoother, shape):
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    Q = mk.Generator(box, wl.model)
    s = wl.shifts[0]
    M = mk.SectorMultigrid(wl.model.domain, box, s,
                           mk.MultigridOptions(precision=precision, smoother=smoother,
                                               **SHAPES[shape]))
    rng = np.random.default_rng(3)
    b = rng.random(box.size())
    try:
        # the host-driven Krylov loop in every arm (a whole hierarchy in LDS would otherwise be
        # solved in one workgroup, with its sums in another order)
        mk.tiny_gmres_enable(False)
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
        mk.tiny_gmres_enable(True)


@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
def test_tail_image_follows_set_shift(precision):
    """The LDS tail stages an image of its constant arrays assembled at setup; the terminal
    inverses change with the shift, so after set_shift the image must be rebuilt: a retargeted
    hierarchy gives the cycle of one built at the new shift (up to the rounding of the two
    inversions; a stale image would be off at the size of the change)."""
    wl = birth_death_1d(32, 4, 0.7, "random")
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    opts = mk.MultigridOptions(precision=precision, smoother=mk.RED_BLACK, pre_sweeps=1,
                               coarse_cycles=2, cycle_levels=2, coarsest_size=100)
    b = np.random.default_rng(5).random(box.size())
    M = mk.SectorMultigrid(wl.model.domain, box, 0.7, opts)
    z_old = M.apply(b)
    M.set_shift(3.0)
    z = M.apply(b)
    z_ref = mk.SectorMultigrid(wl.model.domain, box, 3.0, opts).apply(b)
    err = np.abs(z - z_ref).max() / np.abs(z_ref).max()
    assert err <= (1e-10 if precision == mk.FP64 else 1e-5), err
    assert np.abs(z_old - z_ref).max() / np.abs(z_ref).max() > 1e-2


FUSED_SHAPES = {"V": {}, "W2-c800": dict(pre_sweeps=1, post_sweeps=2, coarse_cycles=2, cycle_levels=2,
                                         coarsest_size=800),
                "C3-L2": dict(pre_sweeps=2, post_sweeps=1, coarse_cycles=3, cycle_levels=2)}


@pytest.mark.parametrize("shape", list(FUSED_SHAPES))
@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("wl", [annihilation_2d(4, (2, 2), 1.0), birth_death_1d(32, 4, 0.7, "random")],
                         ids=["2d", "58905"])
def test_fused_first_half_sweep_is_bitwise(wl, precision, shape):
    """The fused first half-sweep (restriction or residual kernel) against its own rb_first
    launch, with and without the fused tail: cycle and solve bitwise equal"""
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    Q = mk.Generator(box, wl.model)
    s = wl.shifts[0]
    M = mk.SectorMultigrid(wl.model.domain, box, s,
                           mk.MultigridOptions(precision=precision, smoother=mk.RED_BLACK,
                                               **FUSED_SHAPES[shape]))
    b = np.random.default_rng(5).random(box.size())
    try:
        for tail in (True, False):
            mk.fused_tail_enable(tail)
            out = {}
            for fused in (True, False):
                mk.fused_first_enable(fused)
                z = M.apply(b)
                x = np.zeros(box.size())
                r = mk.gmres(mk.ShiftedGenerator(Q, s), M, b, x, tolerance=1e-10 * np.linalg.norm(b))
                assert r.converged
                out[fused] = (z, x, r.iterations)
            assert np.array_equal(out[True][0], out[False][0]), tail
            assert np.array_equal(out[True][1], out[False][1]), tail
            assert out[True][2] == out[False][2]
    finally:
        mk.fused_first_enable(True)
        mk.fused_tail_enable(True)


DENSE_SHAPES = {"V-c800": dict(coarsest_size=800),
                "W1-c800": dict(pre_sweeps=1, post_sweeps=2, coarse_cycles=2, cycle_levels=1, coarsest_size=800)}


@pytest.mark.parametrize("shape", list(DENSE_SHAPES))
@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("wl", [annihilation_2d(4, (2, 2), 1.0), birth_death_1d(32, 4, 0.7, "random")],
                         ids=["2d", "58905"])
def test_coarse_fuse_is_bitwise(wl, precision, shape):
    """A direct coarse level (dense blocks of more than 128 states) solved and prolonged in one
    launch against its terminal solve and prolongation launches: cycle and solve bitwise equal"""
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    Q = mk.Generator(box, wl.model)
    s = wl.shifts[0]
    M = mk.SectorMultigrid(wl.model.domain, box, s,
                           mk.MultigridOptions(precision=precision, smoother=mk.RED_BLACK,
                                               **DENSE_SHAPES[shape]))
    b = np.random.default_rng(6).random(box.size())
    out = {}
    try:
        for fused in (True, False):
            mk.coarse_fuse_enable(fused)
            z = M.apply(b)
            x = np.zeros(box.size())
            r = mk.gmres(mk.ShiftedGenerator(Q, s), M, b, x, tolerance=1e-10 * np.linalg.norm(b))
            assert r.converged
            out[fused] = (z, x, r.iterations)
    finally:
        mk.coarse_fuse_enable(True)
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][1], out[False][1])
    assert out[True][2] == out[False][2]
