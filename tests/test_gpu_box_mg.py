"""The device whole-box multigrid (boxmg.hip, through the C ABI) against the oracle's
restatement of the reference's paper driver (oracle/box_mg.hpp) with the same options.

- Operators: the finest generator's shifted apply equals the oracle's level matrix (the driver's
  assembly, pinned to its recorded runs) to rounding.
- One cycle from a random guess (V, W and K cycles; four-colour Gauss-Seidel and Jacobi;
  Thomas and dense terminal levels; 1D, 2D and 3D; linear and Schlogl rates) equals the
  oracle's cycle to 1e-11 relative (FP64 on both sides; the summation orders of the residual and
  the restriction differ, the formulas do not).
- The standalone iteration takes the oracle's number of cycles and follows its error history.
- The answer of recorded driver runs: the device solve of the same system ends with the
  recorded mass and minimum probability (the solution does not depend on the smoother).
- Graph replay of the standalone iteration gives bitwise the eager iteration.
- Flexible GMRES preconditioned by the cycle (the driver's mg-gmres) converges to the same
  solution.
"""
import json
import os

import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd import boxmg
from oracle.oracle import BoxMultigrid as Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "box_mg_recorded.json")) as f:
    RECORDED = json.load(f)

SCHLOGL = None  # the driver's default rates
LIN = boxmg.linear(1.0)

CASES = {
    "1d-16-cap3": dict(dimension=1, cells=16, cap=3, shift=0.5, rates=LIN),
    "1d-64-cap2": dict(dimension=1, cells=64, cap=2, shift=0.5, rates=LIN),
    "2d-4-cap4-schlogl": dict(dimension=2, cells=4, cap=4, shift=0.5, rates=SCHLOGL),
    "3d-4-cap2": dict(dimension=3, cells=4, cap=2, shift=0.5, rates=LIN),
    "1d-4-cap40-schlogl": dict(dimension=1, cells=4, cap=40, shift=0.05, rates=SCHLOGL),
}

ORACLE_SMOOTHER = {"colour-gs": "colour-gs", "jacobi": "jacobi", "gs": "gs", "block-ilu": "block-ilu"}


def pair(case, smoother="colour-gs", sweeps=2, cycles=2, kcycle=False, terminal_rows=0):
    kw = dict(CASES[case])
    common = dict(sweeps=sweeps, cycles=cycles, kcycle=kcycle, terminal_rows=terminal_rows)
    G = mk.BoxMultigrid(smoother=smoother, **kw, **common)
    O = Oracle(smoother=ORACLE_SMOOTHER[smoother], **kw, **common)
    return G, O


@pytest.mark.parametrize("case", list(CASES))
def test_finest_operator_matches_oracle(case):
    G, O = pair(case)
    assert G.rows() == O.n
    x = np.random.default_rng(2).standard_normal(G.rows())
    Q = G.generator()
    y = Q.apply_shifted(G.shift, x)
    want = O.apply_level(0, x)
    assert np.max(np.abs(y - want)) <= 1e-13 * np.max(np.abs(want))
    for l in range(G.levels()):
        rows, voxels, entries, terminal = G.level(l)
        assert rows == O.rows(l)
        assert terminal == O.terminal(l)


@pytest.mark.parametrize("terminal_rows", [0, 200], ids=["thomas", "dense200"])
@pytest.mark.parametrize("cycle", ["V", "W", "K"])
@pytest.mark.parametrize("smoother", ["colour-gs", "jacobi", "gs", "block-ilu"])
@pytest.mark.parametrize("case", list(CASES))
def test_cycle_matches_oracle(case, smoother, cycle, terminal_rows):
    G, O = pair(case, smoother, sweeps=2, cycles=1 if cycle == "V" else 2, kcycle=cycle == "K",
                terminal_rows=terminal_rows)
    assert G.levels() == O.levels()
    rng = np.random.default_rng(11)
    b = rng.random(G.rows())
    x0 = rng.random(G.rows()) * 0.1
    if G.level(0)[3]:
        got, want = G.apply(b), O.apply(b)
    else:
        got, want = G.cycle(b, x0), O.cycle(b, x0)
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want)), np.max(np.abs(got - want))


@pytest.mark.parametrize("cycle", ["V", "W", "K"])
@pytest.mark.parametrize("smoother", ["colour-gs", "jacobi", "gs", "block-ilu"])
@pytest.mark.parametrize("case", ["1d-64-cap2", "2d-4-cap4-schlogl", "1d-4-cap40-schlogl"])
def test_standalone_iteration_matches_oracle(case, smoother, cycle):
    G, O = pair(case, smoother, sweeps=2, cycles=1 if cycle == "V" else 2, kcycle=cycle == "K")
    b, x0 = O.initial_problem(2)
    xo, ito, erro, ho = O.solve(b, x0, 1e-9, 400)
    xg, itg, errg, hg = G.solve(b, x0, 1e-9, 400)
    assert erro < 1e-9 and errg < 1e-9
    assert abs(itg - ito) <= 1, (itg, ito)
    m = min(len(hg), len(ho))
    rel = np.abs(hg[:m] - ho[:m]) / ho[:m]
    # rounding differences grow as the residual falls: tight while it is within 1e-3 of the
    # start, loose near the tolerance (where the two sums cancel to a few digits)
    early = ho[:m] >= 1e-3 * ho[0]
    assert np.all(rel[early] <= 1e-8) and np.all(rel <= 2e-2), rel.max()
    assert np.abs(xg - xo).sum() <= errg + erro + 1e-14


@pytest.mark.parametrize("name,terminal_rows", [("voxels_1d_64_mg", 0), ("voxels_2d_8_mg", 0),
                                                ("voxels_3d_4_mg", 0), ("block_v16_k4_gs_s1", 0),
                                                ("block_schlogl_v4_k40_gs_s1", 0),
                                                ("voxels_1d_1024_mg", 600)])
def test_recorded_answer(name, terminal_rows):
    """the device solve of a recorded driver run ends with the recorded mass and minimum
    probability (the unique solution to within the l1 bound)"""
    rec = RECORDED[name]
    p = rec["params"]
    G = mk.BoxMultigrid(p["dimension"], p["cells"], p["cap"], p["diffusion"], p["shift"],
                        tuple(p["rates"]), "colour-gs", p["sweeps"], p["cycles"], p["kcycle"],
                        terminal_rows)
    assert G.rows() == rec["states"]
    b, x0 = G.initial_problem(p["initial"])
    x, it, err, hist = G.solve(b, x0, p["tolerance"], p["max_iterations"])
    assert err < p["tolerance"], (it, err)
    assert abs(x.sum() - rec["mass"]) <= 2 * p["tolerance"] + 1e-11
    assert abs(x.min() - rec["min_probability"]) <= 2 * p["tolerance"] + 1e-6 * rec["min_probability"]
    print(f"[{name}] {G.rows()} states: {it} cycles on the device (driver: {rec['iterations']} "
          f"lexicographic-GS cycles)")


@pytest.mark.parametrize("name", ["graph_gap_cap2_w", "voxels_1d_64_mg", "voxels_1d_64_kcycle",
                                  "voxels_2d_8_mg", "voxels_3d_4_mg", "block_v16_k4_gs_s1",
                                  "block_v16_k4_block-ilu_s1", "block_v16_k4_block-ilu_s2",
                                  "block_schlogl_v4_k40_block-ilu_s1_kcycle",
                                  "block_v64_k4_block-ilu_s1"])
def test_driver_smoothers_reproduce_recorded_cycles(name):
    """with the driver's own smoother (lexicographic GS or population-block ILU) the device
    takes the recorded number of cycles (one more or fewer where the last cycle lands at the
    tolerance) and ends with the recorded mass"""
    rec = RECORDED[name]
    p = rec["params"]
    G = mk.BoxMultigrid(p["dimension"], p["cells"], p["cap"], p["diffusion"], p["shift"],
                        tuple(p["rates"]), p["smoother"], p["sweeps"], p["cycles"], p["kcycle"])
    b, x0 = G.initial_problem(p["initial"])
    x, it, err, hist = G.solve(b, x0, p["tolerance"], p["max_iterations"])
    assert err < p["tolerance"]
    assert abs(it - rec["iterations"]) <= 1, (it, rec["iterations"])
    assert abs(x.sum() - rec["mass"]) <= 2 * p["tolerance"] + 1e-11
    if "history" in rec:
        H = np.array(rec["history"])
        m = min(len(H), len(hist))
        assert np.all(np.abs(hist[:5] - H[:5]) <= 1e-9 * H[:5])
    print(f"[{name}] {G.rows()} states, {p['smoother']}: {it} cycles (driver {rec['iterations']})")


@pytest.mark.parametrize("cycle", ["W", "K"])
@pytest.mark.parametrize("smoother", ["colour-gs", "jacobi"])
@pytest.mark.parametrize("case", ["1d-64-cap2", "2d-4-cap4-schlogl", "1d-4-cap40-schlogl", "3d-4-cap2"])
def test_small_level_kernels_match_per_operation_kernels(case, smoother, cycle):
    """levels of at most 2,048 states run each visit as two single-workgroup kernels; the
    cycle equals the per-operation kernels' (same formulas and summation orders)"""
    G, _ = pair(case, smoother, cycles=2, kcycle=cycle == "K")
    rng = np.random.default_rng(17)
    b = rng.random(G.rows())
    x0 = rng.random(G.rows()) * 0.1
    try:
        boxmg.small_levels_enable(False)
        z0 = G.cycle(b, x0)
        boxmg.small_levels_enable(True)
        z1 = G.cycle(b, x0)
    finally:
        boxmg.small_levels_enable(True)
    diff = np.max(np.abs(z1 - z0))
    print(f"[{case} {smoother} {cycle}] fused vs per-operation max diff {diff:.3e}")
    assert np.array_equal(z1, z0)


@pytest.mark.parametrize("smoother", ["colour-gs", "gs", "block-ilu"])
def test_graph_replay_is_bitwise_eager(smoother):
    G = mk.BoxMultigrid(1, 64, 3, 1.0, 0.5, LIN, smoother, 2, 2)
    b, x0 = G.initial_problem(2)
    try:
        boxmg.graphs_enable(False)
        xe, ite, erre, he = G.solve(b, x0, 1e-10, 300)
        boxmg.graphs_enable(True)
        for _ in range(2):
            xg, itg, errg, hg = G.solve(b, x0, 1e-10, 300)
            assert itg == ite
            assert np.array_equal(xg, xe)
            assert np.array_equal(hg, he)
    finally:
        boxmg.graphs_enable(True)


@pytest.mark.parametrize("terminal_rows", [0, 300])
def test_gmres_with_box_cycle(terminal_rows):
    G = mk.BoxMultigrid(2, 8, 3, 1.0, 0.5, None, "colour-gs", 1, 1, terminal_rows=terminal_rows)
    O = Oracle(2, 8, 3, 1.0, 0.5, None, "colour-gs", 1, 1, terminal_rows=terminal_rows)
    b, x0 = O.initial_problem(2)
    xo, ito, erro, _ = O.solve(b, x0, 1e-11, 500)
    x = x0.copy()
    tol = 1e-11 * G.shift / np.sqrt(G.rows())
    r = G.gmres(b, x, tolerance=tol, restart=30)
    assert r.converged
    resid = np.abs(b - G.generator().apply_shifted(G.shift, x)).sum() / G.shift
    assert np.abs(x - xo).sum() <= resid + erro + 1e-13
    with pytest.raises(ValueError):
        K = mk.BoxMultigrid(2, 8, 3, 1.0, 0.5, None, "colour-gs", 1, 2, kcycle=True)
        K.gmres(b, x0.copy(), tolerance=tol)


def test_one_voxel_and_dense_finest_are_exact():
    for kw in (dict(dimension=1, cells=1, cap=30), dict(dimension=2, cells=4, cap=3, terminal_rows=5000)):
        G = mk.BoxMultigrid(shift=0.3, **kw)
        O = Oracle(shift=0.3, **kw)
        b = np.random.default_rng(4).random(G.rows())
        z = G.apply(b)
        assert np.allclose(z, np.linalg.solve(O.dense(0), b), rtol=1e-11, atol=1e-14)


def test_rejects_bad_parameters():
    for kw in (dict(cells=6), dict(cycles=3), dict(dimension=4), dict(shift=0.0), dict(sweeps=0),
               dict(terminal_rows=100000)):
        with pytest.raises(ValueError):
            mk.BoxMultigrid(**{**dict(dimension=1, cells=4, cap=2, shift=0.5), **kw})
    with pytest.raises(ValueError):
        mk.BoxMultigrid(smoother="lexicographic")
