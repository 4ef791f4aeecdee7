"""Time stepping and adaptive FSP with the device whole-box multigrid (four-colour Gauss-Seidel)
against the reference driver's recorded outputs (tests/golden/box_mg_steps.json,
box_mg_adaptive.json). The driver used lexicographic Gauss-Seidel, so cycle counts differ; the
distributions do not (each solve is exact to its l1 error bound):

- time_check_s10 (861 states, Rannacher start + Crank-Nicolson, 11 solves) and
  transient_k80_s20 (1,929,501 states, 21 solves): every solve's mass and high-population mass,
  and for the latter the recorded population marginal after every solve;
- the adaptive FSP runs: the same sequence of attempts (step, cap, states, accepted) and the
  same missing masses.
"""
import json
import os
import time

import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd import boxmg

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "box_mg_steps.json")) as f:
    STEPS = json.load(f)
with open(os.path.join(HERE, "golden", "box_mg_adaptive.json")) as f:
    ADAPTIVE = json.load(f)


def device_for(p, cap=None, terminal_rows=0):
    return mk.BoxMultigrid(p["dimension"], p["cells"], p["cap"] if cap is None else cap,
                           p["diffusion"], p["shift"], tuple(p["rates"]), "colour-gs", p["sweeps"],
                           p["cycles"], p["kcycle"], terminal_rows)


@pytest.mark.parametrize("name", ["time_check_s10", "transient_k80_s20"])
def test_crank_nicolson_transient(name):
    rec = STEPS[name]
    p = rec["params"]
    t0 = time.perf_counter()
    M = device_for(p)
    setup = time.perf_counter() - t0
    assert M.rows() == rec["states"]
    _, p0 = M.initial_problem(p["initial"])
    marg = []
    t0 = time.perf_counter()
    _, records = boxmg.time_steps(M, p0, p["steps"], p["integrator"], p["tolerance"],
                                  p["max_iterations"],
                                  on_step=lambda r, x: marg.append(boxmg.marginal(x, M.voxels, M.cap)))
    wall = time.perf_counter() - t0
    assert len(records) == len(rec["solves"])
    drift = 0.0
    for got, want in zip(records, rec["solves"]):
        assert got["converged"]
        drift += got["error"] + want["error"]
        assert abs(got["mass"] - want["mass"]) <= drift + 1e-12
        assert abs(got["high_mass"] - want["high_mass"]) <= drift + 1e-12
    if "marginals" in rec:
        times = [r["time"] for r in records]
        assert np.allclose(times, rec["marginal_times"])
        for got, want in zip(marg, rec["marginals"]):
            assert np.max(np.abs(np.asarray(got) - np.asarray(want))) <= drift + 1e-12
    cycles = sum(r["iterations"] for r in records)
    print(f"[{name}] {M.rows()} states: setup {setup:.2f}s, {len(records)} solves, {cycles} cycles, "
          f"{wall:.2f}s with host transfers (driver: {rec['wall_s']:.1f}s wall, "
          f"{sum(s['iterations'] for s in rec['solves'])} cycles)")


@pytest.mark.parametrize("name", sorted(ADAPTIVE))
def test_adaptive_fsp_attempts(name):
    rec = ADAPTIVE[name]
    p = rec["params"]
    V = p["cells"] ** p["dimension"]
    _, cap, attempts = boxmg.adaptive_fsp(lambda c: device_for(p, cap=c), V, p["initial"],
                                          p["shift"], p["steps"], p["fsp_tolerance"],
                                          p["tolerance"], p["cap"] if p["cap_given"] else None,
                                          p["max_iterations"])
    want = rec["attempts"]
    assert [(a["step"], a["cap"], a["states"], int(a["accepted"])) for a in attempts] == \
           [(a["step"], a["cap"], a["states"], a["accepted"]) for a in want]
    for a, w in zip(attempts, want):
        assert abs(a["missing_mass"] - w["missing_mass"]) <= 1e-6 * w["missing_mass"] + 1e-11
