"""FGMRES in one workgroup: a problem of at most 1024 states whose sector hierarchy sits whole in
the LDS tail is solved by one kernel launch (restarts, Gram-Schmidt, Givens rotations and the
stopping tests on the device). It runs the host-driven loop's iteration with its sums in
another order, so it is checked against that loop and the CPU oracle through the L1 bound
||x - x*||_1 <= (||b - A x||_1 + ||b - A x*||_1) / s (rdme/cpp/rdme_full_multigrid.cpp:884-887),
with iteration counts within 2 of the oracle's for one FGMRES(50) cycle; and for the edge
cases of the host loop: restarts, an iteration cap, a zero and a non-finite right-hand side,
a warm start, and a problem whose hierarchy does not fit (which must take the host loop)."""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.workloads import annihilation_2d, birth_death_1d, config_small
from oracle.oracle import Problem, SectorMultigrid as OracleMG

pytestmark = pytest.mark.gpu

CASES = {"969": config_small(), "165": birth_death_1d(8, 3, 0.7, "random"),
         "2d_289": annihilation_2d(4, (1, 1), 1.0)}


def setup(wl, precision, smoother, **opts):
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    Q = mk.Generator(box, wl.model)
    s = wl.shifts[0]
    M = mk.SectorMultigrid(wl.model.domain, box, s,
                           mk.MultigridOptions(precision=precision, smoother=smoother, **opts))
    return box, Q, s, M


def solve(Q, M, s, b, x0=None, **kw):
    """solve and report whether it ran as the single one-workgroup launch"""
    x = np.zeros(len(b)) if x0 is None else x0.copy()
    mk.profile_reset()
    mk.profile_enable(True)
    try:
        r = mk.gmres(mk.ShiftedGenerator(Q, s), M, b, x, **kw)
        prof = mk.profile_read()
    finally:
        mk.profile_enable(False)
    one = prof["gmres_other"]["launches"] == 1 and prof["shifted_spmv"]["launches"] == 0
    return x, r, one


def residual(Q, s, b, x):
    return b - Q.apply_shifted(s, x)


@pytest.mark.parametrize("restart", [50, 5], ids=["m50", "m5"])
@pytest.mark.parametrize("smoother", [mk.JACOBI, mk.RED_BLACK], ids=["jacobi", "red_black"])
@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("name", list(CASES))
def test_tiny_matches_host_loop_and_oracle(name, precision, smoother, restart):
    wl = CASES[name]
    box, Q, s, M = setup(wl, precision, smoother)
    n = box.size()
    b = wl.rhs(0, n) if name == "969" else np.random.default_rng(7).random(n)
    tol = 1e-10 * np.linalg.norm(b)
    out = {}
    try:
        for tiny in (True, False):
            mk.tiny_gmres_enable(tiny)
            x, r, one = solve(Q, M, s, b, tolerance=tol, max_iterations=500, restart=restart)
            assert one == tiny, (tiny, one)
            assert r.converged, r
            rr = residual(Q, s, b, x)
            assert np.linalg.norm(rr) <= 1.0001 * tol
            assert abs(r.residual - np.linalg.norm(rr)) <= 1e-3 * tol
            out[tiny] = (x, rr, r.iterations)
    finally:
        mk.tiny_gmres_enable(True)
    (xt, rt, it_t), (xh, rh, it_h) = out[True], out[False]
    assert np.abs(xt - xh).sum() <= (np.abs(rt).sum() + np.abs(rh).sum()) / s
    assert abs(it_t - it_h) <= 2, (it_t, it_h)
    P = Problem(wl.model, wl.max_totals)
    x_ref, info = P.gmres(s, b, mg=OracleMG(P, s), tol=tol, max_iterations=500, restart=50)
    assert info["converged"]
    r_ref = b - P.apply_shifted(s, x_ref)
    assert np.abs(xt - x_ref).sum() <= (np.abs(rt).sum() + np.abs(r_ref).sum()) / s
    if restart == 50 and smoother == mk.JACOBI:
        # the oracle's cycle is the weighted-Jacobi one: one FGMRES(50) cycle, counts up to rounding
        assert abs(it_t - info["iterations"]) <= 2, (it_t, info["iterations"])
    print(f"[tiny {name} m={restart}] its={it_t} host its={it_h} oracle its={info['iterations']}")


@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
def test_tiny_edge_cases(precision):
    wl = config_small()
    box, Q, s, M = setup(wl, precision, mk.RED_BLACK)
    n = box.size()
    b = wl.rhs(0, n)
    tol = 1e-10 * np.linalg.norm(b)
    # an iteration cap inside the second restart cycle: not converged, exactly that many
    x, r, one = solve(Q, M, s, b, tolerance=1e-300, max_iterations=7, restart=4)
    assert one and not r.converged and r.iterations == 7
    assert abs(r.residual - np.linalg.norm(residual(Q, s, b, x))) <= 1e-6 * np.linalg.norm(b)
    # a warm start from that iterate converges, in fewer iterations than from zero
    _, r0, _ = solve(Q, M, s, b, tolerance=tol, restart=50)
    xw, rw, one = solve(Q, M, s, b, x0=x, tolerance=tol, restart=50)
    assert one and rw.converged and rw.iterations < r0.iterations
    assert np.linalg.norm(residual(Q, s, b, xw)) <= 1.0001 * tol
    # a zero right-hand side: converged at once, x stays zero
    x, r, one = solve(Q, M, s, np.zeros(n), tolerance=tol, restart=50)
    assert one and r.converged and r.iterations == 0 and not x.any()
    # a non-finite entry: stops at once, not converged, x untouched
    bad = b.copy()
    bad[n // 3] = np.nan
    x, r, one = solve(Q, M, s, bad, tolerance=tol, restart=50)
    assert one and not r.converged and r.iterations == 0 and not np.isfinite(r.residual)
    assert not x.any()
    # and the next solve is unaffected
    x, r, one = solve(Q, M, s, b, tolerance=tol, restart=50)
    assert one and r.converged and np.linalg.norm(residual(Q, s, b, x)) <= 1.0001 * tol


def test_tiny_not_taken_where_it_does_not_fit():
    """A restart above 64 and a problem above 1024 states take the host-driven loop."""
    wl = config_small()
    box, Q, s, M = setup(wl, mk.FP32, mk.RED_BLACK)
    b = wl.rhs(0, box.size())
    tol = 1e-10 * np.linalg.norm(b)
    x, r, one = solve(Q, M, s, b, tolerance=tol, restart=80, orthogonalization="cgs2")
    assert not one and r.converged
    wl = birth_death_1d(12, 4, 0.7, "random")  # 1,820 states
    box, Q, s, M = setup(wl, mk.FP32, mk.RED_BLACK)
    b = np.random.default_rng(2).random(box.size())
    x, r, one = solve(Q, M, s, b, tolerance=1e-10 * np.linalg.norm(b), restart=50)
    assert not one and r.converged


@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
def test_tiny_batch_matches_single_solves(precision):
    """A batch of full-size systems at the hierarchy's shift runs one launch, a workgroup per
    system: each result is bitwise the single solve of that system (the same kernel code), and
    within the L1 bound of the oracle's and of the batch engine's solution."""
    wl = config_small()
    box, Q, s, M = setup(wl, precision, mk.RED_BLACK)
    n = box.size()
    systems = 40
    B = np.empty((systems, n))
    B[0] = wl.rhs(0, n)
    for j in range(1, systems):
        r = np.random.default_rng(100 + j).random(n)
        B[j] = r / r.sum()
    tols = 1e-10 * np.linalg.norm(B, axis=1)
    batch = mk.BatchSolver(Q, M, [s] * systems)
    mk.profile_reset()
    mk.profile_enable(True)
    try:
        X, res = batch.solve(B, tolerances=tols, restart=50)
        prof = mk.profile_read()
    finally:
        mk.profile_enable(False)
    assert prof["gmres_other"]["launches"] == 1 and prof["shifted_spmv"]["launches"] == 0
    assert all(r.converged for r in res)
    for j in (0, 1, 17, systems - 1):
        x, r, one = solve(Q, M, s, B[j], tolerance=tols[j], restart=50)
        assert one
        assert np.array_equal(x, X[j]), j
        assert r.iterations == res[j].iterations and r.residual == res[j].residual
    try:
        mk.tiny_gmres_enable(False)
        Xe, rese = batch.solve(B, tolerances=tols, restart=50)
    finally:
        mk.tiny_gmres_enable(True)
    assert all(r.converged for r in rese)
    P = Problem(wl.model, wl.max_totals)
    Mo = OracleMG(P, s)
    for j in range(systems):
        rj = B[j] - Q.apply_shifted(s, X[j])
        assert np.linalg.norm(rj) <= 1.0001 * tols[j]
        re = B[j] - Q.apply_shifted(s, Xe[j])
        assert np.abs(X[j] - Xe[j]).sum() <= (np.abs(rj).sum() + np.abs(re).sum()) / s
    for j in (0, 5):
        x_ref, info = P.gmres(s, B[j], mg=Mo, tol=float(tols[j]), restart=50)
        assert info["converged"]
        r_ref = B[j] - P.apply_shifted(s, x_ref)
        rj = B[j] - Q.apply_shifted(s, X[j])
        assert np.abs(X[j] - x_ref).sum() <= (np.abs(rj).sum() + np.abs(r_ref).sum()) / s
