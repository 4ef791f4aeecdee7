"""GPU parity of batched resolvent solves (mk_batch_*): every system of a batch against the
single-system solver and the FP64 CPU oracle on the same inputs, plus the lumped closed form of
the sector sums for 1D linear birth-death boxes.

Tolerances: every system must reach its own absolute tolerance on the true FP64 residual, and
two solutions x, x' of the same system agree within the L1 bound
||x - x'||_1 <= (||b - A x||_1 + ||b - A x'||_1) / s (rdme/cpp/rdme_full_multigrid.cpp:884-887).
"""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.workloads import (annihilation_2d, birth_death_1d, config_multishift,
                                     lumped_sector_sums)
from oracle.oracle import Problem, SectorMultigrid as OracleMG

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def build(w):
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    return box, Q


def rhs_set(n, seeds):
    out = []
    for s in seeds:
        b = np.random.default_rng(s).random(n)
        out.append(b / b.sum())
    return out


def l1_residual(Q, s, b, x):
    return np.abs(b - Q.apply_shifted(s, x)).sum()


@pytest.fixture(scope="module")
def small():
    w = birth_death_1d(16, 3, 0.5, "random")
    box, Q = build(w)
    return w, box, Q


@pytest.mark.parametrize("precision", [mk.FP64, mk.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("restart", [50, 5])
def test_batch_matches_single_solves(small, precision, restart):
    w, box, Q = small
    n = box.size()
    shifts = [0.05, 0.5, 5.0, 50.0, 0.05, 0.5, 5.0, 50.0]
    B = np.array(rhs_set(n, [0, 0, 0, 0, 1, 1, 1, 1]))
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, 0.5, mk.MultigridOptions(precision=precision))
    batch = mk.BatchSolver(Q, M, shifts)
    X, res = batch.solve(B, tolerances=tol, restart=restart, max_iterations=500)
    for j, s in enumerate(shifts):
        assert res[j].converged, (j, res[j])
        r = B[j] - Q.apply_shifted(s, X[j])
        assert np.linalg.norm(r) <= 1.0001 * tol[j], (j, np.linalg.norm(r), tol[j])
        assert abs(res[j].residual - np.linalg.norm(r)) <= 1e-3 * tol[j]
        Ms = mk.SectorMultigrid(w.model.domain, box, s, mk.MultigridOptions(precision=precision))
        x1 = np.zeros(n)
        r1 = mk.gmres(mk.ShiftedGenerator(Q, s), Ms, B[j], x1, tolerance=tol[j], restart=restart,
                      max_iterations=500)
        assert r1.converged
        bound = (l1_residual(Q, s, B[j], X[j]) + l1_residual(Q, s, B[j], x1)) / s
        assert np.abs(X[j] - x1).sum() <= bound, (j, np.abs(X[j] - x1).sum(), bound)
        if precision == mk.FP64:
            # same algorithm and preconditioner: the iteration counts agree up to rounding
            assert abs(res[j].iterations - r1.iterations) <= 1, (j, res[j].iterations, r1.iterations)
        print(f"[batch m={restart}] s={s} its={res[j].iterations} single its={r1.iterations}")


@pytest.mark.parametrize("cycles,levels", [(2, 1), (2, 0)], ids=["W-top1", "W"])
def test_batch_w_cycle_levels(small, cycles, levels):
    """A batch takes its hierarchy's cycle shape: W-cycles on every level, or on the top
    cycle_levels only. Three FGMRES iterations from zero (not converged) give the single-system
    solver's iterate up to rounding, which a different cycle shape would not."""
    w, box, Q = small
    n = box.size()
    shifts = [0.5, 5.0]
    B = np.array(rhs_set(n, [2, 3]))
    opts = dict(precision=mk.FP64, coarse_cycles=cycles, cycle_levels=levels)
    M = mk.SectorMultigrid(w.model.domain, box, 0.5, mk.MultigridOptions(**opts))
    assert M.max_levels() >= 3
    X, res = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=[1e-30] * 2, restart=50,
                                                max_iterations=3)
    for j, s in enumerate(shifts):
        assert res[j].iterations == 3 and not res[j].converged, res[j]
        Ms = mk.SectorMultigrid(w.model.domain, box, s, mk.MultigridOptions(**opts))
        x1 = np.zeros(n)
        mk.gmres(mk.ShiftedGenerator(Q, s), Ms, B[j], x1, tolerance=1e-30, restart=50,
                 max_iterations=3)
        err = np.abs(X[j] - x1).max() / np.abs(x1).max()
        assert err <= 1e-9, (j, err)


def test_batch_against_oracle_and_lumped(small):
    w, box, Q = small
    n = box.size()
    shifts = [0.1, 1.0, 10.0]
    B = np.array(rhs_set(n, [3, 4, 5]))
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, 1.0, mk.MultigridOptions(precision=mk.FP32))
    X, res = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=tol)
    P = Problem(w.model, w.max_totals)
    begins = [box.sector_begin(k) for k in range(box.n_sectors())] + [n]
    for j, s in enumerate(shifts):
        assert res[j].converged
        x_ref, info = P.gmres(s, B[j], mg=OracleMG(P, s), tol=tol[j], restart=50)
        assert info["converged"]
        r = B[j] - P.apply_shifted(s, X[j])
        r_ref = B[j] - P.apply_shifted(s, x_ref)
        bound = (np.abs(r).sum() + np.abs(r_ref).sum()) / s
        assert np.abs(X[j] - x_ref).sum() <= bound
        Bk = np.array([B[j][begins[k]:begins[k + 1]].sum() for k in range(4)])
        Xk = np.array([X[j][begins[k]:begins[k + 1]].sum() for k in range(4)])
        X_closed = lumped_sector_sums(s, 0.05, 0.02, 3, Bk)
        assert np.abs(Xk - X_closed).sum() <= np.abs(r).sum() / s + 1e-15


def test_batch_more_than_one_wave_of_systems(small):
    """70 systems: two 64-lane groups, the second padded; repeated solves reuse the workspace
    and a second call with other right-hand sides gives the same answers as a fresh batch."""
    w, box, Q = small
    n = box.size()
    shifts = list(np.geomspace(0.02, 20.0, 70))
    B = np.array(rhs_set(n, range(70)))
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, 1.0, mk.MultigridOptions(precision=mk.FP32))
    batch = mk.BatchSolver(Q, M, shifts)
    X1, res1 = batch.solve(B[::-1].copy(), tolerances=tol[::-1].copy())
    X, res = batch.solve(B, tolerances=tol)
    Xf, _ = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=tol)
    assert np.array_equal(X, Xf)
    for j, s in enumerate(shifts):
        assert res[j].converged and res1[j].converged
        assert np.linalg.norm(B[j] - Q.apply_shifted(s, X[j])) <= 1.0001 * tol[j]
        assert np.linalg.norm(B[::-1][j] - Q.apply_shifted(s, X1[j])) <= 1.0001 * tol[::-1][j]


def test_batch_converged_guess_and_zero_iterations(small):
    w, box, Q = small
    n = box.size()
    shifts = [0.5, 2.0]
    B = np.array(rhs_set(n, [7, 8]))
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, 0.5, mk.MultigridOptions(precision=mk.FP64))
    batch = mk.BatchSolver(Q, M, shifts)
    X, res = batch.solve(B, tolerances=tol)
    X2 = X.copy()
    _, res2 = batch.solve(B, X2, tolerances=tol)
    assert all(r.converged and r.iterations == 0 for r in res2)
    assert np.array_equal(X, X2)
    _, res3 = batch.solve(B, np.zeros_like(B), tolerances=tol, max_iterations=3)
    assert all(not r.converged and r.iterations == 3 for r in res3)


def test_batch_two_species_2d():
    w = annihilation_2d(4, (2, 2), 1.0)
    box, Q = build(w)
    n = box.size()
    shifts = [0.3, 1.0, 3.0]
    B = np.array(rhs_set(n, [0, 1, 2]))
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, 1.0, mk.MultigridOptions(precision=mk.FP32))
    X, res = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=tol)
    for j, s in enumerate(shifts):
        assert res[j].converged
        assert np.linalg.norm(B[j] - Q.apply_shifted(s, X[j])) <= 1.0001 * tol[j]
        Ms = mk.SectorMultigrid(w.model.domain, box, s, mk.MultigridOptions(precision=mk.FP32))
        x1 = np.zeros(n)
        assert mk.gmres(mk.ShiftedGenerator(Q, s), Ms, B[j], x1, tolerance=tol[j]).converged
        bound = (l1_residual(Q, s, B[j], X[j]) + l1_residual(Q, s, B[j], x1)) / s
        assert np.abs(X[j] - x1).sum() <= bound


def test_multishift_256_solves():
    """1D, 32 voxels, cap 4 (58,905 states): 64 shifts 10^(-2 + 4 i / 63) x 4 random rhs, one
    batch of 256 systems with the FP32 preconditioner; every system reaches 1e-10 relative
    residual and its sector sums match the lumped closed form."""
    w = config_multishift()
    box, Q = build(w)
    n = box.size()
    assert n == 58905
    shifts, B = [], []
    rhs = rhs_set(n, w.seeds)
    for s in w.shifts:
        for b in rhs:
            shifts.append(s)
            B.append(b)
    B = np.array(B)
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, 1.0, mk.MultigridOptions(precision=mk.FP32))
    X, res = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=tol)
    begins = [box.sector_begin(k) for k in range(box.n_sectors())] + [n]
    absA = 2 * Q.exit_rates() - Q.leak_rates()
    its = [r.iterations for r in res]
    print(f"[multishift] iterations min {min(its)} max {max(its)}")
    for j, s in enumerate(shifts):
        assert res[j].converged, (j, res[j])
        r = B[j] - Q.apply_shifted(s, X[j])
        assert np.linalg.norm(r) <= 1.0001 * tol[j]
        Bk = np.array([B[j][begins[k]:begins[k + 1]].sum() for k in range(5)])
        Xk = np.array([X[j][begins[k]:begins[k + 1]].sum() for k in range(5)])
        X_closed = lumped_sector_sums(s, 0.05, 0.02, 4, Bk)
        # ||x - x*||_1 <= ||b - A x||_1 / s, plus the FP64 evaluation error of b - A x
        # (about eps * sum_j (s + 2 exit_j - leak_j) |x_j|)
        bound = (np.abs(r).sum() + 16 * EPS * ((s + absA) * np.abs(X[j])).sum()) / s
        assert np.abs(Xk - X_closed).sum() <= bound, (j, np.abs(Xk - X_closed).sum(), bound)


@pytest.mark.parametrize("precision", [mk.FP64, mk.FP32], ids=["fp64", "fp32"])
def test_batch_red_black(small, precision):
    """Red-black Gauss-Seidel hierarchies in a batch (1 + 2 sweeps, W-cycles on the top level):
    every system reaches its tolerance and lies within the L1 bound of the single-system
    solver's solution with the same hierarchy shape and of the oracle's (Jacobi) solution; FP64
    iteration counts equal the single-system ones up to rounding."""
    w, box, Q = small
    n = box.size()
    shifts = [0.05, 0.5, 5.0, 50.0]
    B = np.array(rhs_set(n, [10, 11, 12, 13]))
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    opts = dict(precision=precision, smoother=mk.RED_BLACK, pre_sweeps=1, post_sweeps=2,
                coarse_cycles=2, cycle_levels=1)
    M = mk.SectorMultigrid(w.model.domain, box, 0.5, mk.MultigridOptions(**opts))
    X, res = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=tol)
    P = Problem(w.model, w.max_totals)
    for j, s in enumerate(shifts):
        assert res[j].converged, (j, res[j])
        assert np.linalg.norm(B[j] - Q.apply_shifted(s, X[j])) <= 1.0001 * tol[j]
        Ms = mk.SectorMultigrid(w.model.domain, box, s, mk.MultigridOptions(**opts))
        x1 = np.zeros(n)
        r1 = mk.gmres(mk.ShiftedGenerator(Q, s), Ms, B[j], x1, tolerance=tol[j])
        assert r1.converged
        bound = (l1_residual(Q, s, B[j], X[j]) + l1_residual(Q, s, B[j], x1)) / s
        assert np.abs(X[j] - x1).sum() <= bound
        if precision == mk.FP64:
            assert abs(res[j].iterations - r1.iterations) <= 1, (j, res[j].iterations, r1.iterations)
        x_ref, _ = P.gmres(s, B[j], mg=OracleMG(P, s), tol=tol[j], restart=50)
        bound = (np.abs(B[j] - P.apply_shifted(s, X[j])).sum() +
                 np.abs(B[j] - P.apply_shifted(s, x_ref)).sum()) / s
        assert np.abs(X[j] - x_ref).sum() <= bound
        print(f"[batch red-black] s={s} its={res[j].iterations} single {r1.iterations}")


@pytest.mark.parametrize("smoother", [mk.JACOBI, mk.RED_BLACK], ids=["jacobi", "red_black"])
def test_batch_uneven_sizes_in_one_launch(smoother):
    """Systems of the boxes of caps 3, 4 and 5 on 16 voxels (969, 4,845 and 20,349 states) in
    one batch over the cap-5 box: system j is the leading rows[j] block, which for the absorbing
    generator is the cap-k box's generator. Each solution equals (within the L1 bound) the
    single-system solve on its own box, its sector sums the lumped closed form of cap k, and
    its entries past its rows are zero."""
    big = birth_death_1d(16, 5, 0.5, "random")
    box, Q = build(big)
    caps = [3, 4, 5, 3, 5, 4]
    shifts = [0.3, 0.3, 0.3, 4.0, 4.0, 1.0]
    rows = [mk.PopulationBox(16, [c]).size() for c in caps]
    assert rows[:3] == [969, 4845, 20349]
    n = box.size()
    B = np.zeros((len(caps), n))
    for j, r in enumerate(rows):
        b = np.random.default_rng(20 + j).random(r)
        B[j, :r] = b / b.sum()
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    opts = mk.MultigridOptions(precision=mk.FP32, smoother=smoother)
    M = mk.SectorMultigrid(big.model.domain, box, 1.0, opts)
    X, res = mk.BatchSolver(Q, M, shifts, rows=rows).solve(B, tolerances=tol)
    for j, (c, s, r) in enumerate(zip(caps, shifts, rows)):
        assert res[j].converged, (j, res[j])
        assert not np.any(X[j, r:]), "entries past the system's rows stay zero"
        w = birth_death_1d(16, c, s, "random")
        bj, Qj = build(w)
        assert bj.size() == r
        x = X[j, :r]
        rj = B[j, :r] - Qj.apply_shifted(s, x)
        assert np.linalg.norm(rj) <= 1.0001 * tol[j], (j, np.linalg.norm(rj), tol[j])
        Mj = mk.SectorMultigrid(w.model.domain, bj, s, opts)
        x1 = np.zeros(r)
        assert mk.gmres(mk.ShiftedGenerator(Qj, s), Mj, B[j, :r], x1, tolerance=tol[j]).converged
        bound = (np.abs(rj).sum() + l1_residual(Qj, s, B[j, :r], x1)) / s
        assert np.abs(x - x1).sum() <= bound
        begins = [bj.sector_begin(k) for k in range(c + 1)] + [r]
        Bk = np.array([B[j][begins[k]:begins[k + 1]].sum() for k in range(c + 1)])
        Xk = np.array([x[begins[k]:begins[k + 1]].sum() for k in range(c + 1)])
        absA = s + 2 * Qj.exit_rates() - Qj.leak_rates()
        bound = (np.abs(rj).sum() + 16 * EPS * (absA * np.abs(x)).sum()) / s
        assert np.abs(Xk - lumped_sector_sums(s, 0.05, 0.02, c, Bk)).sum() <= bound
        print(f"[batch uneven] cap {c} rows {r} s={s} its={res[j].iterations}")


def test_batch_graphs_and_mix_are_bitwise(small):
    """A system's result does not depend on graph replay or on the other systems of its batch:
    the same system alone, in a mix of 70 systems at another slot, and with eager launches
    gives the same solution and iteration count bit for bit."""
    w, box, Q = small
    n = box.size()
    M = mk.SectorMultigrid(w.model.domain, box, 1.0,
                           mk.MultigridOptions(precision=mk.FP32, smoother=mk.RED_BLACK))
    shifts = list(np.geomspace(0.05, 20.0, 70))
    B = np.array(rhs_set(n, range(70)))
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    X, res = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=tol)
    for j in (0, 33, 69):
        Xa, ra = mk.BatchSolver(Q, M, [shifts[j]]).solve(B[j:j + 1].copy(), tolerances=tol[j:j + 1].copy())
        assert np.array_equal(Xa[0], X[j]) and ra[0].iterations == res[j].iterations, j
    try:
        mk.batch_graphs_enable(False)
        Xe, re = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=tol)
    finally:
        mk.batch_graphs_enable(True)
    assert np.array_equal(Xe, X)
    assert [r.iterations for r in re] == [r.iterations for r in res]


@pytest.mark.parametrize("smoother", [mk.JACOBI, mk.RED_BLACK], ids=["jacobi", "red_black"])
def test_block_krylov_groups(small, smoother):
    """Block FGMRES (groups of 4 right-hand sides per shift sharing a basis, block Gram-Schmidt
    on matrix cores): every system reaches its tolerance, lies within the L1 bound of the
    per-system batch's solution and of the oracle's, and its sector sums match the lumped closed
    form; a group takes no more block steps than its slowest system takes per-system
    iterations."""
    w, box, Q = small
    n = box.size()
    group_shifts = [0.05, 0.5, 5.0, 50.0, 0.3]
    shifts = [s for s in group_shifts for _ in range(4)]
    rhs = rhs_set(n, [30, 31, 32, 33])
    B = np.array([b for _ in group_shifts for b in rhs])
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, 1.0,
                           mk.MultigridOptions(precision=mk.FP32, smoother=smoother))
    Xb, rb = mk.BatchSolver(Q, M, shifts, block=4).solve(B, tolerances=tol)
    Xs, rs = mk.BatchSolver(Q, M, shifts).solve(B, tolerances=tol)
    P = Problem(w.model, w.max_totals)
    begins = [box.sector_begin(k) for k in range(box.n_sectors())] + [n]
    absA = 2 * Q.exit_rates() - Q.leak_rates()
    for j, s in enumerate(shifts):
        assert rb[j].converged, (j, rb[j])
        r = B[j] - Q.apply_shifted(s, Xb[j])
        assert np.linalg.norm(r) <= 1.0001 * tol[j]
        bound = (np.abs(r).sum() + l1_residual(Q, s, B[j], Xs[j])) / s
        assert np.abs(Xb[j] - Xs[j]).sum() <= bound
        Bk = np.array([B[j][begins[k]:begins[k + 1]].sum() for k in range(4)])
        Xk = np.array([Xb[j][begins[k]:begins[k + 1]].sum() for k in range(4)])
        lb = (np.abs(r).sum() + 16 * EPS * ((s + absA) * np.abs(Xb[j])).sum()) / s
        assert np.abs(Xk - lumped_sector_sums(s, 0.05, 0.02, 3, Bk)).sum() <= lb
    for g in range(len(group_shifts)):
        blk = [rb[4 * g + q].iterations for q in range(4)]
        per = [rs[4 * g + q].iterations for q in range(4)]
        assert max(blk) <= max(per), (g, blk, per)
        print(f"[block] s={group_shifts[g]} block steps {blk} per-system iterations {per}")
    x_ref, _ = P.gmres(shifts[0], B[0], mg=OracleMG(P, shifts[0]), tol=tol[0], restart=50)
    bound = (np.abs(B[0] - P.apply_shifted(shifts[0], Xb[0])).sum() +
             np.abs(B[0] - P.apply_shifted(shifts[0], x_ref)).sum()) / shifts[0]
    assert np.abs(Xb[0] - x_ref).sum() <= bound


def test_block_krylov_multishift_256():
    """The 256-system multi-shift configuration (64 shifts x 4 rhs, 58,905 states) with block
    groups: every system reaches 1e-10 relative residual."""
    w = config_multishift()
    box, Q = build(w)
    n = box.size()
    rhs = rhs_set(n, w.seeds)
    shifts = [s for s in w.shifts for _ in rhs]
    B = np.array([b for _ in w.shifts for b in rhs])
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, 1.0, mk.MultigridOptions(precision=mk.FP32))
    X, res = mk.BatchSolver(Q, M, shifts, block=4).solve(B, tolerances=tol)
    its = [r.iterations for r in res]
    print(f"[block multishift] iterations min {min(its)} max {max(its)}")
    for j, s in enumerate(shifts):
        assert res[j].converged, (j, res[j])
        assert np.linalg.norm(B[j] - Q.apply_shifted(s, X[j])) <= 1.0001 * tol[j]
