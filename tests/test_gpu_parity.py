"""GPU parity: the HIP library (through its C ABI) against the FP64 CPU oracle on the same
seeded inputs, plus size-independent properties at the full 11.2M-state size.

Tolerances: the shifted-generator SpMV and the FP64 multigrid cycle differ from the oracle
only in summation order (gather by destination vs scatter by source), so they must agree to
64 ulps of the absolute row sums. The FP32 multigrid is compared at 2e-5 relative. Resolvent
solutions must reach the 1e-10 relative residual and agree with the oracle's solution within
the L1 bound ||x - x*||_1 <= ||b - A x||_1 / s (rdme/cpp/rdme_full_multigrid.cpp:884-887) of
both solves.
"""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.workloads import (annihilation_2d, birth_death_1d, config_large,
                                     config_small, lumped_sector_sums)
from oracle.oracle import Problem, SectorMultigrid as OracleMG

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def small_cases():
    return [
        ("1d_v16_cap3", config_small()),
        ("1d_v8_cap2_rand", birth_death_1d(8, 2, 0.7, "random")),
        ("2d_4x4_caps22", annihilation_2d(4, (2, 2), 1.0)),
        ("1d_v32_cap4", birth_death_1d(32, 4, 0.5, "random")),
        ("1d_v12_cap3_odd", birth_death_1d(12, 3, 2.0, "random")),  # 12 -> 6 -> 3 stops coarsening
    ]


@pytest.fixture(scope="module", params=small_cases(), ids=lambda c: c[0])
def case(request):
    name, w = request.param
    print(f"[case {name}] states={w.states}", flush=True)
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    P = Problem(w.model, w.max_totals)
    return dict(name=name, w=w, box=box, Q=Q, P=P)


def abs_rowsum(P, shift, x):
    A = P.matrix(shift, -1.0)
    A.data = np.abs(A.data)
    return A @ np.abs(x)


def test_generator_rates_match(case):
    ex, lk = case["P"].rates()
    assert case["Q"].rows() == case["P"].n
    np.testing.assert_array_equal(case["Q"].exit_rates(), ex)
    np.testing.assert_array_equal(case["Q"].leak_rates(), lk)
    assert case["Q"].n_jumps() == len(case["P"].jumps()[1])


def test_shifted_apply(case):
    P, Q = case["P"], case["Q"]
    rng = np.random.default_rng(7)
    shift = case["w"].shifts[0]
    for trial in range(3):
        x = rng.standard_normal(P.n)
        y = Q.apply_shifted(shift, x)
        y_ref = P.apply_shifted(shift, x)
        scale = abs_rowsum(P, shift, x)
        assert np.all(np.abs(y - y_ref) <= 64 * EPS * scale + 1e-300), case["name"]
    # generator::apply (y = Q x)
    x = rng.random(P.n)
    np.testing.assert_allclose(Q.apply(x), P.apply(x), rtol=0,
                               atol=64 * EPS * abs_rowsum(P, 0.0, x).max())


def test_mass_balance(case):
    """1^T (sI - Q) x = sum_j (s + leak_j) x_j (rdme_full_multigrid.cpp:955-962)."""
    P, Q = case["P"], case["Q"]
    shift = case["w"].shifts[0]
    x = np.random.default_rng(3).random(P.n)
    y = Q.apply_shifted(shift, x)
    lhs = y.sum()
    rhs = ((shift + Q.leak_rates()) * x).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


SMOOTHERS = {mk.JACOBI: "jacobi", mk.AGGREGATE_BLOCK: "aggregate_block", mk.RED_BLACK: "red_black",
             mk.ILU0: "ilu0"}


@pytest.mark.parametrize("precision", [mk.FP64, mk.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("smoother", list(SMOOTHERS), ids=["jacobi", "blocks", "red_black", "ilu0"])
def test_multigrid_cycle(case, precision, smoother):
    w, P, box = case["w"], case["P"], case["box"]
    shift = w.shifts[0]
    M = mk.SectorMultigrid(w.model.domain, box, shift,
                           mk.MultigridOptions(precision=precision, smoother=smoother))
    Mo = OracleMG(P, shift, smoother=SMOOTHERS[smoother])
    assert M.max_levels() == Mo.max_levels()
    assert M.n_sectors() == len(P.sector_begins()) - 1
    rng = np.random.default_rng(11)
    for trial in range(2):
        b = rng.standard_normal(P.n)
        z = M.apply(b)
        z_ref = Mo.apply(b)
        err = np.abs(z - z_ref).max() / np.abs(z_ref).max()
        # FP32 ILU(0): the triangular solves carry rounding along dependency chains of up to
        # total * (voxels - 1) rows (124 at 32 voxels, 4 molecules)
        tol = 1e-12 if precision == mk.FP64 else (1e-4 if smoother == mk.ILU0 else 2e-5)
        assert err <= tol, (case["name"], precision, err)


@pytest.mark.parametrize("precision", [mk.FP64, mk.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("smoother", [mk.JACOBI, mk.RED_BLACK], ids=["jacobi", "red_black"])
@pytest.mark.parametrize("cycles,levels,coarsest", [(2, 0, 4), (2, 1, 4), (2, 2, 4), (3, 1, 4),
                                                    (2, 2, 100), (2, 0, 36), (2, 2, 800)],
                         ids=["W", "W-top1", "W-top2", "3-top1", "W-top2-c100", "W-c36",
                              "W-top2-c800"])
def test_multigrid_cycle_coarse_visits(case, precision, smoother, cycles, levels, coarsest):
    """W-cycles (coarse_cycles visits of every coarse level, or of the top cycle_levels levels
    over a V-cycle below, where the fused tail may take over), with dense bottom blocks of up to
    coarsest states, against the oracle's cycle; blocks of more than 128 states (330 on 32
    voxels with coarsest 800) take the wave-per-row dense kernel"""
    w, P, box = case["w"], case["P"], case["box"]
    shift = w.shifts[0]
    M = mk.SectorMultigrid(w.model.domain, box, shift,
                           mk.MultigridOptions(precision=precision, smoother=smoother,
                                               coarse_cycles=cycles, cycle_levels=levels,
                                               coarsest_size=coarsest))
    Mo = OracleMG(P, shift, smoother=SMOOTHERS[smoother], coarse_cycles=cycles,
                  cycle_levels=levels, coarsest_size=coarsest)
    rng = np.random.default_rng(13)
    b = rng.standard_normal(P.n)
    z = M.apply(b)
    z_ref = Mo.apply(b)
    err = np.abs(z - z_ref).max() / np.abs(z_ref).max()
    assert err <= (1e-12 if precision == mk.FP64 else 2e-5), (case["name"], err)


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_hybrid_cycle_solve(case, levels):
    """FGMRES with the FP32 red-black preconditioner W-cycled on the top levels (dense bottom
    blocks of up to 100 states, the bench's hierarchy) reaches the 1e-10 residual and the
    oracle's answer"""
    w, P, box, Q = case["w"], case["P"], case["box"], case["Q"]
    shift = w.shifts[0]
    b = w.rhs(0, P.n)
    tol = 1e-10 * np.linalg.norm(b)
    M = mk.SectorMultigrid(w.model.domain, box, shift,
                           mk.MultigridOptions(precision=mk.FP32, smoother=mk.RED_BLACK,
                                               coarse_cycles=2, cycle_levels=levels,
                                               coarsest_size=100))
    x = np.zeros(P.n)
    res = mk.gmres(mk.ShiftedGenerator(Q, shift), M, b, x, tolerance=tol, max_iterations=500,
                   restart=50)
    assert res.converged, res
    r = b - P.apply_shifted(shift, x)
    assert np.linalg.norm(r) <= 1.0001 * tol
    x_ref, info = P.gmres(shift, b, mg=OracleMG(P, shift, smoother="red_black", coarse_cycles=2,
                                                cycle_levels=levels, coarsest_size=100),
                          tol=tol, restart=50)
    assert info["converged"]
    bound = (np.abs(r).sum() + np.abs(b - P.apply_shifted(shift, x_ref)).sum()) / shift
    assert np.abs(x - x_ref).sum() <= bound
    assert abs(res.iterations - info["iterations"]) <= 2, (res.iterations, info["iterations"])


@pytest.mark.parametrize("precision", [mk.FP64, mk.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("ortho,restart", [("dcgs2", 50), ("cgs2", 50), ("dcgs2", 6)],
                         ids=["dcgs2", "cgs2", "dcgs2_restart6"])
def test_resolvent_solve(case, precision, ortho, restart):
    w, P, box, Q = case["w"], case["P"], case["box"], case["Q"]
    shift = w.shifts[0]
    b = w.rhs(0, P.n)
    tol = 1e-10 * np.linalg.norm(b)
    A = mk.ShiftedGenerator(Q, shift)
    M = mk.SectorMultigrid(w.model.domain, box, shift, mk.MultigridOptions(precision=precision))
    x = np.zeros(P.n)
    res = mk.gmres(A, M, b, x, tolerance=tol, max_iterations=500, restart=restart,
                   orthogonalization=ortho)
    assert res.converged, res
    r = b - P.apply_shifted(shift, x)
    assert np.linalg.norm(r) <= 1.0001 * tol
    assert abs(res.residual - np.linalg.norm(r)) <= 1e-3 * tol
    Mo = OracleMG(P, shift)
    x_ref, info = P.gmres(shift, b, mg=Mo, tol=tol, max_iterations=500, restart=50)
    assert info["converged"]
    r_ref = b - P.apply_shifted(shift, x_ref)
    bound = (np.abs(r).sum() + np.abs(r_ref).sum()) / shift
    assert np.abs(x - x_ref).sum() <= bound, (np.abs(x - x_ref).sum(), bound)
    print(f"[{case['name']}] {ortho}(m={restart}) gpu its={res.iterations} "
          f"oracle its={info['iterations']} l1={np.abs(x - x_ref).sum():.3e} bound={bound:.3e}")
    if restart == 50:
        # one FGMRES(50) cycle in exact arithmetic: the iteration counts agree up to rounding
        assert abs(res.iterations - info["iterations"]) <= 2, (res.iterations, info["iterations"])


def test_dcgs2_matches_cgs2_iterates(case):
    """Delayed and plain CGS2 span the same Krylov space: after a fixed number of iterations
    (no convergence test) the two FGMRES iterates agree to rounding."""
    w, P, box, Q = case["w"], case["P"], case["box"], case["Q"]
    shift = w.shifts[0]
    b = w.rhs(0, P.n)
    A = mk.ShiftedGenerator(Q, shift)
    M = mk.SectorMultigrid(w.model.domain, box, shift, mk.MultigridOptions(precision=mk.FP64))
    xs = {}
    mk.tiny_gmres_enable(False)  # the host-driven loop, where the two variants differ
    try:
        for ortho in ("dcgs2", "cgs2"):
            x = np.zeros(P.n)
            res = mk.gmres(A, M, b, x, tolerance=1e-300, max_iterations=5, restart=50,
                           orthogonalization=ortho)
            assert res.iterations == 5 and not res.converged
            xs[ortho] = x
    finally:
        mk.tiny_gmres_enable(True)
    scale = np.abs(xs["cgs2"]).sum()
    assert np.abs(xs["dcgs2"] - xs["cgs2"]).sum() <= 1e-9 * scale


@pytest.mark.parametrize("precision", [mk.FP64, mk.FP32], ids=["fp64", "fp32"])
def test_aggregate_block_solve(case, precision):
    """FGMRES preconditioned by the aggregate-block sector multigrid: converges to the oracle's
    answer, in no more iterations than with the Jacobi smoother (the reference's "aggregate
    blocks beat point Jacobi", tests/test_rdme.cpp:207-210, seen on the outer iteration)."""
    w, P, box, Q = case["w"], case["P"], case["box"], case["Q"]
    shift = w.shifts[0]
    b = w.rhs(0, P.n)
    tol = 1e-10 * np.linalg.norm(b)
    A = mk.ShiftedGenerator(Q, shift)
    its = {}
    xs = {}
    for name, sm in (("jacobi", mk.JACOBI), ("blocks", mk.AGGREGATE_BLOCK)):
        M = mk.SectorMultigrid(w.model.domain, box, shift,
                               mk.MultigridOptions(precision=precision, smoother=sm))
        x = np.zeros(P.n)
        res = mk.gmres(A, M, b, x, tolerance=tol, max_iterations=500, restart=50)
        assert res.converged, (name, res)
        its[name] = res.iterations
        xs[name] = x
    r = b - P.apply_shifted(shift, xs["blocks"])
    assert np.linalg.norm(r) <= 1.0001 * tol
    x_ref, info = P.gmres(shift, b, mg=OracleMG(P, shift, smoother="aggregate_block"), tol=tol,
                          restart=50)
    assert info["converged"]
    bound = (np.abs(r).sum() + np.abs(b - P.apply_shifted(shift, x_ref)).sum()) / shift
    assert np.abs(xs["blocks"] - x_ref).sum() <= bound
    assert its["blocks"] <= its["jacobi"], its
    print(f"[{case['name']}] iterations jacobi {its['jacobi']} blocks {its['blocks']} "
          f"oracle blocks {info['iterations']}")


@pytest.mark.parametrize("precision", [mk.FP64, mk.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("smoother", [mk.RED_BLACK, mk.ILU0], ids=["red_black", "ilu0"])
def test_smoother_solve(case, precision, smoother):
    """FGMRES preconditioned by the red-black Gauss-Seidel or ILU(0) sector multigrid reaches
    the 1e-10 residual and the oracle's answer (the smoother changes iterates, not the
    answer), in no more iterations than the oracle's Jacobi-smoothed solve."""
    w, P, box, Q = case["w"], case["P"], case["box"], case["Q"]
    shift = w.shifts[0]
    b = w.rhs(0, P.n)
    tol = 1e-10 * np.linalg.norm(b)
    M = mk.SectorMultigrid(w.model.domain, box, shift,
                           mk.MultigridOptions(precision=precision, smoother=smoother))
    x = np.zeros(P.n)
    res = mk.gmres(mk.ShiftedGenerator(Q, shift), M, b, x, tolerance=tol, max_iterations=500,
                   restart=50)
    assert res.converged, res
    r = b - P.apply_shifted(shift, x)
    assert np.linalg.norm(r) <= 1.0001 * tol
    x_ref, info = P.gmres(shift, b, mg=OracleMG(P, shift, smoother=SMOOTHERS[smoother]), tol=tol,
                          restart=50)
    assert info["converged"]
    bound = (np.abs(r).sum() + np.abs(b - P.apply_shifted(shift, x_ref)).sum()) / shift
    assert np.abs(x - x_ref).sum() <= bound
    _, jac = P.gmres(shift, b, mg=OracleMG(P, shift), tol=tol, restart=50)
    assert res.iterations <= jac["iterations"] + 1, (res.iterations, jac["iterations"])
    print(f"[{case['name']}] {SMOOTHERS[smoother]} iterations {res.iterations} "
          f"(oracle {info['iterations']}, oracle jacobi {jac['iterations']})")


def test_plain_gmres(case):
    w, P, Q = case["w"], case["P"], case["Q"]
    if P.n > 20000:
        pytest.skip("plain GMRES is only checked on the small boxes")
    shift = w.shifts[0]
    b = w.rhs(0, P.n)
    tol = 1e-10 * np.linalg.norm(b)
    x = np.zeros(P.n)
    res = mk.gmres(mk.ShiftedGenerator(Q, shift), b, x, tolerance=tol, max_iterations=2000,
                   restart=100)
    assert res.converged
    r = b - P.apply_shifted(shift, x)
    assert np.linalg.norm(r) <= 1.0001 * tol


def test_nonconvergence_is_reported(case):
    w, Q, box = case["w"], case["Q"], case["box"]
    shift = w.shifts[0]
    b = w.rhs(0, Q.rows())
    x = np.zeros(Q.rows())
    res = mk.gmres(mk.ShiftedGenerator(Q, shift), b, x, tolerance=1e-300, max_iterations=2,
                   restart=50)
    assert not res.converged and res.iterations == 2


def test_large_box_lumped_solution():
    """11,238,513-state 1D box (V=64, cap 5), s=0.5: the FP32-multigrid FGMRES solution's
    sector sums equal the lumped closed form, and the true FP64 residual is <= 1e-10 ||b||."""
    w = config_large()
    print(f"[large] building {w.states} states", flush=True)
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    assert box.size() == 11238513
    Q = mk.Generator(box, w.model)
    shift = w.shifts[0]
    b = w.rhs(0, box.size())
    tol = 1e-10 * np.linalg.norm(b)
    M = mk.SectorMultigrid(w.model.domain, box, shift, mk.MultigridOptions(precision=mk.FP32))
    x = np.zeros(box.size())
    res = mk.gmres(mk.ShiftedGenerator(Q, shift), M, b, x, tolerance=tol, max_iterations=300,
                   restart=50)
    print(f"[large] its={res.iterations} residual={res.residual:.3e}", flush=True)
    assert res.converged
    r = b - Q.apply_shifted(shift, x)
    assert np.linalg.norm(r) <= 1.0001 * tol
    # mass balance on the solution: 1^T b = sum (s + leak) x
    assert abs(b.sum() - ((shift + Q.leak_rates()) * x).sum()) <= 1e-9
    begins = [box.sector_begin(k) for k in range(box.n_sectors())] + [box.size()]
    Bk = np.array([b[begins[k]:begins[k + 1]].sum() for k in range(6)])
    Xk = np.array([x[begins[k]:begins[k + 1]].sum() for k in range(6)])
    X = lumped_sector_sums(shift, 0.05, 0.02, 5, Bk)
    # ||x - x*||_1 <= ||b - A x||_1 / s, with the FP64 evaluation error of b - A x
    # (about eps * || |A| |x| ||_1, where || |A| |x| ||_1 = sum_j (s + 2 exit_j - leak_j) |x_j|)
    absAx = ((shift + 2 * Q.exit_rates() - Q.leak_rates()) * np.abs(x)).sum()
    bound = (np.abs(r).sum() + 16 * EPS * absAx) / shift
    print(f"[large] sector-sum error {np.abs(Xk - X).sum():.3e} bound {bound:.3e}")
    assert np.abs(Xk - X).sum() <= bound


def test_plain_layout_generator():
    """A generator whose rates take more than 256 distinct values is stored in the plain
    layout (u32 columns, FP64 rates); its SpMV and resolvent solve must match too."""
    w = birth_death_1d(8, 3, 0.5, "random")
    P = Problem(w.model, w.max_totals)
    first, to, rate = P.jumps()
    rng = np.random.default_rng(5)
    rate = rate * (1.0 + 0.5 * rng.random(len(rate)))  # distinct rates: plain layout
    ex, lk = P.rates()
    exit_ = lk + np.bincount(np.repeat(np.arange(P.n), np.diff(first)), weights=rate,
                             minlength=P.n)
    Q = mk.Generator.from_jumps(first, to, rate, exit_, lk)
    import scipy.sparse as sp
    src = np.repeat(np.arange(P.n), np.diff(first))
    C = sp.csr_matrix((rate, (to, src)), shape=(P.n, P.n))
    x = rng.standard_normal(P.n)
    y = Q.apply_shifted(0.5, x)
    y_ref = (0.5 + exit_) * x - C @ x
    scale = (0.5 + exit_) * np.abs(x) + abs(C) @ np.abs(x)
    assert np.all(np.abs(y - y_ref) <= 64 * EPS * scale)
    b = w.rhs(0, P.n)
    xs = np.zeros(P.n)
    res = mk.gmres(mk.ShiftedGenerator(Q, 0.5), b, xs, tolerance=1e-10 * np.linalg.norm(b),
                   max_iterations=500, restart=100)
    assert res.converged
    r = b - ((0.5 + exit_) * xs - C @ xs)
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("bad", [np.nan, np.inf], ids=["nan", "inf"])
def test_non_finite_rhs_is_reported(bad):
    """A NaN or infinite entry in b makes the initial residual non-finite: the solve stops at
    once and reports converged = False (non-convergence is not an error in the reference);
    the library stays usable for the next solve."""
    w = config_small()
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    shift = w.shifts[0]
    M = mk.SectorMultigrid(w.model.domain, box, shift, mk.MultigridOptions(precision=mk.FP32))
    b = w.rhs(0, box.size())
    b_bad = b.copy()
    b_bad[box.size() // 2] = bad
    x = np.zeros(box.size())
    res = mk.gmres(mk.ShiftedGenerator(Q, shift), M, b_bad, x, tolerance=1e-10, restart=50)
    assert not res.converged and not np.isfinite(res.residual)
    assert res.iterations == 0
    x = np.zeros(box.size())
    res = mk.gmres(mk.ShiftedGenerator(Q, shift), M, b, x, tolerance=1e-10, restart=50)
    assert res.converged
