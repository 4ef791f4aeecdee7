"""The 2D two-species annihilation resolvent at full size (8 x 8 voxels, h = 1/8, D_A = D_B = 1,
births 0.1 * volume each, A + B -> 0 at n_A n_B / volume, caps {2, 2}: 2145^2 = 4,601,025
states, s = 1, b = e_0), the model family of examples/rdme_two_particle_annihilation.cpp:61-93.

The chain is not lumpable, so at full size the solution is checked through what the system
fixes exactly (tests/test_rdme.cpp:253-297 checks the same kind of identity on 25 states):
- the true FP64 residual ||b - A x||_2 <= 1e-10 ||b||_2;
- mass balance: the column sums of A = sI - Q are s + leak_j, so
  1^T b - sum_j (s + leak_j) x_j = 1^T (b - A x), at most ||b - A x||_1 plus rounding;
- positivity: A is a nonsingular M-matrix and b >= 0, so x* >= 0 and x_j >= -||x - x*||_1;
- the mixed (FP32 multigrid) and plain FP64 solutions agree within the L1 bound
  (||r_mixed||_1 + ||r_fp64||_1 + 16 eps sum_j (s + 2 exit_j - leak_j)(|x_mixed| + |x_fp64|)) / s.
A box on the same 8 x 8 grid small enough for the oracle (caps {2, 1}, 139,425 states) is solved
against the oracle's FP64 solution state by state.
"""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.workloads import annihilation_2d, config_2d
from oracle.oracle import Problem, SectorMultigrid as OracleMG

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def solve(Q, w, box, precision, b, smoother=mk.RED_BLACK):
    shift = w.shifts[0]
    M = mk.SectorMultigrid(w.model.domain, box, shift,
                           mk.MultigridOptions(precision=precision, smoother=smoother))
    x = np.zeros(box.size())
    tol = 1e-10 * np.linalg.norm(b)
    res = mk.gmres(mk.ShiftedGenerator(Q, shift), M, b, x, tolerance=tol, max_iterations=500,
                   restart=50)
    return x, res


def test_c3_full_size():
    w = config_2d()
    print(f"[c3] building {w.states} states", flush=True)
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    assert box.size() == 4601025 == 2145 ** 2
    assert box.n_sectors() == 9
    Q = mk.Generator(box, w.model)
    shift = w.shifts[0]
    b = w.rhs(0, box.size())
    ex, lk = Q.exit_rates(), Q.leak_rates()
    abs_A = shift + 2 * ex - lk
    out = {}
    for name, prec in (("mixed", mk.FP32), ("fp64", mk.FP64)):
        x, res = solve(Q, w, box, prec, b)
        print(f"[c3] {name}: {res.iterations} iterations, residual {res.residual:.3e}", flush=True)
        assert res.converged, res
        r = b - Q.apply_shifted(shift, x)
        assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b)
        rounding = 16 * EPS * (abs_A * np.abs(x)).sum()
        bound = (np.abs(r).sum() + rounding) / shift
        mass = abs(b.sum() - ((shift + lk) * x).sum())
        assert mass <= np.abs(r).sum() + 64 * EPS * ((shift + lk) * np.abs(x)).sum(), mass
        assert x.min() >= -bound, (x.min(), bound)
        out[name] = (x, r)
    (xm, rm), (xd, rd) = out["mixed"], out["fp64"]
    bound = (np.abs(rm).sum() + np.abs(rd).sum() +
             16 * EPS * (abs_A * (np.abs(xm) + np.abs(xd))).sum()) / shift
    l1 = np.abs(xm - xd).sum()
    print(f"[c3] |x_mixed - x_fp64|_1 = {l1:.3e}, bound {bound:.3e}")
    assert l1 <= bound


@pytest.fixture(scope="module")
def grid_8x8():
    """8 x 8 voxels, caps {2, 1} (2145 x 65 = 139,425 states), random L1-normalised b, and the
    oracle's FP64 solution (Jacobi V-cycle, GMRES(50))."""
    w = annihilation_2d(8, (2, 1), 1.0)
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    assert box.size() == 2145 * 65
    Q = mk.Generator(box, w.model)
    P = Problem(w.model, w.max_totals)
    shift = w.shifts[0]
    b = np.random.default_rng(11).random(box.size())
    b /= b.sum()
    tol = 1e-10 * np.linalg.norm(b)
    x_ref, info = P.gmres(shift, b, mg=OracleMG(P, shift), tol=tol, restart=50)
    assert info["converged"]
    return dict(w=w, box=box, Q=Q, P=P, b=b, tol=tol, x_ref=x_ref, info=info)


@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("smoother", [mk.RED_BLACK, mk.JACOBI], ids=["red_black", "jacobi"])
def test_8x8_grid_against_oracle(grid_8x8, precision, smoother):
    """The C3 grid and hierarchy (8 -> 4 -> 2 -> 1 voxels per axis) on a box the oracle solves
    in seconds: the GPU solution against the oracle's state by state."""
    g = grid_8x8
    w, P, b, tol, x_ref = g["w"], g["P"], g["b"], g["tol"], g["x_ref"]
    shift = w.shifts[0]
    x, res = solve(g["Q"], w, g["box"], precision, b, smoother)
    assert res.converged
    r = b - P.apply_shifted(shift, x)
    r_ref = b - P.apply_shifted(shift, x_ref)
    assert np.linalg.norm(r) <= 1.0001 * tol
    ex, lk = P.rates()
    rounding = 16 * EPS * ((shift + 2 * ex - lk) * (np.abs(x) + np.abs(x_ref))).sum()
    bound = (np.abs(r).sum() + np.abs(r_ref).sum() + rounding) / shift
    l1 = np.abs(x - x_ref).sum()
    print(f"[8x8 caps21] {res.iterations} iterations (oracle {g['info']['iterations']}), "
          f"|x - x_oracle|_1 = {l1:.3e} <= {bound:.3e}")
    assert l1 <= bound
