"""Pins of the oracle's red-black Gauss-Seidel and ILU(0) smoothers (the paper driver's
smoothers, rdme/cpp/rdme_full_multigrid.cpp:178-195 and :204-296, placed in
population_multigrid's V-cycle).

* colours: the oracle's colour of every state equals the parity of the sum over molecules of
  their voxels' coordinate sums, computed here from the state list, and every off-diagonal
  entry of the level operator joins the two colours (the graph is bipartite);
* red-black Gauss-Seidel is textbook lexicographic Gauss-Seidel on the dense matrix in the
  order "colour 0 states ascending, then colour 1 states" (reversed colour order for
  post-smoothing), one and two sweeps. A simultaneous (Jacobi-like) update of both colours, a
  wrong colour order or a relaxation weight fails this;
* ILU(0): the pivots equal the diagonal of U from a textbook IKJ ILU(0) on the dense matrix
  restricted to its pattern (Saad, Iterative Methods, Alg. 10.4), whose off-diagonal factors
  then equal A's own entries as the bipartite argument says; one smoothing sweep equals
  x + (L U)^-1 (b - A x) by dense triangular solves;
* every smoother preconditions GMRES to the same answer as a direct sparse solve.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from markovkit_amd.model import Grid, ReactionDiffusion, VoxelDomain
from markovkit_amd.workloads import annihilation_2d, birth_death_1d
from oracle.oracle import Problem, SectorMultigrid


def diffusion(dims, h, species=1):
    return ReactionDiffusion(VoxelDomain(Grid(list(dims), [h] * len(dims)), [1.0] * species), [])


CASES = [((8,), 1.0 / 8, [3], 3),     # 1D, sector of 3 molecules: 120 states
         ((4, 4), 0.25, [1, 1], 3),   # 2D, two species, one molecule each: 256 states
         ((4, 2), 0.5, [2], 2)]       # 2D 4 x 2, two molecules: 36 states


def sector(case):
    dims, h, caps, k = case
    P = Problem(diffusion(dims, h, len(caps)), caps)
    b = P.sector_begins()
    return P, slice(int(b[k]), int(b[k + 1])), k, dims


def colour_formula(states, dims):
    # states: rows of V * S counts (voxel-major); voxel v = c0 + d0 * c1 (axis 0 fastest)
    V = int(np.prod(dims))
    S = states.shape[1] // V
    coords = np.array(np.unravel_index(np.arange(V), dims, order="F")).sum(axis=0)
    counts = states.reshape(len(states), V, S).sum(axis=2)
    return ((counts * coords[None, :]).sum(axis=1) & 1).astype(np.uint8)


@pytest.mark.parametrize("case", CASES, ids=["1d", "2d-two-species", "2d-4x2"])
def test_colours_split_every_jump(case):
    P, sl, k, dims = sector(case)
    M = SectorMultigrid(P, 0.7, smoother="red_black")
    col = M.colours(k)
    assert np.array_equal(col, colour_formula(P.states()[sl], dims))
    A = P.matrix(0.7, -1.0).toarray()[sl, sl]
    i, j = np.nonzero(A - np.diag(np.diag(A)))
    assert len(i) > 0 and np.all(col[i] != col[j])
    assert 0 < col.sum() < len(col)


def lexicographic_gs(A, b, x, order):
    x = x.copy()
    for i in order:
        x[i] = (b[i] - A[i] @ x + A[i, i] * x[i]) / A[i, i]
    return x


@pytest.mark.parametrize("case", CASES, ids=["1d", "2d-two-species", "2d-4x2"])
@pytest.mark.parametrize("reverse", [False, True])
def test_red_black_is_lexicographic_gs_in_colour_order(case, reverse):
    P, sl, k, _ = sector(case)
    s = 0.7
    M = SectorMultigrid(P, s, smoother="red_black")
    col = M.colours(k)
    A = P.matrix(s, -1.0).toarray()[sl, sl]
    rng = np.random.default_rng(3)
    b = rng.random(len(col))
    x0 = rng.random(len(col))
    first, second = (1, 0) if reverse else (0, 1)
    order = np.concatenate([np.nonzero(col == first)[0], np.nonzero(col == second)[0]])
    for sweeps in (1, 2):
        want = x0
        for _ in range(sweeps):
            want = lexicographic_gs(A, b, want, order)
        got = M.relax(k, b, x0.copy(), sweeps=sweeps, reverse=reverse)
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    # a plausible mistake fails: both colours updated from the same old iterate
    d = np.diag(A)
    jac = x0 + (b - A @ x0) / d
    assert np.abs(jac - lexicographic_gs(A, b, x0, order)).max() > 1e-6


def ilu0_dense(A):
    """textbook IKJ ILU(0) on the pattern of A; returns unit-lower L and U"""
    n = len(A)
    W = A.copy()
    nz = A != 0
    for i in range(1, n):
        for kk in range(i):
            if not nz[i, kk]:
                continue
            W[i, kk] /= W[kk, kk]
            for j in range(kk + 1, n):
                if nz[i, j]:
                    W[i, j] -= W[i, kk] * W[kk, j]
    return np.tril(W, -1) + np.eye(n), np.triu(W)


@pytest.mark.parametrize("case", CASES, ids=["1d", "2d-two-species", "2d-4x2"])
def test_ilu0_matches_textbook_factors(case):
    P, sl, k, _ = sector(case)
    s = 0.7
    M = SectorMultigrid(P, s, smoother="ilu0")
    A = P.matrix(s, -1.0).toarray()[sl, sl]
    L, U = ilu0_dense(A)
    piv = M.ilu_pivots(k)
    assert np.allclose(piv, np.diag(U), rtol=1e-13, atol=0)
    # bipartite: the off-diagonal factors are A's own entries
    assert np.allclose(np.triu(U, 1), np.triu(A, 1), rtol=0, atol=0)
    assert np.allclose(np.tril(L, -1) * np.diag(U)[None, :], np.tril(A, -1), rtol=1e-13, atol=1e-300)
    rng = np.random.default_rng(4)
    b = rng.random(len(piv))
    x0 = rng.random(len(piv))
    r = b - A @ x0
    e = sla.solve_triangular(U, sla.solve_triangular(L, r, lower=True, unit_diagonal=True))
    for sweeps in (1, 2):
        want = x0 + e if sweeps == 1 else None
        if sweeps == 2:
            x1 = x0 + e
            r1 = b - A @ x1
            want = x1 + sla.solve_triangular(U, sla.solve_triangular(L, r1, lower=True,
                                                                    unit_diagonal=True))
        got = M.relax(k, b, x0.copy(), sweeps=sweeps)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # a plausible mistake fails: the plain diagonal instead of the ILU pivots
    wrong = sla.solve_triangular(np.triu(A), sla.solve_triangular(
        np.tril(A, -1) / np.diag(A)[None, :] + np.eye(len(A)), r, lower=True, unit_diagonal=True))
    assert np.abs(x0 + wrong - (x0 + e)).max() > 1e-8


@pytest.mark.parametrize("wl", [birth_death_1d(16, 3, 0.5, "random"), annihilation_2d(4, (1, 1), 1.0)],
                         ids=["1d", "2d"])
def test_every_smoother_gives_the_direct_answer(wl):
    P = Problem(wl.model, wl.max_totals)
    s = wl.shifts[0]
    b = wl.rhs(0, P.n)
    exact = spla.spsolve(P.matrix(s, -1.0).tocsc(), b)
    its = {}
    for sm in SectorMultigrid.SMOOTHERS:
        M = SectorMultigrid(P, s, smoother=sm)
        x, info = P.gmres(s, b, mg=M, tol=1e-12 * np.linalg.norm(b), restart=50)
        assert info["converged"], sm
        assert np.abs(x - exact).sum() <= 1e-9 * np.abs(exact).sum(), sm
        its[sm] = info["iterations"]
    # the stronger smoothers need no more outer iterations than weighted Jacobi here
    assert its["red_black"] <= its["jacobi"] and its["ilu0"] <= its["jacobi"], its


def test_cycle_levels_semantics():
    """cycle_levels: 0 and any count at or beyond the hierarchy's depth give the full W-cycle,
    one coarse visit gives the V-cycle whatever cycle_levels, and on a two-level hierarchy
    cycle_levels = 1 is the full W-cycle (the oracle cycle's results, bitwise)"""
    from oracle.oracle import Problem, SectorMultigrid as OracleMG
    from markovkit_amd.workloads import birth_death_1d
    w = birth_death_1d(16, 3, 0.5, "random")
    P = Problem(w.model, w.max_totals)
    s = w.shifts[0]
    b = np.random.default_rng(2).standard_normal(P.n)
    for sm in ("jacobi", "red_black"):
        full = OracleMG(P, s, smoother=sm, coarse_cycles=2).apply(b)
        assert np.array_equal(OracleMG(P, s, smoother=sm, coarse_cycles=2, cycle_levels=32).apply(b), full)
        v = OracleMG(P, s, smoother=sm).apply(b)
        for k in (1, 2, 5):
            assert np.array_equal(OracleMG(P, s, smoother=sm, cycle_levels=k).apply(b), v)
        two = OracleMG(P, s, smoother=sm, coarse_cycles=2, max_levels=2)
        assert np.array_equal(OracleMG(P, s, smoother=sm, coarse_cycles=2, max_levels=2,
                                       cycle_levels=1).apply(b), two.apply(b))
        top1 = OracleMG(P, s, smoother=sm, coarse_cycles=2, cycle_levels=1).apply(b)
        assert not np.array_equal(top1, full) and not np.array_equal(top1, v)
