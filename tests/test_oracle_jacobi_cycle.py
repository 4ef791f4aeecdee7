"""Value pin of the oracle's default preconditioner: the weighted-Jacobi sector multigrid cycle
(include/markovkit/rdme/multigrid.hpp:201-228 cycle, :309-317 Jacobi smoothing, :145-178 the
hierarchy, :47-93 the transfer) against a textbook V- and W-cycle written here in numpy from
independent parts:
- each level's operator A_l = s I - Q_l is the dense diffusion generator of the brute-force
  enumeration (tests/bruteforce.py: itertools states, an explicit move list and a dictionary
  lookup) on the sector, on the grid with every axis longer than one cell halved and the
  diffusivity doubled;
- restriction R sums the fine states into the coarse state with the same counts per coarse
  voxel; prolongation P splits it by the multinomial probability
  prod_a n_a! / prod_{v in a} n_v! * children_a^(-n_a), computed in exact integer arithmetic;
- the cycle: x = 0, pre-sweeps x += w (b - A x) / diag(A) with w = 2/3, the residual, coarse
  cycles x_c += cycle(R r - A_c x_c) from x_c = 0 (one for V, two for W), x += P x_c,
  post-sweeps; on the last level (at most coarsest_size states, or an axis that cannot halve)
  the coarse correction is the dense solve of the residual.
The oracle's cycle must equal it to 1e-13 relative; the same textbook cycle with w = 1, with no
post-smoothing, or with one coarse cycle in place of two must differ by far more, so each of
those mistakes in the oracle would fail this test.
"""
import itertools
import math

import numpy as np
import pytest

from markovkit_amd.model import Grid, ReactionDiffusion, VoxelDomain, birth_death
from oracle.oracle import Problem, SectorMultigrid
from tests import bruteforce


def sector_levels(dims, spacing, D, totals, coarsest_size=4, max_levels=32):
    """[(A_l without shift, coarse_of_l, weight_l)] of one sector, finest first; the last level
    has no transfer (coarse_of None)."""
    S = len(totals)
    levels = []
    while True:
        grid = Grid(dims, spacing)
        model = ReactionDiffusion(VoxelDomain(grid, list(D)), [])
        Qd, _, _, states = bruteforce.dense_generator(model, list(totals))
        V = grid.n_vertices()
        keep = [i for i, st in enumerate(states)
                if tuple(np.array(st).reshape(V, S).sum(axis=0)) == tuple(totals)]
        Q = Qd[np.ix_(keep, keep)]
        sec = [states[i] for i in keep]
        can = all(d == 1 or d % 2 == 0 for d in dims) and any(d > 1 for d in dims)
        if len(sec) <= coarsest_size or len(levels) + 1 >= max_levels or not can:
            levels.append((-Q, None, None))
            return levels
        cdims = [d // 2 if d > 1 else 1 for d in dims]
        cgrid = Grid(cdims, [h * 2 if d > 1 else h for h, d in zip(spacing, dims)])
        parent = []
        for v in range(V):
            c = [a // 2 if d > 1 else a for a, d in zip(grid.coordinate(v), dims)]
            p, stride = 0, 1
            for a, d in zip(c, cdims):
                p += a * stride
                stride *= d
            parent.append(p)
        children = np.bincount(parent, minlength=cgrid.n_vertices())
        coarse_of, weight = [], []
        for st in sec:
            x = np.array(st).reshape(V, S)
            c = np.zeros((cgrid.n_vertices(), S), np.int64)
            for v in range(V):
                c[parent[v]] += x[v]
            key = tuple(c.ravel())
            coarse_of.append(key)
            num, den = 1, 1
            for a in range(cgrid.n_vertices()):
                for s in range(S):
                    num *= math.factorial(int(c[a, s]))
                    den *= int(children[a]) ** int(c[a, s])
            for v in range(V):
                for s in range(S):
                    den *= math.factorial(int(x[v, s]))
            weight.append(num / den)
        levels.append((-Q, coarse_of, np.array(weight)))
        dims, spacing = cdims, cgrid.spacing
        D = [2 * d for d in D]


def textbook_cycle(levels, shift, b, l=0, w=2.0 / 3.0, pre=2, post=2, cycles=1, _index=None):
    A = levels[l][0] + shift * np.eye(len(b))
    d = np.diag(A)
    x = np.zeros_like(b)
    for _ in range(pre):
        x = x + w * (b - A @ x) / d
    r = b - A @ x
    co, wt = levels[l][1], levels[l][2]
    if co is None:
        x = x + np.linalg.solve(A, r)
    else:
        idx = _index[l]
        nc = len(levels[l + 1][0])
        rc = np.zeros(nc)
        np.add.at(rc, idx, r)
        Ac = levels[l + 1][0] + shift * np.eye(nc)
        xc = np.zeros(nc)
        for _ in range(cycles):
            xc = xc + textbook_cycle(levels, shift, rc - Ac @ xc, l + 1, w, pre, post, cycles, _index)
        x = x + wt * xc[idx]
    for _ in range(post):
        x = x + w * (b - A @ x) / d
    return x


def coarse_index(levels, dims, S, totals):
    """Per level, the index of each fine state's coarse state in the next level's order."""
    out = []
    grid_dims = list(dims)
    for l in range(len(levels) - 1):
        grid_dims = [d // 2 if d > 1 else 1 for d in grid_dims]
        Vc = int(np.prod(grid_dims))
        nxt = [st for st in bruteforce.box_states(Vc, list(totals))
               if tuple(np.array(st).reshape(Vc, S).sum(axis=0)) == tuple(totals)]
        pos = {st: i for i, st in enumerate(nxt)}
        assert len(nxt) == len(levels[l + 1][0])
        out.append(np.array([pos[k] for k in levels[l][1]]))
    return out


CASES = {
    # name: (dims, spacing, diffusivities, caps)
    "1d_v8_cap3": ([8], [1.0 / 8], [1.0], [3]),  # sector 3: 120 -> 20 -> 4 states
    "1d_v8_cap2_D2": ([8], [1.0 / 8], [2.0], [2]),
    "1d_v4_cap2": ([4], [0.25], [4.0], [2]),
    "2d_2x4_cap2": ([2, 4], [0.5, 0.25], [1.0], [2]),  # axes 2, 4 -> 1, 2 -> 1, 1
    "1d_v4_two_species": ([4], [0.25], [1.0, 0.5], [1, 1]),
}


def sector_list(caps):
    return list(itertools.product(*[range(c + 1) for c in caps]))


def textbook_apply(name, shift, b, **kw):
    dims, spacing, D, caps = CASES[name]
    S = len(caps)
    out = np.zeros_like(b)
    lo = 0
    for totals in sector_list(caps):
        levels = sector_levels(dims, spacing, D, totals)
        n = len(levels[0][0])
        idx = coarse_index(levels, dims, S, totals)
        out[lo:lo + n] = textbook_cycle(levels, shift, b[lo:lo + n], _index=idx, **kw)
        lo += n
    assert lo == len(b)
    return out


def oracle_mg(name, shift, **kw):
    dims, spacing, D, caps = CASES[name]
    model = ReactionDiffusion(VoxelDomain(Grid(dims, spacing), D), birth_death(0.05, 0.02)
                              if len(caps) == 1 else [])
    P = Problem(model, caps)
    return P, SectorMultigrid(P, shift, **kw)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("cycles", [1, 2], ids=["V", "W"])
def test_jacobi_cycle_equals_textbook(name, cycles):
    shift = 0.7
    P, M = oracle_mg(name, shift, coarse_cycles=cycles)
    b = np.random.default_rng(2).random(P.n) - 0.3
    z = M.apply(b)
    ref = textbook_apply(name, shift, b, cycles=cycles)
    scale = np.abs(ref).max()
    assert np.abs(z - ref).max() <= 1e-13 * scale, np.abs(z - ref).max() / scale
    # the mistakes this pin must catch: each variant is far from the oracle's cycle (V and W
    # coincide where the level below the finest is solved exactly, so the cycle count is only
    # checked on hierarchies of three or more levels)
    dims, spacing, D, caps = CASES[name]
    deep = max(len(sector_levels(dims, spacing, D, t)) for t in sector_list(caps)) >= 3
    wrongs = [dict(w=1.0, cycles=cycles), dict(post=0, cycles=cycles), dict(pre=1, cycles=cycles)]
    if deep:
        wrongs.append(dict(cycles=3 - cycles))
    for wrong in wrongs:
        other = textbook_apply(name, shift, b, **wrong)
        assert np.abs(z - other).max() > 1e-6 * scale, (name, wrong)


def test_levels_match_the_oracle_hierarchy():
    """The independent hierarchy has the oracle's level sizes sector by sector."""
    for name, (dims, spacing, D, caps) in CASES.items():
        _, M = oracle_mg(name, 1.0)
        rows = M.level_rows()
        for k, totals in enumerate(sector_list(caps)):
            levels = sector_levels(dims, spacing, D, totals)
            got = [int(r) for r in rows[k] if r > 0]
            assert got == [len(L[0]) for L in levels], (name, totals, got)
