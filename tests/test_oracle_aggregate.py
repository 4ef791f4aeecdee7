"""Pins of the oracle's aggregate-block smoother (include/markovkit/rdme/multigrid.hpp:239-342).

* the reference's own check (tests/test_rdme.cpp:191-210): on the 36-state sector of 8 voxels
  with 2 molecules, a Jacobi V-cycle reduces the residual and the aggregate-block V-cycle
  reduces it further;
* exactness: on 2 voxels the only aggregate holds the whole sector, so one block sweep is a
  direct solve and the cycle returns A^{-1} b;
* a dense brute force of block Gauss-Seidel over aggregates (block Jacobi inside one
  aggregate, groups = states that agree outside the aggregate) written from the matrix and the
  state list, checked against the oracle's relax() forwards and in reverse.
"""
import numpy as np
import pytest

from markovkit_amd.model import Grid, ReactionDiffusion, VoxelDomain
from oracle.oracle import Problem, SectorMultigrid


def diffusion_only(cells, h):
    return ReactionDiffusion(VoxelDomain(Grid([cells], [h]), [1.0]), [])


def sector_slice(P, k):
    b = P.sector_begins()
    return slice(int(b[k]), int(b[k + 1]))


def test_reference_aggregate_blocks_beat_jacobi():
    P = Problem(diffusion_only(8, 1.0), [2])
    sl = sector_slice(P, 2)
    assert sl.stop - sl.start == 36
    A = P.matrix(1.0, -1.0).toarray()[sl, sl]
    b = np.zeros(P.n)
    b[sl] = 0.25 + np.sin(np.arange(36.0))
    initial = float(b[sl] @ b[sl])

    def residual(smoother):
        M = SectorMultigrid(P, 1.0, smoother=smoother)
        x = M.apply(b)[sl]
        r = b[sl] - A @ x
        return float(r @ r)

    jacobi = residual("jacobi")
    block = residual("aggregate_block")
    assert jacobi < initial
    assert block < jacobi, (block, jacobi)


def test_single_aggregate_block_is_a_direct_solve():
    P = Problem(diffusion_only(2, 0.5), [4])
    sl = sector_slice(P, 4)
    assert sl.stop - sl.start == 5
    A = P.matrix(0.7, -1.0).toarray()[sl, sl]
    rng = np.random.default_rng(3)
    rhs = rng.random(5)
    M = SectorMultigrid(P, 0.7, pre_sweeps=1, post_sweeps=1, smoother="aggregate_block")
    x = M.relax(4, rhs, np.zeros(5), sweeps=1)
    exact = np.linalg.solve(A, rhs)
    assert np.abs(x - exact).max() <= 1e-13 * np.abs(exact).max()
    b = np.zeros(P.n)
    b[sl] = rhs
    assert np.abs(M.apply(b)[sl] - exact).max() <= 1e-13 * np.abs(exact).max()


def brute_block_gs(A, states, parent, rhs, x, sweeps, reverse):
    """block Gauss-Seidel over aggregates from the dense matrix: groups of states agreeing
    outside aggregate a (ascending outside configuration), solved together on the residual
    before the aggregate's update"""
    x = x.copy()
    n_agg = max(parent) + 1
    blocks = []
    for a in range(n_agg):
        groups = {}
        for i, st in enumerate(states):
            out = tuple(0 if parent[v] == a else int(c) for v, c in enumerate(st))
            groups.setdefault(out, []).append(i)
        blocks.append([groups[k] for k in sorted(groups) if len(groups[k]) > 1])
    for _ in range(sweeps):
        for step in range(n_agg):
            a = n_agg - 1 - step if reverse else step
            r = rhs - A @ x
            for members in blocks[a]:
                m = np.array(members)
                x[m] += np.linalg.solve(A[np.ix_(m, m)], r[m])
    return x


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("cells,total", [(4, 2), (8, 3)])
def test_block_gauss_seidel_brute_force(cells, total, reverse):
    P = Problem(diffusion_only(cells, 1.0 / cells), [total])
    k = total
    sl = sector_slice(P, k)
    states = P.states()[sl]
    A = P.matrix(0.5, -1.0).toarray()[sl, sl]
    parent = [v // 2 for v in range(cells)]
    rng = np.random.default_rng(cells + total)
    rhs = rng.random(sl.stop - sl.start)
    x0 = rng.random(sl.stop - sl.start)
    M = SectorMultigrid(P, 0.5, smoother="aggregate_block")
    got = M.relax(k, rhs, x0.copy(), sweeps=2, reverse=reverse)
    want = brute_block_gs(A, states, parent, rhs, x0, 2, reverse)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # a plausible mistake fails: visiting the aggregates in the other order changes the result
    other = brute_block_gs(A, states, parent, rhs, x0, 2, not reverse)
    assert np.abs(other - want).max() > 1e-8 * np.abs(want).max()
