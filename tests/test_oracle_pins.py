"""Pins of the FP64 CPU oracle (oracle/mk_oracle.hpp) against what the reference fixes:
the exact values of tests/test_rdme.cpp, the reference's recorded transfer weights
(rdme/paper/data/transfer_sparsity/*.csv, copied under tests/golden/), its recorded masses
(rdme/paper/data/full_multigrid/voxels_1d_64_mg.csv:2, block_v64_k4_gs_s1.csv:2) through the
lumped closed form, an independent brute-force generator, and direct sparse solves.
"""
import csv
import math
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spl

from markovkit_amd.model import Graph, Grid, LocalReaction, ReactionDiffusion, VoxelDomain, birth_death
from markovkit_amd.workloads import birth_death_1d, config_small, lumped_sector_sums
from oracle.oracle import Problem, SectorMultigrid
from tests import bruteforce

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def diffusion(V, h, D, S=1):
    return ReactionDiffusion(VoxelDomain(Grid([V], [h]), [D] * S), [])


# ---------------------------------------------------------------- geometry and rates
def test_structured_voxel_domain():
    """tests/test_rdme.cpp:22-33"""
    dom = VoxelDomain(Grid([3, 2], [0.5, 2.0]), [0.25, 1.0])
    assert dom.n_voxels() == 6 and dom.n_species() == 2
    assert abs(dom.volumes[0] - 1.0) < 1e-15
    nb = dom.graph.neighbors(0)
    assert len(nb) == 2
    assert abs(dom.jump_rate(0, 0, nb[0][1]) - 1.0) < 1e-15
    assert abs(dom.jump_rate(0, 0, nb[1][1]) - 0.0625) < 1e-15


def test_explicit_graph_voxel_domain():
    """tests/test_rdme.cpp:35-43"""
    g = Graph(2)
    g.add_edge(0, 1, 3.0)
    dom = VoxelDomain(g, [2.0, 4.0], [0.5])
    assert abs(dom.jump_rate(0, 0, g.neighbors(0)[0][1]) - 0.75) < 1e-15
    assert abs(dom.jump_rate(0, 1, g.neighbors(1)[0][1]) - 0.375) < 1e-15


def jumps_from(P, x):
    first, to, rate = P.jumps()
    j = P.find(x)
    states = P.states()
    return {tuple(states[to[k]]): rate[k] for k in range(first[j], first[j + 1])}


def test_reaction_diffusion_transitions():
    """tests/test_rdme.cpp:45-64: h = 0.5 gives 4 per molecule; births at 0.25"""
    model = ReactionDiffusion(VoxelDomain(Grid([2], [0.5]), [1.0]),
                              [LocalReaction([1], lambda n, vol: 0.25)])
    P = Problem(model, [6])
    J = jumps_from(P, [3, 2])
    assert len(J) == 4
    assert abs(J[(2, 3)] - 12.0) < 1e-15
    assert abs(J[(4, 1)] - 8.0) < 1e-15
    assert abs(J[(4, 2)] - 0.25) < 1e-15
    assert abs(J[(3, 3)] - 0.25) < 1e-15


# ---------------------------------------------------------------- state spaces
def sector_size(P, k):
    b = P.sector_begins()
    return b[k + 1] - b[k]


def test_population_spaces():
    """tests/test_rdme.cpp:75-98"""
    assert sector_size(Problem(diffusion(3, 1.0, 1.0), [4]), 4) == 15
    assert sector_size(Problem(diffusion(3, 1.0, 1.0), [16]), 16) == 153
    assert math.comb(64 + 2, 2) == 2145  # population_sector(3, {64})
    two = Problem(diffusion(2, 1.0, 1.0, S=2), [1, 2])
    k = 1 * 3 + 2
    assert sector_size(two, k) == 6
    b = two.sector_begins()
    states = two.states()
    for i in range(b[k], b[k + 1]):
        x = states[i]
        assert x[0] + x[2] == 1 and x[1] + x[3] == 2
        assert two.find(x) == i
    assert two.find([2, 0, 0, 2]) == -1
    box = Problem(diffusion(3, 1.0, 1.0, S=2), [2, 2])
    states = box.states()
    begins = box.sector_begins()
    for i in range(len(states)):
        assert box.find(states[i]) == i
        k = np.searchsorted(begins, i, side="right") - 1
        assert begins[k] <= i < begins[k + 1]
    assert box.find([3, 0, 0, 0, 0, 0]) == -1


@pytest.mark.parametrize("V,caps", [(16, [3]), (64, [5]), (8, [2, 2]), (32, [4])])
def test_box_sizes(V, caps):
    """closed-form sizes of the benchmark boxes: sum over sectors of prod C(t+V-1, V-1)"""
    from markovkit_amd.model import box_size
    expected = {(16, (3,)): 969, (64, (5,)): 11238513, (8, (2, 2)): None, (32, (4,)): 58905}
    n = box_size(V, caps)
    if (V, tuple(caps)) == (8, (2, 2)):
        assert box_size(64, caps) == 2145 ** 2
    else:
        assert n == expected[(V, tuple(caps))]


# ---------------------------------------------------------------- generator
def test_diffusion_generator():
    """tests/test_rdme.cpp:100-120"""
    P = Problem(diffusion(3, 0.5, 0.25), [4])
    b = P.sector_begins()
    source = P.find([4, 0, 0])
    ex, _ = P.rates()
    assert abs(ex[source] - 4.0) < 1e-15
    p = np.zeros(P.n)
    p[source] = 1.0
    free = P.apply(p)
    assert abs(free[source] + 4.0) < 1e-15
    assert abs(free.sum()) < 1e-15
    assembled = P.matrix() @ p
    assert np.all(np.abs(free - assembled) < 1e-15)


def test_reactions_couple_neighboring_sectors():
    """tests/test_rdme.cpp:122-158"""
    model = ReactionDiffusion(VoxelDomain(Grid([2], [1.0]), [0.5]), birth_death(2.0, 0.25))
    P = Problem(model, [3])
    assert P.n == 10
    first, to, rate = P.jumps()
    begins = P.sector_begins()
    sector = lambda i: np.searchsorted(begins, i, side="right") - 1
    seen = set()
    for j in range(P.n):
        for k in range(first[j], first[j + 1]):
            change = sector(to[k]) - sector(j)
            assert abs(change) <= 1
            seen.add(change)
    assert seen == {-1, 0, 1}
    ex, _ = P.rates()
    empty, full = P.find([0, 0]), P.find([3, 0])
    assert abs(ex[empty] - 4.0) < 1e-15
    assert abs(ex[full] - 6.25) < 1e-15
    p = np.zeros(P.n)
    p[full] = 1.0
    free = P.apply(p)
    assert abs(free.sum() + 4.0) < 1e-15
    assert np.all(np.abs(free - P.matrix() @ p) < 1e-15)
    R = Problem(model, [3], absorbing=False)
    exr, lkr = R.rates()
    assert abs(exr[full] - 2.25) < 1e-15
    assert abs(lkr[full] - 4.0) < 1e-15


@pytest.mark.parametrize("absorbing", [True, False])
@pytest.mark.parametrize("case", ["1d_bd", "2d_annihilation", "graph"])
def test_generator_matches_bruteforce(case, absorbing):
    """ordering, rates, exit and leak against the independent brute-force enumeration"""
    if case == "1d_bd":
        model, caps = ReactionDiffusion(VoxelDomain(Grid([3], [0.5]), [0.7]),
                                        birth_death(0.3, 0.2)), [3]
    elif case == "2d_annihilation":
        model = ReactionDiffusion(VoxelDomain(Grid([2, 2], [0.5, 0.25]), [1.0, 0.5]), [
            LocalReaction([1, 0], lambda n, v: 0.1 * v),
            LocalReaction([0, 1], lambda n, v: 0.2 * v),
            LocalReaction([-1, -1], lambda n, v: 1.5 * n[0] * n[1] / v)])
        caps = [2, 1]
    else:
        g = Graph(3)
        g.add_edge(0, 1, 3.0)
        g.add_edge(1, 2, 1.5)
        model = ReactionDiffusion(VoxelDomain(g, [2.0, 4.0, 1.0], [0.5]), birth_death(0.4, 0.3))
        caps = [2]
    Qb, exb, lkb, states = bruteforce.dense_generator(model, caps, absorbing)
    P = Problem(model, caps, absorbing=absorbing)
    assert P.n == len(states)
    np.testing.assert_array_equal(P.states(), np.array(states))
    ex, lk = P.rates()
    np.testing.assert_allclose(ex, exb, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(lk, lkb, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(P.matrix().toarray(), Qb, rtol=1e-15, atol=1e-14)


# ---------------------------------------------------------------- multigrid
def multigrid_case():
    model = diffusion(8, 1.0, 1.0)
    P = Problem(model, [2])
    return P, SectorMultigrid(P, 1.0)


def test_population_multigrid_structure():
    """tests/test_rdme.cpp:160-166, :214-217"""
    P, M = multigrid_case()
    rows = M.level_rows()
    assert list(rows[2]) == [36, 10, 3]
    assert M.max_levels() == 3
    assert P.n == 45 and len(P.sector_begins()) - 1 == 3


def test_transfer_identities():
    """R P = I, mass conservation, R A P = A_c (tests/test_rdme.cpp:170-189)"""
    P, M = multigrid_case()
    co, w = M.transfer(2, 0)
    coarse = np.sin(np.arange(10.0))
    fine = w * coarse[co]
    restricted = np.zeros(10)
    np.add.at(restricted, co, fine)
    assert np.all(np.abs(restricted - coarse) < 1e-14)
    assert abs(fine.sum() - coarse.sum()) < 1e-14
    fine_image = M.level_apply(2, 0, fine)
    projected = np.zeros(10)
    np.add.at(projected, co, fine_image)
    coarse_image = M.level_apply(2, 1, coarse)
    assert np.all(np.abs(projected - coarse_image) < 1e-12)


def test_jacobi_cycle_reduces_residual():
    """tests/test_rdme.cpp:191-208 (through the sector multigrid on the total-2 sector)"""
    P, M = multigrid_case()
    b = np.zeros(P.n)
    lo, hi = P.sector_begins()[2], P.sector_begins()[3]
    b[lo:hi] = 0.25 + np.sin(np.arange(36.0))
    x = M.apply(b)
    r = M.level_apply(2, 0, x[lo:hi]) - b[lo:hi]
    assert (r ** 2).sum() < (b ** 2).sum()
    ones = M.apply(np.ones(P.n))
    assert np.all(np.isfinite(ones))


def read_transfer(path):
    with open(path) as f:
        rows = list(csv.DictReader(f))
    return (np.array([int(r["fine_state"]) for r in rows]),
            np.array([int(r["coarse_state"]) for r in rows]),
            np.array([int(r["population"]) for r in rows]),
            np.array([float(r["weight"]) for r in rows]))


def test_transfer_weights_match_recorded_2to1():
    """rdme/paper/data/transfer_sparsity/schlogl_k40_level1.csv: 2 -> 1 voxels, cap 40"""
    fine, coarse, pop, weight = read_transfer(os.path.join(GOLDEN, "transfer_2to1_cap40.csv"))
    P = Problem(diffusion(2, 0.5, 1.0), [40])
    M = SectorMultigrid(P, 1.0, coarsest_size=0)
    begins = P.sector_begins()
    got_c, got_w = [], []
    for k in range(41):
        co, w = M.transfer(k, 0)
        got_c.append(co + k)  # one coarse state per population on one voxel
        got_w.append(w)
    got_c = np.concatenate(got_c)
    got_w = np.concatenate(got_w)
    assert np.array_equal(fine, np.arange(len(fine)))
    assert np.array_equal(coarse, got_c)
    assert np.array_equal(pop, np.searchsorted(begins, np.arange(P.n), side="right") - 1)
    np.testing.assert_allclose(got_w, weight, rtol=5e-14, atol=0)  # lgamma vs the driver's products


def test_transfer_weights_match_recorded_4to2_population4():
    """rdme/paper/data/transfer_sparsity/schlogl_k40_population4.csv: 4 -> 2 voxels,
    population-4 block"""
    fine, coarse, pop, weight = read_transfer(
        os.path.join(GOLDEN, "transfer_4to2_population4.csv"))
    P = Problem(diffusion(4, 0.25, 1.0), [4])
    M = SectorMultigrid(P, 1.0, coarsest_size=0)
    co, w = M.transfer(4, 0)
    assert np.all(pop == 4)
    assert np.array_equal(fine, np.arange(35))
    assert np.array_equal(coarse, co)
    np.testing.assert_allclose(w, weight, rtol=5e-14, atol=0)


# ---------------------------------------------------------------- solves
def test_multigrid_preconditions_resolvent():
    """tests/test_rdme.cpp:225-251"""
    model = ReactionDiffusion(VoxelDomain(Grid([4], [1.0]), [4.0]), birth_death(0.05, 0.02))
    P = Problem(model, [3])
    shift = 0.5
    b = 1.0 + 0.1 * np.cos(np.arange(P.n, dtype=float))
    x_plain, plain = P.gmres(shift, b, tol=1e-10, max_iterations=200, restart=P.n)
    x_mg, mg = P.gmres(shift, b, mg=SectorMultigrid(P, shift), tol=1e-10, max_iterations=200,
                       restart=P.n)
    assert plain["converged"] and mg["converged"]
    assert mg["iterations"] <= plain["iterations"]
    assert np.all(np.abs(x_mg - x_plain) < 1e-8)


def test_two_particle_annihilation_transform():
    """tests/test_rdme.cpp:253-297"""
    model = ReactionDiffusion(VoxelDomain(Grid([4], [0.25]), [1.0, 1.0]),
                              [LocalReaction([-1, -1], lambda n, v: 2.0 * n[0] * n[1] / v)])
    P = Problem(model, [1, 1])
    assert P.n == 25
    shift = 0.5
    M = SectorMultigrid(P, shift)
    begins = P.sector_begins()
    joint = 1 * 2 + 1
    vacuum = P.find([0] * 8)
    states = P.states()

    def transform(a_voxel, b_voxel):
        init = [0] * 8
        init[a_voxel * 2 + 0] = 1
        init[b_voxel * 2 + 1] = 1
        b = np.zeros(P.n)
        b[P.find(init)] = 1.0
        x, info = P.gmres(shift, b, mg=M, tol=1e-12, max_iterations=100, restart=P.n)
        assert info["converged"]
        survival = x[begins[joint]:begins[joint + 1]].sum()
        flux = sum(2.0 * states[i][2 * v] * states[i][2 * v + 1] / 0.25 * x[i]
                   for i in range(begins[joint], begins[joint + 1]) for v in range(4))
        absorbed = shift * x[vacuum]
        assert abs(absorbed - flux) < 1e-10
        assert abs(absorbed + shift * survival - 1.0) < 1e-10
        return absorbed

    colocated = transform(0, 0)
    separated = transform(0, 3)
    assert colocated > separated
    assert separated > 0.0 and colocated < 1.0


def test_lumped_closed_form_reproduces_recorded_masses():
    """mass = 0.337423312883 (cap 2) and 0.82128451753 (cap 4): linear birth-death with total
    birth rate 2, death 1 per molecule, s = 0.5, two molecules initially
    (rdme/paper/data/full_multigrid/voxels_1d_64_mg.csv:2, block_v64_k4_gs_s1.csv:2). The
    driver's recorded mass is that of the resolvent scaled by s, s * sum(x) with
    (sI - Q) x = e_initial."""
    for cap, recorded in [(2, 0.337423312883), (4, 0.82128451753)]:
        B = np.zeros(cap + 1)
        B[2] = 1.0
        mass = 0.5 * lumped_sector_sums(0.5, 2.0, 1.0, cap, B).sum()
        assert abs(mass - recorded) < 5e-13, (cap, mass)


def test_oracle_solution_matches_recorded_mass():
    """the oracle's resolvent solve on the recorded 1D, 64-voxel, cap-2 case reproduces the
    recorded mass 0.337423312883 (its sector sums match the closed form)"""
    model = ReactionDiffusion(VoxelDomain(Grid([64], [1.0 / 64]), [1.0]), birth_death(2.0, 1.0))
    P = Problem(model, [2])
    assert P.n == 2145
    b = np.zeros(P.n)
    init = [1, 1] + [0] * 62
    b[P.find(init)] = 1.0
    x, info = P.gmres(0.5, b, mg=SectorMultigrid(P, 0.5), tol=1e-12, restart=50)
    assert info["converged"]
    assert abs(0.5 * x.sum() - 0.337423312883) < 5e-13
    begins = P.sector_begins()
    X = lumped_sector_sums(0.5, 2.0, 1.0, 2, np.eye(3)[2])
    # |sum over a sector of (x - x*)| <= ||x - x*||_1 <= ||b - A x||_1 / s
    bound = np.abs(b - P.apply_shifted(0.5, x)).sum() / 0.5 + 1e-15
    assert bound < 1e-10
    for k in range(3):
        assert abs(x[begins[k]:begins[k + 1]].sum() - X[k]) <= bound


def test_small_config_sector_sums_and_direct_solve():
    """969-state config: sector sums equal the lumped closed form; the solution equals the
    committed direct-solve fixture (tests/golden/make_golden.py, brute-force matrix)"""
    w = config_small()
    P = Problem(w.model, w.max_totals)
    b = w.rhs()
    x, info = P.gmres(0.5, b, mg=SectorMultigrid(P, 0.5), tol=1e-12, restart=50)
    assert info["converged"]
    begins = P.sector_begins()
    Xk = [x[begins[k]:begins[k + 1]].sum() for k in range(4)]
    expected = [1.8240352155078126, 0.16096842646485102, 0.013756057739361041,
                0.001127545716341069]
    # the sector sums of x - x* are bounded by ||x - x*||_1 <= ||b - A x||_1 / s
    bound = np.abs(b - P.apply_shifted(0.5, x)).sum() / 0.5
    assert np.abs(np.array(Xk) - expected).sum() <= bound + 1e-15
    X = lumped_sector_sums(0.5, 0.05, 0.02, 3, np.eye(4)[0])
    np.testing.assert_allclose(X, expected, rtol=1e-14)
    ref = np.load(os.path.join(GOLDEN, "rdme1d_v16_cap3_direct.npy"))
    assert np.abs(x - ref).sum() <= np.abs(b - P.apply_shifted(0.5, x)).sum() / 0.5 + 1e-14


def test_gmres_can_fail():
    """a solve against an impossible tolerance reports non-convergence (tests/check.hpp:9-10)"""
    w = config_small()
    P = Problem(w.model, w.max_totals)
    _, info = P.gmres(0.5, w.rhs(), tol=1e-300, max_iterations=3, restart=50)
    assert not info["converged"] and info["iterations"] == 3


def test_shift_must_be_positive():
    P = Problem(diffusion(4, 1.0, 1.0), [2])
    with pytest.raises(ValueError):
        SectorMultigrid(P, 0.0)
