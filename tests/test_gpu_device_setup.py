"""Device-side generator and multigrid-level assembly (assemble.hip) against the host assembly
(setup.cpp), which the oracle pins: for each box both are built through the C ABI, and the exit
and leak rates, the jump count and the operator (a shifted apply of random vectors, which reads
every stored entry) must be bitwise equal. Cases: 1D birth-death boxes, the 2D two-species
annihilation box, a reflecting boundary, two reactions with the same change vector (rows with
merged entries), an explicit voxel graph with uneven volumes, a three-species box, and the
11.2M-state box. Hierarchies whose level operators are assembled on the device give bitwise
the host hierarchy's cycle, for every smoother and precision.
"""
import time

import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.model import Graph, Grid, LocalReaction, ReactionDiffusion, VoxelDomain
from markovkit_amd.workloads import annihilation_2d, birth_death_1d, config_large

pytestmark = pytest.mark.gpu


def build(box, model, boundary, device):
    mk.device_setup_enable(device)
    try:
        t = time.perf_counter()
        Q = mk.Generator(box, model, boundary)
        mk.synchronize()
        return Q, time.perf_counter() - t
    finally:
        mk.device_setup_enable(True)


def compare(box, model, boundary=mk.ABSORBING, trials=2, name=""):
    Qh, th = build(box, model, boundary, False)
    Qd, td = build(box, model, boundary, True)
    assert Qd.rows() == Qh.rows()
    assert Qd.n_jumps() == Qh.n_jumps()
    assert np.array_equal(Qd.exit_rates(), Qh.exit_rates())
    assert np.array_equal(Qd.leak_rates(), Qh.leak_rates())
    rng = np.random.default_rng(5)
    for _ in range(trials):
        x = rng.standard_normal(Qh.rows())
        assert np.array_equal(Qd.apply_shifted(0.7, x), Qh.apply_shifted(0.7, x))
    print(f"[{name}] {Qh.rows()} states: host {th:.3f}s device {td:.3f}s")


def two_births():
    # two reactions with the same change: their jumps into a state share a source and merge
    return [LocalReaction([1], lambda n, vol: 0.3 * vol),
            LocalReaction([1], lambda n, vol: 0.1 + 0.2 * n[0]),
            LocalReaction([-1], lambda n, vol: 0.05 * n[0] * n[0])]


@pytest.mark.parametrize("wl", [birth_death_1d(16, 3, 0.5, "empty"), birth_death_1d(8, 2, 0.7, "random"),
                                birth_death_1d(32, 4, 0.5, "random"), annihilation_2d(4, (2, 2), 1.0),
                                annihilation_2d(8, (1, 1), 1.0)],
                         ids=["969", "45", "58905", "2d-4x4", "2d-8x8-caps11"])
@pytest.mark.parametrize("boundary", [mk.ABSORBING, mk.REFLECTING], ids=["absorbing", "reflecting"])
def test_device_assembly_matches_host(wl, boundary):
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    compare(box, wl.model, boundary, name=wl.name)


def test_merged_rows_and_graph_domain():
    g = Graph(5)
    for u, v, w in [(0, 1, 1.0), (1, 2, 0.5), (2, 3, 2.0), (3, 4, 1.5), (0, 4, 0.25), (1, 3, 0.75)]:
        g.add_edge(u, v, w)
    dom = VoxelDomain(g, np.array([1.0, 0.5, 2.0, 0.8, 1.3]), [1.0])
    model = ReactionDiffusion(dom, two_births())
    compare(mk.PopulationBox(5, [6]), model, name="graph-merged")


def test_three_species():
    dom = VoxelDomain(Grid([4, 2], [0.5, 0.5]), [1.0, 0.5, 2.0])
    reactions = [LocalReaction([1, 0, 0], lambda n, vol: 0.2 * vol),
                 LocalReaction([-1, 1, 0], lambda n, vol: 0.7 * n[0]),
                 LocalReaction([0, -1, 1], lambda n, vol: 0.3 * n[1] * (n[2] + 1)),
                 LocalReaction([0, 0, -1], lambda n, vol: 0.1 * n[2])]
    compare(mk.PopulationBox(8, [2, 2, 1]), ReactionDiffusion(dom, reactions), name="three-species")


def test_full_size_box():
    w = config_large()
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    compare(box, w.model, trials=1, name=w.name)


def hierarchy(w, box, smoother, precision, device):
    mk.device_setup_enable(device)
    try:
        t = time.perf_counter()
        M = mk.SectorMultigrid(w.model.domain, box, w.shifts[0],
                               mk.MultigridOptions(precision=precision, smoother=smoother))
        mk.synchronize()
        return M, time.perf_counter() - t
    finally:
        mk.device_setup_enable(True)


@pytest.mark.parametrize("smoother", [mk.JACOBI, mk.RED_BLACK, mk.ILU0], ids=["jacobi", "red_black", "ilu0"])
@pytest.mark.parametrize("precision", [mk.FP32, mk.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("wl", [birth_death_1d(32, 4, 0.5, "random"), annihilation_2d(4, (2, 2), 1.0)],
                         ids=["58905", "2d-4x4"])
def test_device_hierarchy_matches_host(wl, precision, smoother):
    """level operators assembled on the device give the same hierarchy: the cycles agree bitwise"""
    box = mk.PopulationBox(wl.n_voxels, wl.max_totals)
    Mh, th = hierarchy(wl, box, smoother, precision, False)
    Md, td = hierarchy(wl, box, smoother, precision, True)
    b = np.random.default_rng(9).standard_normal(box.size())
    assert np.array_equal(Md.apply(b), Mh.apply(b))
    print(f"[{wl.name}] hierarchy host {th:.3f}s device {td:.3f}s")


def test_device_hierarchy_full_size():
    w = config_large()
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Mh, th = hierarchy(w, box, mk.RED_BLACK, mk.FP32, False)
    b = np.random.default_rng(9).random(box.size())
    zh = Mh.apply(b)
    del Mh
    Md, td = hierarchy(w, box, mk.RED_BLACK, mk.FP32, True)
    assert np.array_equal(Md.apply(b), zh)
    print(f"[{w.name}] red-black FP32 hierarchy host {th:.2f}s device {td:.2f}s")
