"""CPU checks of the HIP library's host setup (markovkit_amd/csrc/setup.cpp, through
tests/native/setup_probe.cpp) against the oracle: state ordering and rank, the gather-form
generator (bit-identical rates, exit and leak), the multigrid hierarchy (level sizes,
transfers, terminal inverses), and the C ABI's exported symbols."""
import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from markovkit_amd.model import Graph, LocalReaction, ReactionDiffusion, VoxelDomain, birth_death
from markovkit_amd.workloads import annihilation_2d, birth_death_1d, config_small
from oracle.oracle import Problem, SectorMultigrid
from tests import probe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cases():
    g = Graph(4)
    g.add_edge(0, 1, 3.0)
    g.add_edge(1, 2, 1.5)
    g.add_edge(1, 3, 0.5)
    graph_model = ReactionDiffusion(VoxelDomain(g, [2.0, 4.0, 1.0, 0.5], [0.5]),
                                    birth_death(0.4, 0.3))
    return [
        ("1d_v16_cap3", config_small().model, [3]),
        ("1d_v12_cap3", birth_death_1d(12, 3).model, [3]),
        ("2d_4x4_caps22", annihilation_2d(4, (2, 2)).model, [2, 2]),
        ("2d_3x2_caps21", annihilation_2d(2, (2, 1)).model, [2, 1]),
        ("graph_cap3", graph_model, [3]),
    ]


@pytest.mark.parametrize("name,model,caps", cases(), ids=[c[0] for c in cases()])
def test_box_states_and_rank(name, model, caps):
    states, begin, found = probe.box_states(model.n_voxels(), caps)
    P = Problem(model, caps)
    np.testing.assert_array_equal(states, P.states())
    np.testing.assert_array_equal(begin, P.sector_begins())
    np.testing.assert_array_equal(found, np.arange(P.n))


@pytest.mark.parametrize("absorbing", [True, False])
@pytest.mark.parametrize("name,model,caps", cases(), ids=[c[0] for c in cases()])
def test_gather_generator_is_reference_generator(name, model, caps, absorbing):
    ptr, col, val, ex, lk, n_jumps = probe.generator(model, caps, absorbing)
    P = Problem(model, caps, absorbing=absorbing)
    first, to, rate = P.jumps()
    ex_o, lk_o = P.rates()
    np.testing.assert_array_equal(ex, ex_o)
    np.testing.assert_array_equal(lk, lk_o)
    assert n_jumps == len(to)
    src = np.repeat(np.arange(P.n), np.diff(first))
    M = sp.csr_matrix((rate, (to, src)), shape=(P.n, P.n))
    M.sum_duplicates()
    M.sort_indices()
    np.testing.assert_array_equal(ptr, M.indptr)
    np.testing.assert_array_equal(col, M.indices)
    np.testing.assert_array_equal(val, M.data)


@pytest.mark.parametrize("name,model,caps", cases()[:4], ids=[c[0] for c in cases()[:4]])
@pytest.mark.parametrize("shift", [0.5, 3.0])
def test_hierarchy_matches_oracle(name, model, caps, shift):
    P = Problem(model, caps)
    Mo = SectorMultigrid(P, shift)
    H = probe.Hierarchy(model, caps, shift)
    rows = Mo.level_rows()
    for k in range(rows.shape[0]):
        depth = int((rows[k] > 0).sum())
        assert H.levels(k) == depth
        for l in range(depth):
            n, terminal, _ = H.info(k, l)
            assert n == rows[k, l]
            assert terminal == (l == depth - 1)
            if terminal:
                # terminal inverse = (shift I - Q_diffusion)^-1 on the level
                ptr, col, val, ex = H.operator(k, l)
                A = np.diag(shift + ex)
                for i in range(n):
                    for q in range(ptr[i], ptr[i + 1]):
                        A[i, col[q]] -= val[q]
                np.testing.assert_allclose(H.inverse(k, l) @ A, np.eye(n), atol=1e-12)
                e = np.eye(n)[0]
                np.testing.assert_allclose(A @ e, Mo.level_apply(k, l, e), rtol=1e-15, atol=1e-15)
            else:
                co, w = H.transfer(k, l)
                co_o, w_o = Mo.transfer(k, l)
                np.testing.assert_array_equal(co, co_o)
                np.testing.assert_array_equal(w, w_o)
                x = np.random.default_rng(k * 10 + l).random(n)
                ptr, col, val, ex = H.operator(k, l)
                y = (shift + ex) * x - np.add.reduceat(val * x[col], ptr[:-1]) * (np.diff(ptr) > 0)
                np.testing.assert_allclose(y, Mo.level_apply(k, l, x), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("name,model,caps", cases(), ids=[c[0] for c in cases()])
def test_sliced_layout_indexing(name, model, caps):
    """the device kernel's sliced-layout index arithmetic, emulated on the host, reproduces the
    oracle's shifted generator (every entry read lies inside its slice)"""
    S = probe.sliced(model, caps)
    assert S is not None
    P = Problem(model, caps)
    x = np.random.default_rng(1).standard_normal(P.n)
    y = probe.sliced_apply(S, 0.7, x)
    y_ref = P.apply_shifted(0.7, x)
    A = P.matrix(0.7, -1.0)
    A.data = np.abs(A.data)
    assert np.all(np.abs(y - y_ref) <= 64 * np.finfo(float).eps * (A @ np.abs(x)))
    # padding: every slice is as wide as its longest row
    for s in range(len(S["soff"]) - 1):
        rows = S["rlen"][64 * s: 64 * (s + 1)]
        assert S["soff"][s + 1] - S["soff"][s] == 64 * int(rows.max())


@pytest.mark.parametrize("name,model,caps", cases()[:4], ids=[c[0] for c in cases()[:4]])
def test_restriction_as_sliced_operator(name, model, caps):
    """restriction = the children of every coarse row summed in ascending order; its sliced
    encoding (columns index the finer level) reproduces the oracle's transfer bit for bit"""
    Mo = SectorMultigrid(Problem(model, caps), 1.0)
    rows = Mo.level_rows()
    for k in range(rows.shape[0]):
        if rows[k, 1] == 0:
            continue
        co, _ = Mo.transfer(k, 0)
        nc = rows[k, 1]
        order = np.argsort(co, kind="stable")
        ptr = np.concatenate([[0], np.cumsum(np.bincount(co, minlength=nc))])
        S = probe.sliced_csr(ptr, order, np.ones(len(order)), np.zeros(nc))
        if S is None:  # a coarse row with more than 255 children keeps the plain layout
            assert np.diff(ptr).max() > 255
            continue
        fine = np.random.default_rng(k).standard_normal(len(co))
        y = probe.sliced_apply(S, 0.0, fine, gather_sum=True)
        ref = np.zeros(nc)
        for i in range(len(co)):
            ref[co[i]] += fine[i]
        np.testing.assert_array_equal(y, ref)
    # a column range beyond 24 bits does not fit the coded layout
    assert probe.sliced_csr(np.array([0, 1]), np.array([1 << 24]), np.ones(1), np.zeros(1)) is None


def test_multigrid_errors():
    model = config_small().model
    with pytest.raises(ValueError):
        probe.Hierarchy(model, [3], 0.0)
    g = Graph(2)
    g.add_edge(0, 1, 1.0)
    with pytest.raises((ValueError, RuntimeError)):
        probe.Hierarchy(ReactionDiffusion(VoxelDomain(g, [1.0, 1.0], [1.0]), []), [2], 1.0)


def test_oversized_box_overflows():
    with pytest.raises(OverflowError):
        probe.box_states(4096, [40])


# ---------------------------------------------------------------- C ABI surface
def header_symbols():
    text = open(os.path.join(ROOT, "include", "markovkit_amd.h")).read()
    return sorted(set(re.findall(r"\b(mk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from markovkit_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    names = header_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_box_through_the_abi_on_host():
    """population_box entry points are host-only: they run without a GPU"""
    import markovkit_amd as mk
    box = mk.PopulationBox(16, [3])
    assert box.size() == 969 and box.n_sectors() == 4
    assert [box.sector_begin(k) for k in range(4)] == [0, 1, 17, 153]
    assert box.find(box[500]) == 500
    assert box.find([4] + [0] * 15) is None
    assert box.sector_totals(2) == [2]
    with pytest.raises(ValueError):
        mk.PopulationBox(0, [3])
    with pytest.raises(OverflowError):
        mk.PopulationBox(4096, [40])
