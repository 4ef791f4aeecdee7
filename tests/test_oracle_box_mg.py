"""Pins of the oracle's whole-box multigrid (oracle/box_mg.hpp), the CPU restatement of the
reference's paper driver rdme/cpp/rdme_full_multigrid.cpp.

- The driver's recorded runs (tests/golden/box_mg_recorded.json, collected from
  rdme/paper/data/full_multigrid/): the oracle repeats each run's standalone iteration and
  must take exactly the recorded number of cycles, end with the recorded mass and minimum
  probability, and follow the recorded error history (tight on the first cycles; the last
  cycles sit near 1e-9, where the two builds' different rounding shows).
- The driver's --check invariants (:923-979): no positive off-diagonal, column sums of
  shift I - Q equal shift + leak, R P = I, and R A P = A_coarse on every level.
- The state order is the library's population box order (total first, voxel 0 most
  significant), checked against markovkit_amd.PopulationBox.
- The error bound: |x - A^-1 b|_1 <= |b - A x|_1 / shift, against a dense solve.
- The two library options: the four colours separate every coupling, and a terminal level
  with at least as many rows as the fine level makes one cycle an exact solve.
"""
import json
import os

import numpy as np
import pytest

from oracle.oracle import BoxMultigrid

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "box_mg_recorded.json")) as f:
    RECORDED = json.load(f)

FAST = ["graph_gap_cap2_w", "voxels_1d_64_mg", "voxels_1d_64_kcycle", "voxels_2d_4_kcycle",
        "voxels_2d_8_mg", "voxels_3d_2_kcycle", "voxels_3d_4_mg", "voxels_1d_128_mg",
        "block_v16_k4_gs_s1", "block_v16_k4_gs_s2", "block_v16_k4_block-ilu_s1",
        "block_v16_k4_block-ilu_s2", "block_schlogl_v4_k40_block-ilu_s1_kcycle"]
SLOW = ["block_schlogl_v4_k40_gs_s1", "block_schlogl_v4_k40_gs_s1_kcycle", "block_v16_k8_gs_s2"]


def build(p, **kw):
    args = dict(dimension=p["dimension"], cells=p["cells"], cap=p["cap"], diffusion=p["diffusion"],
                shift=p["shift"], rates=tuple(p["rates"]), smoother=p["smoother"],
                sweeps=p["sweeps"], cycles=p["cycles"], kcycle=p["kcycle"])
    args.update(kw)
    return BoxMultigrid(**args)


def check_recorded(name):
    rec = RECORDED[name]
    p = rec["params"]
    M = build(p)
    assert M.n == rec["states"]
    b, x0 = M.initial_problem(p["initial"])
    x, it, err, hist = M.solve(b, x0, p["tolerance"], p["max_iterations"])
    assert it == rec["iterations"], (name, it, rec["iterations"])
    assert err < p["tolerance"]
    assert abs(err - rec["error"]) <= 2e-3 * rec["error"]
    assert abs(x.sum() - rec["mass"]) <= 1e-11
    assert abs(x.min() - rec["min_probability"]) <= 1e-6 * abs(rec["min_probability"])
    if "history" in rec:
        H = np.array(rec["history"])
        assert len(hist) == len(H)
        rel = np.abs(hist - H) / H
        assert np.all(rel[:5] <= 1e-9), rel[:5]
        assert np.all(rel <= 2e-3), rel.max()


@pytest.mark.parametrize("name", FAST)
def test_recorded_run(name):
    check_recorded(name)


@pytest.mark.slow
@pytest.mark.parametrize("name", SLOW)
def test_recorded_run_large(name):
    check_recorded(name)


def csr(M, level):
    import scipy.sparse as sp
    ptr, col, val = M.matrix(level)
    n = len(ptr) - 1
    return sp.csr_matrix((val, col, ptr), shape=(n, n))


@pytest.mark.parametrize("dimension,cells,cap,rates", [
    (1, 16, 3, BoxMultigrid.linear(1.0)), (2, 4, 3, BoxMultigrid.linear(0.7)),
    (3, 2, 3, BoxMultigrid.linear(1.0)), (1, 4, 12, None), (2, 2, 10, None)],
    ids=["1d", "2d", "3d", "schlogl-1d", "schlogl-2d"])
def test_check_invariants(dimension, cells, cap, rates):
    """the driver's --check: sign pattern, mass balance, R P = I and R A P = A_coarse"""
    shift = 0.5
    M = BoxMultigrid(dimension, cells, cap, 1.0, shift, rates, "gs", 2, 2)
    for level in range(M.levels()):
        A = csr(M, level)
        n = A.shape[0]
        off = A - __import__("scipy.sparse", fromlist=["diags"]).diags(A.diagonal())
        assert off.max() <= 0.0
        colsum = np.asarray(A.sum(axis=0)).ravel()
        leak = M.leak(level)
        assert np.all(np.abs(colsum - shift - leak) <= 1e-10 * A.diagonal())
        if M.terminal(level):
            continue
        parent, weight = M.transfer(level)
        nc = M.rows(level + 1)
        q = np.sin(0.7 * np.arange(nc)) + 0.25 * np.cos(0.31 * np.arange(nc))
        y = weight * q[parent]
        rp = np.bincount(parent, weights=y, minlength=nc)
        assert np.max(np.abs(rp - q)) <= 1e-12
        rap = np.bincount(parent, weights=A @ y, minlength=nc)
        Ac = csr(M, level + 1) @ q
        assert np.max(np.abs(rap - Ac)) <= 1e-10 * max(1.0, np.max(np.abs(Ac)))
        assert n > nc


def test_order_is_population_box_order():
    import markovkit_amd as mk
    for V, cap in [(4, 2), (8, 3), (16, 2)]:
        M = BoxMultigrid(1, V, cap, 1.0, 0.5, BoxMultigrid.linear(1.0))
        box = mk.PopulationBox(V, [cap])
        assert box.size() == M.n
        for i in range(M.n):
            assert M.rank(np.asarray(box[i])) == i


def test_error_bound_against_direct_solve():
    M = BoxMultigrid(1, 16, 2, 1.0, 0.5, BoxMultigrid.linear(1.0), "gs", 2, 2)
    b, x0 = M.initial_problem(2)
    x, it, err, _ = M.solve(b, x0, 1e-6, 100)
    exact = np.linalg.solve(M.dense(0), b)
    assert np.abs(x - exact).sum() <= err + 1e-14
    # and with the tolerance of the recorded runs the solution agrees to the bound
    x, it, err, _ = M.solve(b, x0, 1e-9, 100)
    assert np.abs(x - exact).sum() <= err + 1e-14


@pytest.mark.parametrize("dimension,cells,cap,rates", [
    (1, 16, 4, BoxMultigrid.linear(1.0)), (2, 4, 4, None), (3, 4, 2, BoxMultigrid.linear(1.0))])
def test_four_colours_separate_couplings(dimension, cells, cap, rates):
    M = BoxMultigrid(dimension, cells, cap, 1.0, 0.5, rates, "colour-gs", 2, 2)
    for level in range(M.levels()):
        if M.terminal(level):
            continue
        ptr, col, _ = M.matrix(level)
        c = M.colours(level)
        rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
        off = col != rows
        assert np.all(c[col[off]] != c[rows[off]])
        assert set(np.unique(c)) <= {0, 1, 2, 3}


@pytest.mark.parametrize("reverse", [False, True], ids=["pre", "post"])
@pytest.mark.parametrize("smoother", ["colour-gs", "gs", "jacobi"])
def test_point_smoothers_against_dense(smoother, reverse):
    """one sweep of each point smoother against its dense matrix form: Gauss-Seidel in the
    colour-sorted (colours descending for post-smoothing) or index order solves
    (D + L_order) x_new = b - U_order x_old; Jacobi is x + (2/3) D^-1 (b - A x)"""
    M = BoxMultigrid(2, 4, 3, 1.0, 0.5, None, smoother, 1, 1)
    A = M.dense(0)
    rng = np.random.default_rng(3)
    b, x0 = rng.random(M.n), rng.random(M.n)
    got = M.smooth(0, b, x0, reverse)
    d = np.diag(A)
    if smoother == "jacobi":
        want = x0 + (2.0 / 3.0) * (b - A @ x0) / d
    else:
        if smoother == "colour-gs":
            c = M.colours(0).astype(np.int64)
            order = np.argsort(-c if reverse else c, kind="stable")
        else:
            order = np.arange(M.n)[::-1] if reverse else np.arange(M.n)
        P = A[np.ix_(order, order)]
        lower = np.tril(P)
        upper = P - lower
        want = np.empty(M.n)
        want[order] = np.linalg.solve(lower, b[order] - upper @ x0[order])
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14)


def test_block_ilu_pivots_against_generic_ilu():
    """the structured population-block ILU(0) sweep against a dense ILU(0) of each block (the
    driver's own --check comparison, :923-951): block Gauss-Seidel over the population blocks,
    each block's correction the ILU(0) solve of its residual"""
    M = BoxMultigrid(1, 8, 3, 1.0, 0.5, BoxMultigrid.linear(1.0), "block-ilu", 1, 1)
    A = M.dense(0)
    n = M.n
    rng = np.random.default_rng(5)
    b, x = rng.random(n), rng.random(n)
    got = M.smooth(0, b, x)
    import markovkit_amd as mk
    box = mk.PopulationBox(8, [3])
    pop = np.array([int(np.sum(box[i])) for i in range(n)])
    want = x.copy()
    for k in range(4):
        idx = np.flatnonzero(pop == k)
        B = A[np.ix_(idx, idx)].copy()
        m = len(idx)
        pattern = B != 0
        for i in range(1, m):           # IKJ ILU(0) on the block's pattern
            for kk in range(i):
                if not pattern[i, kk]:
                    continue
                B[i, kk] /= B[kk, kk]
                for j in range(kk + 1, m):
                    if pattern[i, j]:
                        B[i, j] -= B[i, kk] * B[kk, j]
        r = (b - A @ want)[idx]
        Lw = np.tril(B, -1) + np.eye(m)
        Uw = np.triu(B)
        want[idx] += np.linalg.solve(Uw, np.linalg.solve(Lw, r))
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14)


def test_terminal_level_solves_exactly():
    M = BoxMultigrid(2, 4, 3, 1.0, 0.5, None, "colour-gs", 2, 2, terminal_rows=10_000)
    assert M.terminal(0) == 2
    b = np.random.default_rng(1).random(M.n)
    z = M.apply(b)
    assert np.allclose(z, np.linalg.solve(M.dense(0), b), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("smoother", ["colour-gs", "jacobi"])
@pytest.mark.parametrize("terminal", [0, 200])
def test_library_options_converge(smoother, terminal):
    M = BoxMultigrid(1, 32, 2, 1.0, 0.5, BoxMultigrid.linear(1.0), smoother, 2, 2,
                     terminal_rows=terminal)
    b, x0 = M.initial_problem(2)
    x, it, err, hist = M.solve(b, x0, 1e-9, 300)
    assert err < 1e-9 and it < 300
    assert abs(x.sum() - RECORDED["voxels_1d_64_mg"]["mass"]) <= 1e-10
