"""mk_gmres_device_noalloc (the allocation-free FGMRES, its Krylov workspace in the library's
device scratch array) against mk_gmres_device, both called through the C ABI with b and x in
HBM: the same kernels and arithmetic, so solutions, iteration counts and residuals are the same
bit for bit. Two calls in a row on different systems, the second smaller (the scratch array keeps
the first call's contents), with an FP32 and an FP64 hierarchy and without a preconditioner;
requests beyond the scratch array's size or restart 64 fail with MK_ERR_INVALID_ARGUMENT.
"""
import ctypes as C

import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd._lib import lib
from markovkit_amd.api import _GmresOptions, _SolverResult
from markovkit_amd.workloads import birth_death_1d, config_small

pytestmark = pytest.mark.gpu

MK_ERR_INVALID_ARGUMENT = 1


def system(w, precision, smoother=mk.RED_BLACK, seed=0):
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    shift = w.shifts[0]
    M = None
    if precision is not None:
        M = mk.SectorMultigrid(w.model.domain, box, shift,
                               mk.MultigridOptions(precision=precision, smoother=smoother))
    b = np.random.default_rng(seed).random(box.size())
    b /= b.sum()
    return Q, M, shift, b


def call(name, Q, M, shift, b, restart):
    n = Q.rows()
    bd, xd = mk.DeviceVector(n, b), mk.DeviceVector(n, np.zeros(n))  # x = 0: the initial guess
    opts = _GmresOptions(1e-10 * float(np.linalg.norm(b)), 1000, restart, 0)
    res = _SolverResult()
    st = getattr(lib(), name)(Q._h, shift, M._h if M is not None else None, bd.ptr, xd.ptr,
                              C.byref(opts), C.byref(res))
    return st, xd.download(), res


def check_same(Q, M, shift, b, restart):
    st0, x0, r0 = call("mk_gmres_device", Q, M, shift, b, restart)
    st1, x1, r1 = call("mk_gmres_device_noalloc", Q, M, shift, b, restart)
    assert st0 == 0 and st1 == 0, lib().mk_last_error()
    assert r0.converged and r1.converged
    assert r0.iterations == r1.iterations and r0.residual == r1.residual
    assert np.array_equal(x0, x1)
    r = b - Q.apply_shifted(shift, x1)
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b)
    return r1.iterations


def test_noalloc_matches_the_allocating_entry_point():
    print("[noalloc] 58,905 states, FP32 red-black hierarchy, restart 50", flush=True)
    big = system(birth_death_1d(32, 4, 0.5, "random"), mk.FP32)
    it_big = check_same(*big, restart=50)
    print("[noalloc] 969 states, FP64 Jacobi hierarchy, restart 30 (smaller, after the first)",
          flush=True)
    small = system(config_small(), mk.FP64, mk.JACOBI, seed=3)
    it_small = check_same(*small, restart=30)
    print("[noalloc] 969 states, no preconditioner, restart 64", flush=True)
    plain = system(config_small(), None, seed=4)
    check_same(*plain, restart=64)
    print("[noalloc] 58,905 states again after the small calls", flush=True)
    assert check_same(*big, restart=50) == it_big
    print(f"[noalloc] iterations {it_big}, {it_small}")


def test_noalloc_rejects_what_does_not_fit():
    Q, M, shift, b = system(config_small(), mk.FP32)
    st, _, _ = call("mk_gmres_device_noalloc", Q, M, shift, b, 65)
    assert st == MK_ERR_INVALID_ARGUMENT
    assert b"restart" in lib().mk_last_error()
    print("[noalloc] 435,897 states, restart 50: above the scratch array", flush=True)
    Q, M, shift, b = system(birth_death_1d(32, 5, 0.5, "random"), mk.FP32)
    st, _, _ = call("mk_gmres_device_noalloc", Q, M, shift, b, 50)
    assert st == MK_ERR_INVALID_ARGUMENT
    assert b"scratch" in lib().mk_last_error()
    # the library still solves on the allocating path afterwards
    st, x, res = call("mk_gmres_device", Q, M, shift, b, 50)
    assert st == 0 and res.converged
