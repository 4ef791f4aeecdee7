"""Whole-box multigrid of the reference's paper driver (rdme/cpp/rdme_full_multigrid.cpp) on the
device: every level holds the complete single-species reaction-diffusion operator of a
Cartesian grid (Schlogl births and deaths, diffusion, absorbing population cap), coarsened by
halving the cells and doubling the diffusivity.

Parameters follow the driver's command line (rdme_full_multigrid.cpp:49-63, :1366-1451):
dimension, cells, cap, diffusion, shift, the Schlogl constants c1..c4 (``linear(s)`` gives the
driver's --linear s), sweeps, cycles (1 V, 2 W) and kcycle (--cycle-type k). Smoothers:
"colour-gs" (Gauss-Seidel in four-colour order: one launch per colour), "jacobi" (the driver's,
w = 2/3), and the driver's sequential "gs" (lexicographic) and "block-ilu" (population-block
ILU(0)), run as level-scheduled triangular solves: the driver's cycle counts, but long
dependency chains per sweep. terminal_rows > 0 solves the
first level with at most that many states exactly (dense inverse) instead of recursing to the
one-voxel Thomas solve.

The state order is the population box order (total population first, voxel 0 most
significant), the driver's Space order: ``markovkit_amd.PopulationBox(cells**dimension, [cap])``
indexes the same states.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from ._lib import BoxMgParams as _Params
from ._lib import BoxMgResult as _Result
from ._lib import GmresOptions as _GmresOptions
from ._lib import SolverResult as _SolverResult
from ._lib import check, lib
from .api import DeviceVector, Generator, SolverResult

SMOOTHERS = {"colour-gs": 0, "jacobi": 1, "gs": 2, "block-ilu": 3}
DEFAULT_RATES = (0.0133, 0.0006, 0.6924125, 0.114725)  # rdme_full_multigrid.cpp:57


def linear(scale: float = 1.0):
    """the driver's --linear scale: c1 = c2 = 0, c3 = 2 scale, c4 = scale"""
    return (0.0, 0.0, 2.0 * scale, scale)


class _BorrowedGenerator(Generator):
    """the hierarchy's finest operator; owned by the hierarchy"""

    def __init__(self, owner, handle):
        self.box = None
        self.model = None
        self._owner = owner
        self._h = handle
        r, j = C.c_int64(), C.c_int64()
        check(lib().mk_generator_size(self._h, C.byref(r), C.byref(j)))
        self._rows, self._jumps = r.value, j.value

    def __del__(self):
        self._h = None


class BoxMultigrid:
    def __init__(self, dimension: int = 1, cells: int = 4, cap: int = 40, diffusion: float = 1.0,
                 shift: float = 0.05, rates=None, smoother: str = "colour-gs", sweeps: int = 2,
                 cycles: int = 1, kcycle: bool = False, terminal_rows: int = 0):
        if smoother not in SMOOTHERS:
            raise ValueError(f"smoother must be one of {sorted(SMOOTHERS)}")
        c1, c2, c3, c4 = DEFAULT_RATES if rates is None else rates
        self._params = _Params(int(dimension), int(cells), int(cap), float(diffusion), float(shift),
                               float(c1), float(c2), float(c3), float(c4), SMOOTHERS[smoother],
                               int(sweeps), int(cycles), 1 if kcycle else 0, int(terminal_rows))
        self._h = C.c_void_p()
        check(lib().mk_box_mg_create(C.byref(self._params), C.byref(self._h)))
        rows, levels = C.c_int64(), C.c_int32()
        check(lib().mk_box_mg_info(self._h, C.byref(rows), C.byref(levels)))
        self.n = rows.value
        self._levels = levels.value
        self.shift = float(shift)
        self.cap = int(cap)
        self.voxels = int(cells) ** int(dimension)
        self.kcycle = bool(kcycle)
        g = C.c_void_p()
        check(lib().mk_box_mg_generator(self._h, C.byref(g)))
        self._gen = _BorrowedGenerator(self, g)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mk_box_mg_destroy(self._h)
            self._h = None

    def rows(self) -> int:
        return self.n

    def levels(self) -> int:
        return self._levels

    def level(self, l: int):
        """(rows, voxels, stored entries, terminal kind: 0 smoothed, 1 Thomas, 2 dense)"""
        r, v, e, t = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().mk_box_mg_level(self._h, int(l), C.byref(r), C.byref(v), C.byref(e), C.byref(t)))
        return r.value, v.value, e.value, t.value

    def generator(self) -> Generator:
        """the finest operator's generator Q (shift I - Q is the level operator)"""
        return self._gen

    def apply_operator(self, x):
        """(shift I - Q) x on the finest level (host arrays)"""
        return self._gen.apply_shifted(self.shift, np.ascontiguousarray(x, np.float64))

    def initial_problem(self, initial: int):
        """the driver's first step: b = shift p0, x0 = p0, p0 the state with `initial` molecules
        dealt round-robin over the voxels (rdme_full_multigrid.cpp:1121-1124)"""
        from .api import PopulationBox
        x = np.zeros(self.voxels, np.int32)
        for k in range(initial):
            x[k % self.voxels] += 1
        i = PopulationBox(self.voxels, [self.cap]).find(x)
        if i is None:
            raise ValueError("the initial population exceeds the cap")
        p0 = np.zeros(self.n)
        p0[i] = 1.0
        return self.shift * p0, p0

    def cycle(self, b, x):
        """one cycle of A x = b from the guess x; returns the new x (host arrays)"""
        db = DeviceVector(self.n, np.ascontiguousarray(b, np.float64))
        dx = DeviceVector(self.n, np.ascontiguousarray(x, np.float64))
        self.cycle_device(db, dx)
        return dx.download()

    def cycle_device(self, b: DeviceVector, x: DeviceVector):
        check(lib().mk_box_mg_cycle_device(self._h, b.ptr, x.ptr))

    def apply(self, b):
        """M b: one cycle from x = 0"""
        db = DeviceVector(self.n, np.ascontiguousarray(b, np.float64))
        dz = DeviceVector(self.n)
        check(lib().mk_box_mg_apply_device(self._h, db.ptr, dz.ptr))
        return dz.download()

    def solve(self, b, x0, tolerance: float = 1e-9, max_iterations: int = 1000):
        """the driver's standalone iteration: (x, iterations, error, history)"""
        x = np.array(x0, np.float64)
        hist = np.zeros(max_iterations + 2)
        r = _Result()
        check(lib().mk_box_mg_solve(self._h, np.ascontiguousarray(b, np.float64).ctypes.data_as(C.c_void_p),
                                    x.ctypes.data_as(C.c_void_p), float(tolerance), int(max_iterations),
                                    hist.ctypes.data_as(C.c_void_p), len(hist), C.byref(r)))
        return x, r.iterations, r.error, hist[:min(r.history, len(hist))]

    def solve_device(self, b: DeviceVector, x: DeviceVector, tolerance: float = 1e-9,
                     max_iterations: int = 1000):
        """the same with b, x in HBM: (iterations, error, converged)"""
        r = _Result()
        check(lib().mk_box_mg_solve_device(self._h, b.ptr, x.ptr, float(tolerance), int(max_iterations),
                                           None, 0, C.byref(r)))
        return r.iterations, r.error, bool(r.converged)

    def step_device(self, p: DeviceVector, crank_nicolson: bool, tolerance: float = 1e-9,
                    max_iterations: int = 1000):
        """one backward-Euler or Crank-Nicolson step in place on p (HBM); (iterations, error,
        converged); p is left unchanged when the solve does not converge"""
        r = _Result()
        check(lib().mk_box_mg_step_device(self._h, p.ptr, 1 if crank_nicolson else 0,
                                          float(tolerance), int(max_iterations), C.byref(r)))
        return r.iterations, r.error, bool(r.converged)

    def gmres(self, b, x, tolerance: float, max_iterations: int = 1000, restart: int = 30,
              orthogonalization: str = "dcgs2") -> SolverResult:
        """flexible GMRES on (shift I - Q) x = b preconditioned by one cycle (the driver's
        mg-gmres; its tolerance is tolerance * shift / sqrt(rows)); x (numpy array or
        DeviceVector) is the guess on input"""
        from .api import ORTHOGONALIZATION
        opts = _GmresOptions(float(tolerance), int(max_iterations), int(restart),
                             ORTHOGONALIZATION[orthogonalization])
        res = _SolverResult()
        g = self._gen._h
        if isinstance(b, DeviceVector):
            check(lib().mk_gmres_box_device(g, self.shift, self._h, b.ptr, x.ptr, C.byref(opts),
                                            C.byref(res)))
        else:
            ba = np.ascontiguousarray(b, np.float64)
            if x.dtype != np.float64 or not x.flags.c_contiguous:
                raise TypeError("x must be a contiguous float64 array")
            check(lib().mk_gmres_box(g, self.shift, self._h, ba.ctypes.data_as(C.c_void_p),
                                     x.ctypes.data_as(C.c_void_p), C.byref(opts), C.byref(res)))
        return SolverResult(res.iterations, res.residual, bool(res.converged))


def small_levels_enable(on: bool) -> None:
    """run small levels' smoothing and transfers in two single-workgroup kernels per visit
    (default on)"""
    check(lib().mk_box_mg_small_enable(1 if on else 0))


def graphs_enable(on: bool) -> None:
    """replay the standalone iteration's cycle from a captured HIP graph (default on)"""
    check(lib().mk_box_mg_graphs_enable(1 if on else 0))


# ---------------------------------------------------------------------- time stepping and FSP
# The two drivers below follow run() and run_adaptive_fsp() of rdme_full_multigrid.cpp. They only
# need a solver object with rows(), shift, cap, solve(b, x0, tolerance, max_iterations) ->
# (x, iterations, error, history) and apply_operator(x) = (shift I - Q) x: the device
# BoxMultigrid, or the oracle's restatement on the CPU.

HIGH_POPULATION = 21  # the driver's high_mass threshold (rdme_full_multigrid.cpp:1227)


def population_offsets(voxels: int, cap: int) -> np.ndarray:
    """first state index of every total population 0..cap, and the size (box order)"""
    from math import comb
    return np.array([comb(k + voxels - 1, voxels) if k > 0 else 0 for k in range(cap + 2)],
                    np.int64)


def marginal(p, voxels: int, cap: int) -> np.ndarray:
    """probability of each total population 0..cap"""
    off = population_offsets(voxels, cap)
    return np.add.reduceat(np.asarray(p), off[:-1]) if len(p) else np.zeros(cap + 1)


def time_steps(M, p0, steps: int, integrator: str = "be", tolerance: float = 1e-9,
               max_iterations: int = 1000, on_step=None):
    """run() with --steps (rdme_full_multigrid.cpp:1174-1245): backward Euler steps of length
    1 / shift, (shift I - Q) p_next = shift p; or "cn": two backward-Euler half steps (Rannacher
    start) and then Crank-Nicolson steps of 2 / shift, (shift I - Q) p_next = 2 shift p -
    (shift I - Q) p; steps + 1 solves. Each solve starts from the previous distribution. Stops
    at the first solve that does not reach the tolerance. Returns (p, records); on_step(record,
    p) is called after every solve."""
    if integrator not in ("be", "cn"):
        raise ValueError("integrator must be 'be' or 'cn'")
    s = M.shift
    solves = steps + (1 if integrator == "cn" else 0)
    p = np.array(p0, np.float64)
    pop = np.repeat(np.arange(M.cap + 1), np.diff(population_offsets(M.voxels, M.cap)))
    records = []
    for step in range(1, solves + 1):
        if integrator == "cn" and step > 2:
            b = 2 * s * p - M.apply_operator(p)
        else:
            b = s * p
        x, it, err, _ = M.solve(b, p, tolerance, max_iterations)
        t = ((1.0 if step == 1 else 2.0 * (step - 1)) if integrator == "cn" else float(step)) / s
        rec = {"step": step, "time": t, "iterations": it, "error": err,
               "converged": bool(err < tolerance), "mass": float(x.sum()),
               "high_mass": float(x[pop >= HIGH_POPULATION].sum()), "min_probability": float(x.min())}
        records.append(rec)
        if on_step:
            on_step(rec, x)
        if not rec["converged"]:
            break
        p = x
    return p, records


def adaptive_fsp(make, voxels: int, initial: int, shift: float, steps: int,
                 fsp_tolerance: float = 1e-6, tolerance: float = 1e-9, cap: Optional[int] = None,
                 max_iterations: int = 1000, on_attempt=None):
    """run_adaptive_fsp() (rdme_full_multigrid.cpp:1248-1363): backward-Euler steps of 1 / shift
    on a total-population box whose cap grows until a step loses little enough probability.
    make(cap) builds the solver for a cap (the hierarchy is rebuilt on growth: device setup is
    fast, and the state order keeps every old state's index, so the distribution is extended by
    zeros). A trial is accepted when missing mass + 2 (accumulated + this solve's l1 error
    bound) <= fsp_tolerance * step / steps; otherwise the cap grows to max(cap + 4,
    ceil(1.25 (cap + 1))) and the step is retried from the accepted distribution. Returns (p,
    cap, attempts)."""
    import math
    if cap is None:
        cap = max(2, initial)
    solve_tol = min(tolerance, fsp_tolerance / (100.0 * steps))
    M = None
    p = None
    acc = 0.0
    attempts = []
    for step in range(1, steps + 1):
        accepted = False
        while not accepted:
            if M is None or M.cap != cap:
                M = make(cap)
                q = np.zeros(M.rows())
                if p is None:
                    x0 = np.zeros(voxels, np.int64)
                    for k in range(initial):
                        x0[k % voxels] += 1
                    if initial > cap:
                        raise ValueError("the initial population exceeds the cap")
                    q[_rank(x0, voxels)] = 1.0
                else:
                    q[:len(p)] = p
                p = q
            b = shift * p
            trial, it, err, _ = M.solve(b, p, solve_tol, max_iterations)
            if not (err <= solve_tol) or not np.isfinite(err):
                raise RuntimeError("adaptive FSP linear solve failed; no step was accepted")
            mass = float(trial.sum())
            negative = float(-trial[trial < 0].sum())
            if not np.isfinite(mass) or negative > err + 1e-13:
                raise RuntimeError("adaptive FSP produced invalid probabilities")
            nxt = acc + err
            missing = max(0.0, 1.0 - mass)
            budget = fsp_tolerance * step / steps
            accepted = missing + 2 * nxt <= budget
            rec = {"attempt": len(attempts) + 1, "step": step, "time": step / shift, "cap": cap,
                   "states": M.rows(), "accepted": accepted, "iterations": it, "error": err,
                   "missing_mass": missing, "error_bound": nxt, "budget": budget}
            attempts.append(rec)
            if on_attempt:
                on_attempt(rec)
            if accepted:
                p = trial
                acc = nxt
            else:
                cap = max(cap + 4, math.ceil(1.25 * (cap + 1)))
                if cap > 10000:
                    raise RuntimeError("adaptive expansion exceeded the population guard")
    return p, cap, attempts


def _rank(x, voxels: int) -> int:
    """index of the composition x in the box order (total first, voxel 0 most significant)"""
    from math import comb
    total = int(sum(x))
    r = comb(total + voxels - 1, voxels) if total > 0 else 0
    left = total
    for v in range(voxels - 1):
        rest = voxels - v - 1
        r += comb(left + rest, rest) - comb(left - int(x[v]) + rest, rest)
        left -= int(x[v])
    return r
