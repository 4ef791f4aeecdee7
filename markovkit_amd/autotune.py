"""Autotuning of the resolvent solve's launch and sweep shapes by timing trial solves.

A shape is (pre_sweeps, post_sweeps, xcd_chunk):
- the sweep counts trade multigrid work per V-cycle against outer FGMRES iterations, whose
  Arnoldi cost grows with the square of their number;
- the XCD chunk is the number of consecutive row blocks each of the MI355X's 8 XCDs takes in
  turn. It sets how neighbouring rows share an L2.

The tuner builds the hierarchy of every sweep candidate, runs a warm-up solve and `repeats`
timed solves per XCD chunk with b and x in HBM, and keeps the fastest shape. The solve is the
one being tuned, so no model of the hardware is assumed. Ties are broken towards the earlier
candidate; the reference's defaults (2, 2) come first.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

from ._lib import check, lib

SWEEPS: Tuple[Tuple[int, int], ...] = ((2, 2), (1, 2), (2, 1), (1, 1), (3, 3))
# (coarse_cycles, cycle_levels): the V-cycle, and W-cycles on the top levels over a V-cycle tail
CYCLES: Tuple[Tuple[int, int], ...] = ((1, 0), (2, 2), (2, 3))
XCD_CHUNKS: Tuple[int, ...] = (16, 0, 8, 32)


def xcd_chunk() -> int:
    c = C.c_int32()
    check(lib().mk_get_xcd_chunk(C.byref(c)))
    return c.value


def set_xcd_chunk(chunk: int) -> None:
    check(lib().mk_set_xcd_chunk(int(chunk)))


@dataclass
class Trial:
    pre: int
    post: int
    xcd_chunk: int
    seconds: float
    iterations: int
    coarse_cycles: int = 1
    cycle_levels: int = 0
    smoother: int = None


@dataclass
class TuneResult:
    pre: int
    post: int
    xcd_chunk: int
    seconds: float
    trials: List[Trial] = field(default_factory=list)
    coarse_cycles: int = 1
    cycle_levels: int = 0
    smoother: int = None

    def options(self, **kw):
        """MultigridOptions with the tuned sweep counts, cycle shape and (when tuned) smoother;
        other fields from kw"""
        from .api import MultigridOptions
        if self.smoother is not None:
            kw.setdefault("smoother", self.smoother)
        return MultigridOptions(pre_sweeps=self.pre, post_sweeps=self.post,
                                coarse_cycles=self.coarse_cycles, cycle_levels=self.cycle_levels, **kw)


def tune_solve(Q, domain, box, shift: float, b, *, smoother=None, precision=None,
               sweeps: Sequence[Tuple[int, int]] = SWEEPS, chunks: Sequence[int] = XCD_CHUNKS,
               cycles: Sequence[Tuple[int, int]] = ((1, 0),),
               rel_tolerance: float = 1e-10, restart: int = 50, repeats: int = 2,
               apply: bool = True) -> TuneResult:
    """Time FGMRES(restart) + sector multigrid on (shift I - Q) x = b for every shape (sweep
    counts x cycle shapes (coarse_cycles, cycle_levels) x XCD chunks) and return the fastest.
    With apply=True the XCD chunk of the result is left set (the sweep counts and cycle shape go
    into the hierarchy options the caller builds)."""
    from . import api as mk
    smoother = mk.RED_BLACK if smoother is None else smoother
    precision = mk.FP32 if precision is None else precision
    b = np.ascontiguousarray(b, np.float64)
    tol = rel_tolerance * float(np.linalg.norm(b))
    n = box.size()
    b_dev = mk.DeviceVector(n, b)
    x_dev = mk.DeviceVector(n)
    A = mk.ShiftedGenerator(Q, shift)
    start_chunk = xcd_chunk()
    trials: List[Trial] = []
    try:
        for (pre, post), (cyc, lev) in [(sw, cy) for sw in sweeps for cy in cycles]:
            M = mk.SectorMultigrid(domain, box, shift,
                                   mk.MultigridOptions(precision=precision, smoother=smoother,
                                                       pre_sweeps=pre, post_sweeps=post,
                                                       coarse_cycles=cyc, cycle_levels=lev))
            for chunk in chunks:
                set_xcd_chunk(chunk)
                x_dev.zero()
                mk.gmres(A, M, b_dev, x_dev, tolerance=tol, restart=restart)  # warm-up
                best = float("inf")
                its = 0
                for _ in range(repeats):
                    x_dev.zero()
                    mk.synchronize()
                    t = time.perf_counter()
                    r = mk.gmres(A, M, b_dev, x_dev, tolerance=tol, restart=restart)
                    mk.synchronize()
                    best = min(best, time.perf_counter() - t)
                    if not r.converged:
                        best = float("inf")
                        break
                    its = r.iterations
                trials.append(Trial(pre, post, chunk, best, its, cyc, lev))
            del M
    finally:
        set_xcd_chunk(start_chunk)
    win = min(trials, key=lambda t: t.seconds)  # first minimum: earlier candidates win ties
    if apply:
        set_xcd_chunk(win.xcd_chunk)
    return TuneResult(win.pre, win.post, win.xcd_chunk, win.seconds, trials, win.coarse_cycles,
                      win.cycle_levels)


def tune_batch(Q, domain, box, shifts, B, *, precision=None, smoothers=None,
               sweeps: Sequence[Tuple[int, int]] = SWEEPS, chunks: Sequence[int] = (16,),
               rel_tolerance: float = 1e-10, restart: int = 50, repeats: int = 1,
               apply: bool = True) -> TuneResult:
    """The same search for a batched multi-shift solve (BatchSolver) over smoothers (Jacobi and
    red-black Gauss-Seidel by default) and sweep counts: every system to
    rel_tolerance * ||b_j||."""
    from . import api as mk
    precision = mk.FP32 if precision is None else precision
    smoothers = (mk.JACOBI, mk.RED_BLACK) if smoothers is None else tuple(smoothers)
    B = np.ascontiguousarray(B, np.float64)
    tols = rel_tolerance * np.linalg.norm(B, axis=1)
    start_chunk = xcd_chunk()
    trials: List[Trial] = []
    try:
        for smoother in smoothers:
            for pre, post in sweeps:
                M = mk.SectorMultigrid(domain, box, 1.0,
                                       mk.MultigridOptions(precision=precision, smoother=smoother,
                                                           pre_sweeps=pre, post_sweeps=post))
                solver = mk.BatchSolver(Q, M, shifts)
                for chunk in chunks:
                    set_xcd_chunk(chunk)
                    solver.solve(B, tolerances=tols, restart=restart)  # warm-up
                    best = float("inf")
                    its = 0
                    for _ in range(repeats):
                        mk.synchronize()
                        t = time.perf_counter()
                        _, res = solver.solve(B, tolerances=tols, restart=restart)
                        mk.synchronize()
                        best = min(best, time.perf_counter() - t)
                        if not all(r.converged for r in res):
                            best = float("inf")
                            break
                        its = max(r.iterations for r in res)
                    trials.append(Trial(pre, post, chunk, best, its, smoother=smoother))
                del solver, M
    finally:
        set_xcd_chunk(start_chunk)
    win = min(trials, key=lambda t: t.seconds)
    if apply:
        set_xcd_chunk(win.xcd_chunk)
    return TuneResult(win.pre, win.post, win.xcd_chunk, win.seconds, trials, smoother=win.smoother)
