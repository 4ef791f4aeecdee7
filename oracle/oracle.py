"""ORACLE binding — test infrastructure only.

ctypes access to ``oracle/liboracle.so``, the FP64 CPU restatement of MarkovKit's RDME
resolvent path (``oracle/mk_oracle.hpp``; reference citations there). Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this module, and only
as the checker or the timed CPU baseline; the product (``markovkit_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.argtypes = [C.c_int, P, P, C.c_int64, P, P, P, P, C.c_int, P, P, C.c_int,
                                 P, P, C.c_int, C.POINTER(C.c_void_p)]
        L.orc_destroy.argtypes = [P]
        for name in ("orc_size", "orc_n_sectors", "orc_n_jumps"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = C.c_int64
        L.orc_sector_begins.argtypes = [P, i64p]
        L.orc_states.argtypes = [P, i32p]
        L.orc_find.argtypes = [P, i32p]
        L.orc_find.restype = C.c_int64
        L.orc_rates.argtypes = [P, f64p, f64p]
        L.orc_jumps.argtypes = [P, i64p, i64p, f64p]
        L.orc_apply.argtypes = [P, f64p, f64p]
        L.orc_apply_shifted.argtypes = [P, C.c_double, f64p, f64p]
        L.orc_mg_create.argtypes = [P, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int64,
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_void_p)]
        L.orc_mg_relax.argtypes = [P, C.c_int64, f64p, f64p, C.c_int, C.c_int]
        L.orc_mg_colours.argtypes = [P, C.c_int64, np.ctypeslib.ndpointer(np.uint8, flags="C")]
        L.orc_mg_ilu_pivots.argtypes = [P, C.c_int64, f64p]
        L.orc_expv_jumps.argtypes = [C.c_int64, i64p, i64p, f64p, f64p, C.c_double, f64p, f64p,
                                     C.c_int64, C.c_double, i64p, i64p]
        L.orc_mg_destroy.argtypes = [P]
        L.orc_mg_max_levels.argtypes = [P]
        L.orc_mg_max_levels.restype = C.c_int64
        L.orc_mg_apply.argtypes = [P, f64p, f64p]
        L.orc_mg_level_rows.argtypes = [P, i64p]
        L.orc_mg_transfer.argtypes = [P, C.c_int64, C.c_int64, i64p, f64p]
        L.orc_mg_level_apply.argtypes = [P, C.c_int64, C.c_int64, f64p, f64p]
        L.orc_gmres.argtypes = [P, P, C.c_double, f64p, f64p, C.c_double, C.c_int64, C.c_int64,
                                C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.orc_gmres.restype = C.c_int
        D = C.c_double
        L.orc_box_create.argtypes = [C.c_int, C.c_int, C.c_int, D, D, D, D, D, D, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int64, C.POINTER(C.c_void_p)]
        L.orc_box_destroy.argtypes = [P]
        for name in ("orc_box_rows", "orc_box_nnz", "orc_box_voxels"):
            getattr(L, name).argtypes = [P, C.c_int64]
            getattr(L, name).restype = C.c_int64
        L.orc_box_levels.argtypes = [P]
        L.orc_box_levels.restype = C.c_int64
        L.orc_box_terminal.argtypes = [P, C.c_int64]
        L.orc_box_rank.argtypes = [P, i32p]
        L.orc_box_rank.restype = C.c_int64
        L.orc_box_initial.argtypes = [P, C.c_int]
        L.orc_box_initial.restype = C.c_int64
        L.orc_box_matrix.argtypes = [P, C.c_int64, i64p, i64p, f64p]
        L.orc_box_leak.argtypes = [P, C.c_int64, f64p]
        L.orc_box_colours.argtypes = [P, C.c_int64, np.ctypeslib.ndpointer(np.uint8, flags="C")]
        L.orc_box_transfer.argtypes = [P, C.c_int64, i64p, f64p]
        L.orc_box_cycle.argtypes = [P, f64p, f64p]
        L.orc_box_smooth.argtypes = [P, C.c_int64, f64p, f64p, C.c_int]
        L.orc_box_solve.argtypes = [P, f64p, f64p, D, C.c_int, f64p, C.c_int64,
                                    C.POINTER(C.c_int), C.POINTER(D)]
        L.orc_box_solve.restype = C.c_int64
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        msg = lib().orc_last_error().decode()
        if status == 1:
            raise ValueError(msg)
        if status == 2:
            raise OverflowError(msg)
        raise RuntimeError(msg)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Problem:
    """population_box + reaction_diffusion + generator on the CPU (FP64)."""

    def __init__(self, model, max_totals, absorbing=True):
        from markovkit_amd.model import Grid
        L = lib()
        self.model = model
        self.max_totals = np.asarray(max_totals, np.int32)
        dom = model.domain
        S = dom.n_species()
        V = dom.n_voxels()
        self._keep = []
        if isinstance(dom.graph, Grid):
            g = dom.graph
            dims = np.asarray(g.dims, np.int64)
            spacing = np.asarray(g.spacing, np.float64)
            vols = np.ascontiguousarray(dom.volumes, np.float64)
            args = [g.dim, _ptr(dims), _ptr(spacing), V, None, None, None, _ptr(vols)]
            self._keep += [dims, spacing, vols]
        else:
            off, to, w = dom.graph.adjacency()
            vols = np.ascontiguousarray(dom.volumes, np.float64)
            args = [0, None, None, V, _ptr(off), _ptr(to), _ptr(w), _ptr(vols)]
            self._keep += [off, to, w, vols]
        diff = np.ascontiguousarray(dom.diffusivities, np.float64)
        changes = np.ascontiguousarray(model.changes(), np.int32)
        rates = np.ascontiguousarray(model.rate_tables(self.max_totals), np.float64)
        self._keep += [diff, changes, rates]
        h = C.c_void_p()
        _check(L.orc_create(*args, S, _ptr(diff), _ptr(self.max_totals), len(model.reactions),
                            _ptr(changes), _ptr(rates), 1 if absorbing else 0, C.byref(h)))
        self.h = h
        self.n = L.orc_size(h)
        self.S = S
        self.V = V

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def sector_begins(self):
        k = lib().orc_n_sectors(self.h)
        out = np.zeros(k + 1, np.int64)
        lib().orc_sector_begins(self.h, out)
        return out

    def states(self):
        out = np.zeros(self.n * self.V * self.S, np.int32)
        lib().orc_states(self.h, out)
        return out.reshape(self.n, self.V * self.S)

    def find(self, x):
        return lib().orc_find(self.h, np.ascontiguousarray(x, np.int32))

    def rates(self):
        e = np.zeros(self.n)
        l = np.zeros(self.n)
        lib().orc_rates(self.h, e, l)
        return e, l

    def jumps(self):
        m = lib().orc_n_jumps(self.h)
        first = np.zeros(self.n + 1, np.int64)
        to = np.zeros(m, np.int64)
        rate = np.zeros(m)
        lib().orc_jumps(self.h, first, to, rate)
        return first, to, rate

    def apply(self, x):
        y = np.zeros(self.n)
        lib().orc_apply(self.h, np.ascontiguousarray(x, np.float64), y)
        return y

    def apply_shifted(self, shift, x):
        y = np.zeros(self.n)
        lib().orc_apply_shifted(self.h, float(shift), np.ascontiguousarray(x, np.float64), y)
        return y

    def matrix(self, shift=0.0, sign=1.0):
        """sign * (Q - shift I) as a scipy CSR matrix (generator.hpp:93-107)."""
        import scipy.sparse as sp
        first, to, rate = self.jumps()
        exit_, _ = self.rates()
        cols = np.repeat(np.arange(self.n), np.diff(first))
        rows = np.concatenate([np.arange(self.n), to])
        cc = np.concatenate([np.arange(self.n), cols])
        vals = np.concatenate([-sign * (exit_ + shift), sign * rate])
        return sp.csr_matrix((vals, (rows, cc)), shape=(self.n, self.n))

    def gmres(self, shift, b, x0=None, mg=None, tol=1e-10, max_iterations=1000, restart=30):
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64)
        it = C.c_int64()
        res = C.c_double()
        conv = lib().orc_gmres(self.h, mg.h if mg is not None else None, float(shift),
                               np.ascontiguousarray(b, np.float64), x, float(tol),
                               int(max_iterations), int(restart), C.byref(it), C.byref(res))
        return x, dict(iterations=it.value, residual=res.value, converged=bool(conv))


class SectorMultigrid:
    """sector_multigrid on the CPU (smoother "jacobi", "aggregate_block", "red_black" or
    "ilu0")."""

    SMOOTHERS = ("jacobi", "aggregate_block", "red_black", "ilu0")

    def __init__(self, problem, shift, pre_sweeps=2, post_sweeps=2, jacobi_weight=2.0 / 3.0,
                 coarsest_size=4, max_levels=32, coarse_cycles=1, sweep_growth=0,
                 smoother="jacobi", cycle_levels=0):
        if smoother not in self.SMOOTHERS:
            raise ValueError(f"smoother must be one of {self.SMOOTHERS}")
        h = C.c_void_p()
        _check(lib().orc_mg_create(problem.h, float(shift), pre_sweeps, post_sweeps,
                                   jacobi_weight, coarsest_size, max_levels, coarse_cycles,
                                   sweep_growth, self.SMOOTHERS.index(smoother), cycle_levels,
                                   C.byref(h)))
        self.h = h
        self.problem = problem
        self.n = problem.n

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_destroy(self.h)
            self.h = None

    def max_levels(self):
        return lib().orc_mg_max_levels(self.h)

    def level_rows(self):
        k = len(self.problem.sector_begins()) - 1
        L = self.max_levels()
        out = np.zeros(k * L, np.int64)
        lib().orc_mg_level_rows(self.h, out)
        return out.reshape(k, L)

    def relax(self, k, rhs, x, sweeps=1, reverse=False):
        """population_multigrid::relax on sector k's finest level (multigrid.hpp:231-236):
        `sweeps` smoother sweeps on A x = rhs starting from x (updated in place)"""
        _check(lib().orc_mg_relax(self.h, k, np.ascontiguousarray(rhs, np.float64), x, sweeps,
                                  1 if reverse else 0))
        return x

    def colours(self, k):
        """red-black colours of sector k's finest level"""
        out = np.zeros(self.level_rows()[k, 0], np.uint8)
        _check(lib().orc_mg_colours(self.h, k, out))
        return out

    def ilu_pivots(self, k):
        """ILU(0) pivots of sector k's finest level (rank order)"""
        out = np.zeros(self.level_rows()[k, 0])
        _check(lib().orc_mg_ilu_pivots(self.h, k, out))
        return out

    def apply(self, b):
        x = np.zeros(self.n)
        lib().orc_mg_apply(self.h, np.ascontiguousarray(b, np.float64), x)
        return x

    def transfer(self, sector, level):
        rows = self.level_rows()[sector, level]
        co = np.zeros(rows, np.int64)
        w = np.zeros(rows)
        _check(lib().orc_mg_transfer(self.h, sector, level, co, w))
        return co, w

    def level_apply(self, sector, level, x):
        y = np.zeros(len(x))
        _check(lib().orc_mg_level_apply(self.h, sector, level,
                                        np.ascontiguousarray(x, np.float64), y))
        return y


def expv_jumps(first, to, rate, exit_, t, v, krylov_dim=30, tol=1e-8):
    """w = exp(t Q) v on the CPU (Sidje's expv restated) for a generator given by source-major
    jump lists; returns (w, steps, rejected)."""
    first = np.ascontiguousarray(first, np.int64)
    n = len(first) - 1
    w = np.zeros(n)
    steps, rej = np.zeros(1, np.int64), np.zeros(1, np.int64)
    _check(lib().orc_expv_jumps(n, first, np.ascontiguousarray(to, np.int64),
                                np.ascontiguousarray(rate, np.float64),
                                np.ascontiguousarray(exit_, np.float64), float(t),
                                np.ascontiguousarray(v, np.float64), w, int(krylov_dim), float(tol),
                                steps, rej))
    return w, int(steps[0]), int(rej[0])


class BoxMultigrid:
    """The paper driver's whole-box multigrid (box_mg.hpp) on the CPU (FP64).

    Parameters follow the driver's command line (rdme_full_multigrid.cpp:49-63): Schlogl
    constants c1..c4, smoother "gs" | "jacobi" | "block-ilu" (the driver's) or "colour-gs"
    (the GPU library's four-colour order), cycles 1 (V) or 2 (W), kcycle for the K-cycle, and
    terminal_rows > 0 for an exact dense solve at the first level with at most that many states.
    ``linear(scale)`` gives the driver's --linear rates (c1 = c2 = 0, c3 = 2 scale, c4 = scale).
    """

    SMOOTHERS = {"gs": 0, "jacobi": 1, "block-ilu": 2, "colour-gs": 3}
    DEFAULT_RATES = (0.0133, 0.0006, 0.6924125, 0.114725)

    @staticmethod
    def linear(scale=1.0):
        return (0.0, 0.0, 2.0 * scale, scale)

    def __init__(self, dimension=1, cells=4, cap=40, diffusion=1.0, shift=0.05, rates=None,
                 smoother="gs", sweeps=2, cycles=1, kcycle=False, terminal_rows=0):
        L = lib()
        c1, c2, c3, c4 = self.DEFAULT_RATES if rates is None else rates
        self.shift = shift
        h = C.c_void_p()
        _check(L.orc_box_create(dimension, cells, cap, diffusion, shift, c1, c2, c3, c4,
                                self.SMOOTHERS[smoother], sweeps, cycles, 1 if kcycle else 0,
                                terminal_rows, C.byref(h)))
        self.h = h
        self.n = L.orc_box_rows(h, 0)
        self.V = L.orc_box_voxels(h, 0)
        self.voxels = self.V
        self.cap = cap

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_box_destroy(self.h)
            self.h = None

    def levels(self):
        return lib().orc_box_levels(self.h)

    def rows(self, level=0):
        return lib().orc_box_rows(self.h, level)

    def terminal(self, level):
        return lib().orc_box_terminal(self.h, level)

    def rank(self, counts):
        return lib().orc_box_rank(self.h, np.ascontiguousarray(counts, np.int32))

    def initial_index(self, initial):
        return lib().orc_box_initial(self.h, initial)

    def matrix(self, level=0):
        """(ptr, col, val) of shift I - Q by destination on that level"""
        L = lib()
        n, nnz = L.orc_box_rows(self.h, level), L.orc_box_nnz(self.h, level)
        ptr = np.zeros(n + 1, np.int64)
        col = np.zeros(nnz, np.int64)
        val = np.zeros(nnz, np.float64)
        L.orc_box_matrix(self.h, level, ptr, col, val)
        return ptr, col, val

    def dense(self, level=0):
        ptr, col, val = self.matrix(level)
        n = len(ptr) - 1
        A = np.zeros((n, n))
        for i in range(n):
            A[i, col[ptr[i]:ptr[i + 1]]] = val[ptr[i]:ptr[i + 1]]
        return A

    def apply_level(self, level, x):
        import scipy.sparse as sp
        ptr, col, val = self.matrix(level)
        n = len(ptr) - 1
        return sp.csr_matrix((val, col, ptr), shape=(n, n)) @ np.asarray(x, np.float64)

    def apply_operator(self, x):
        """(shift I - Q) x on the finest level"""
        return self.apply_level(0, x)

    def leak(self, level=0):
        out = np.zeros(self.rows(level))
        lib().orc_box_leak(self.h, level, out)
        return out

    def colours(self, level=0):
        out = np.zeros(self.rows(level), np.uint8)
        lib().orc_box_colours(self.h, level, out)
        return out

    def transfer(self, level=0):
        n = self.rows(level)
        parent = np.zeros(n, np.int64)
        weight = np.zeros(n, np.float64)
        _check(lib().orc_box_transfer(self.h, level, parent, weight))
        return parent, weight

    def smooth(self, level, b, x, reverse=False):
        """the level's smoother (all its sweeps) from the guess x (returned)"""
        x = np.array(x, np.float64)
        _check(lib().orc_box_smooth(self.h, level, np.ascontiguousarray(b, np.float64), x,
                                    1 if reverse else 0))
        return x

    def cycle(self, b, x):
        """one cycle of A x = b from the guess x (returned; x is not modified)"""
        x = np.array(x, np.float64)
        lib().orc_box_cycle(self.h, np.ascontiguousarray(b, np.float64), x)
        return x

    def apply(self, b):
        return self.cycle(b, np.zeros(self.n))

    def solve(self, b, x0, tolerance=1e-9, max_iterations=1000):
        """the driver's standalone iteration: (x, iterations, error, history)"""
        x = np.array(x0, np.float64)
        hist = np.zeros(max_iterations + 2)
        it = C.c_int()
        err = C.c_double()
        m = lib().orc_box_solve(self.h, np.ascontiguousarray(b, np.float64), x, tolerance,
                                max_iterations, hist, len(hist), C.byref(it), C.byref(err))
        return x, it.value, err.value, hist[:m]

    def initial_problem(self, initial):
        """the driver's first step: b = shift p0, x0 = p0 with p0 the initial state"""
        p0 = np.zeros(self.n)
        p0[self.initial_index(initial)] = 1.0
        return self.shift * p0, p0
