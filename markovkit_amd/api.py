"""Python mirror of MarkovKit's resolvent-solve interface over the HIP library's C ABI.

Names, argument meaning and error behaviour follow the reference (paths relative to the
markovkit repository root):

* ``PopulationBox``      population_box        include/markovkit/rdme/population_space.hpp:131-196
* ``Generator``          generator             include/markovkit/generator.hpp:19-112
* ``ShiftedGenerator``   shifted_generator     include/markovkit/generator.hpp:116-133
* ``SectorMultigrid``    sector_multigrid<D>   include/markovkit/rdme/multigrid.hpp:371-408
* ``gmres``              num::gmres            (numerics; call sites tests/test_rdme.cpp:242-243)

Vectors are numpy arrays on the host (copied to and from HBM per call) or ``DeviceVector``
objects that stay resident in HBM.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._lib import ExpvOptions as _ExpvOptions
from ._lib import ExpvResult as _ExpvResult
from ._lib import GmresOptions as _GmresOptions
from ._lib import ServerResult as _ServerResult
from ._lib import MultigridOptions as _MultigridOptions
from ._lib import SolverResult as _SolverResult
from ._lib import check, lib
from .model import Grid, ReactionDiffusion, VoxelDomain

REFLECTING = 0
ABSORBING = 1
JACOBI = 0
AGGREGATE_BLOCK = 1
RED_BLACK = 2  # multicolour Gauss-Seidel
ILU0 = 3       # level-scheduled ILU(0) in rank order
FP64 = 0
FP32 = 1


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class PopulationBox:
    """population_box(n_voxels, max_totals)."""

    def __init__(self, n_voxels: int, max_totals: Sequence[int]):
        self.max_totals = np.ascontiguousarray(max_totals, np.int32)
        self._h = C.c_void_p()
        check(lib().mk_box_create(int(n_voxels), len(self.max_totals), _p(self.max_totals),
                                  C.byref(self._h)))
        self.voxels = int(n_voxels)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mk_box_destroy(self._h)
            self._h = None

    def size(self) -> int:
        v = C.c_int64()
        check(lib().mk_box_size(self._h, C.byref(v)))
        return v.value

    def n_voxels(self) -> int:
        return self.voxels

    def n_species(self) -> int:
        return len(self.max_totals)

    def n_sectors(self) -> int:
        v = C.c_int64()
        check(lib().mk_box_n_sectors(self._h, C.byref(v)))
        return v.value

    def _sector(self, k):
        b, e = C.c_int64(), C.c_int64()
        t = np.zeros(self.n_species(), np.int32)
        check(lib().mk_box_sector(self._h, int(k), C.byref(b), C.byref(e), _p(t)))
        return b.value, e.value, t

    def sector_begin(self, k: int) -> int:
        return self._sector(k)[0]

    def sector_end(self, k: int) -> int:
        return self._sector(k)[1]

    def sector_totals(self, k: int):
        return list(self._sector(k)[2])

    def sector_of(self, i: int) -> int:
        lo, hi = 0, self.n_sectors() - 1
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if self.sector_begin(mid) <= i:
                lo = mid
            else:
                hi = mid - 1
        return lo

    def __len__(self):
        return self.size()

    def __getitem__(self, i: int):
        x = np.zeros(self.voxels * self.n_species(), np.int32)
        check(lib().mk_box_state(self._h, int(i), _p(x)))
        return list(x)

    def find(self, x) -> Optional[int]:
        arr = np.ascontiguousarray(x, np.int32)
        if arr.size != self.voxels * self.n_species():
            return None
        v = C.c_int64()
        check(lib().mk_box_find(self._h, _p(arr), C.byref(v)))
        return None if v.value < 0 else v.value

    def find_sector(self, totals) -> Optional[int]:
        k = 0
        for s, t in enumerate(totals):
            if t < 0 or t > self.max_totals[s]:
                return None
            k = k * (int(self.max_totals[s]) + 1) + int(t)
        return k


class PopulationSector(PopulationBox):
    """population_sector(n_voxels, totals): the states with exactly totals[s] molecules of each
    species, in rank order (a space of one sector; every box query applies)."""

    def __init__(self, n_voxels: int, totals: Sequence[int]):
        self.max_totals = np.ascontiguousarray(totals, np.int32)
        self.totals = [int(t) for t in totals]
        self._h = C.c_void_p()
        check(lib().mk_sector_space_create(int(n_voxels), len(self.max_totals), _p(self.max_totals),
                                           C.byref(self._h)))
        self.voxels = int(n_voxels)


class _Domain:
    """voxel_domain handle on the library side."""

    def __init__(self, domain: VoxelDomain):
        self.domain = domain
        self._h = C.c_void_p()
        diff = np.ascontiguousarray(domain.diffusivities, np.float64)
        vols = np.ascontiguousarray(domain.volumes, np.float64)
        g = domain.graph
        if isinstance(g, Grid):
            dims = np.asarray(g.dims, np.int64)
            spacing = np.asarray(g.spacing, np.float64)
            check(lib().mk_domain_create_grid(g.dim, _p(dims), _p(spacing), _p(vols), len(diff),
                                              _p(diff), C.byref(self._h)))
        else:
            off, to, w = g.adjacency()
            check(lib().mk_domain_create_graph(g.n_vertices(), _p(off), _p(to), _p(w), _p(vols),
                                               len(diff), _p(diff), C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mk_domain_destroy(self._h)
            self._h = None


class DeviceVector:
    """A float64 vector resident in HBM (allocated by the library)."""

    def __init__(self, n: int, host: Optional[np.ndarray] = None):
        self.n = int(n)
        self._p = C.c_void_p()
        check(lib().mk_device_alloc(self.n * 8, C.byref(self._p)))
        if host is not None:
            self.upload(host)

    def upload(self, host: np.ndarray):
        a = np.ascontiguousarray(host, np.float64)
        assert a.size == self.n
        check(lib().mk_copy_to_device(self._p, _p(a), self.n * 8))

    def zero(self):
        check(lib().mk_device_zero(self._p, self.n * 8))

    def download(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        check(lib().mk_copy_to_host(_p(out), self._p, self.n * 8))
        return out

    @property
    def ptr(self):
        return self._p

    def __del__(self):
        if getattr(self, "_p", None):
            lib().mk_device_free(self._p)
            self._p = None


class Generator:
    """generator(space, model, boundary): Q assembled by destination and resident in HBM."""

    def __init__(self, box: PopulationBox, model: ReactionDiffusion, boundary: int = ABSORBING):
        self.box = box
        self.model = model
        self._domain = _Domain(model.domain)
        changes = np.ascontiguousarray(model.changes(), np.int32)
        rates = np.ascontiguousarray(model.rate_tables(box.max_totals), np.float64)
        self._h = C.c_void_p()
        check(lib().mk_generator_create_rdme(box._h, self._domain._h, len(model.reactions),
                                             _p(changes), _p(rates), int(boundary),
                                             C.byref(self._h)))
        r, j = C.c_int64(), C.c_int64()
        check(lib().mk_generator_size(self._h, C.byref(r), C.byref(j)))
        self._rows, self._jumps = r.value, j.value

    @classmethod
    def from_jumps(cls, first, to, rate, exit_rates, leak_rates):
        """A generator from source-major jump lists (generator.hpp:109-112 layout)."""
        self = cls.__new__(cls)
        self.box = None
        self.model = None
        first = np.ascontiguousarray(first, np.int64)
        to = np.ascontiguousarray(to, np.int64)
        rate = np.ascontiguousarray(rate, np.float64)
        e = np.ascontiguousarray(exit_rates, np.float64)
        l = np.ascontiguousarray(leak_rates, np.float64)
        self._h = C.c_void_p()
        check(lib().mk_generator_create_jumps(len(e), _p(first), _p(to), _p(rate), _p(e), _p(l),
                                              C.byref(self._h)))
        r, j = C.c_int64(), C.c_int64()
        check(lib().mk_generator_size(self._h, C.byref(r), C.byref(j)))
        self._rows, self._jumps = r.value, j.value
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mk_generator_destroy(self._h)
            self._h = None

    def rows(self) -> int:
        return self._rows

    def cols(self) -> int:
        return self._rows

    def n_jumps(self) -> int:
        return self._jumps

    def _rates(self):
        e = np.zeros(self._rows)
        l = np.zeros(self._rows)
        check(lib().mk_generator_rates(self._h, _p(e), _p(l)))
        return e, l

    def entries(self):
        """(ptr, col, rate): the stored jumps by destination (row i: sources j and rates of
        j -> i, sources ascending, jumps of one source to one destination summed)."""
        e = C.c_int64()
        check(lib().mk_generator_entries(self._h, C.byref(e), None, None, None))
        ptr = np.zeros(self.rows() + 1, np.int64)
        col = np.zeros(e.value, np.int64)
        rate = np.zeros(e.value)
        check(lib().mk_generator_entries(self._h, C.byref(e), _p(ptr), _p(col), _p(rate)))
        return ptr, col, rate

    def exit_rates(self) -> np.ndarray:
        return self._rates()[0]

    def leak_rates(self) -> np.ndarray:
        return self._rates()[1]

    def apply(self, x) -> np.ndarray:
        xa = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self._rows)
        check(lib().mk_generator_apply(self._h, _p(xa), _p(y)))
        return y

    def apply_shifted(self, shift: float, x) -> np.ndarray:
        xa = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self._rows)
        check(lib().mk_shifted_apply(self._h, float(shift), _p(xa), _p(y)))
        return y

    def apply_shifted_device(self, shift: float, x: DeviceVector, y: DeviceVector):
        check(lib().mk_shifted_apply_device(self._h, float(shift), x.ptr, y.ptr))


class ShiftedGenerator:
    """shifted_generator(Q, shift): the operator shift I - Q; it reads Q by reference."""

    def __init__(self, Q: Generator, shift: float):
        self.Q = Q
        self._shift = float(shift)

    def rows(self) -> int:
        return self.Q.rows()

    def cols(self) -> int:
        return self.Q.rows()

    def shift(self) -> float:
        return self._shift

    def apply(self, x) -> np.ndarray:
        return self.Q.apply_shifted(self._shift, x)


@dataclass
class MultigridOptions:
    """multigrid_options (multigrid.hpp:125-134) plus the device working precision."""
    pre_sweeps: int = 2
    post_sweeps: int = 2
    jacobi_weight: float = 2.0 / 3.0
    coarsest_size: int = 4
    max_levels: int = 32
    coarse_cycles: int = 1
    sweep_growth: int = 0
    smoother: int = JACOBI
    precision: int = FP64
    cycle_levels: int = 0  # 0: coarse_cycles on every level; k: on the top k levels only

    def to_c(self) -> _MultigridOptions:
        return _MultigridOptions(self.pre_sweeps, self.post_sweeps, self.jacobi_weight,
                                 self.coarsest_size, self.max_levels, self.coarse_cycles,
                                 self.sweep_growth, self.smoother, self.precision,
                                 self.cycle_levels)


class SectorMultigrid:
    """sector_multigrid(domain, box, shift, options): one population multigrid cycle per
    sector, all sectors of a level in one launch."""

    def __init__(self, domain: VoxelDomain, box: PopulationBox, shift: float,
                 options: Optional[MultigridOptions] = None):
        self.options = options or MultigridOptions()
        self._domain = _Domain(domain)
        self.box = box
        self._shift = float(shift)
        self._h = C.c_void_p()
        opts = self.options.to_c()
        check(lib().mk_sector_mg_create(self._domain._h, box._h, float(shift), C.byref(opts),
                                        C.byref(self._h)))
        r, k, L = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().mk_sector_mg_info(self._h, C.byref(r), C.byref(k), C.byref(L)))
        self._rows, self._sectors, self._levels = r.value, k.value, L.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mk_sector_mg_destroy(self._h)
            self._h = None

    def rows(self) -> int:
        return self._rows

    def cols(self) -> int:
        return self._rows

    def n_sectors(self) -> int:
        return self._sectors

    def max_levels(self) -> int:
        return self._levels

    def shift(self) -> float:
        return self._shift

    def set_shift(self, shift: float) -> None:
        """Re-target the hierarchy to another shift (terminal blocks re-inverted; Jacobi only)."""
        check(lib().mk_sector_mg_set_shift(self._h, float(shift)))
        self._shift = float(shift)

    def apply(self, b) -> np.ndarray:
        ba = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self._rows)
        check(lib().mk_sector_mg_apply(self._h, _p(ba), _p(x)))
        return x

    def apply_device(self, b: DeviceVector, x: DeviceVector):
        check(lib().mk_sector_mg_apply_device(self._h, b.ptr, x.ptr))

    def relax(self, rhs, x, sweeps: int = 1, reverse: bool = False) -> np.ndarray:
        """`sweeps` smoother sweeps on the finest level from the guess x (returned as a new
        array), on the device in the hierarchy's precision."""
        ra = np.ascontiguousarray(rhs, np.float64)
        xa = np.array(x, np.float64, copy=True)
        check(lib().mk_sector_mg_relax(self._h, _p(ra), _p(xa), int(sweeps), int(bool(reverse))))
        return xa

    def level(self, level: int):
        """(rows, jumps, terminal) of a level of a hierarchy over one population_sector."""
        r, j, t = C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().mk_sector_mg_level(self._h, int(level), C.byref(r), C.byref(j), C.byref(t)))
        return r.value, j.value, bool(t.value)

    def transfer(self, level: int = 0):
        """(coarse_of, weight) from level `level` to the next of a hierarchy over one sector."""
        n = self.level(level)[0]
        co = np.zeros(n, np.int64)
        w = np.zeros(n)
        check(lib().mk_sector_mg_transfer(self._h, int(level), _p(co), _p(w)))
        return co, w


@dataclass
class SolverResult:
    iterations: int
    residual: float
    converged: bool


ORTHOGONALIZATION = {"dcgs2": 0, "cgs2": 1}


def gmres(A: ShiftedGenerator, *args, tolerance: float = 1e-10, max_iterations: int = 1000,
          restart: int = 30, orthogonalization: str = "dcgs2") -> SolverResult:
    """num::gmres(A, [M,] b, x, options): x (numpy array or DeviceVector) is the initial
    guess on input and the solution on output; the tolerance is an absolute Euclidean
    residual norm; non-convergence is reported, not raised. orthogonalization "dcgs2"
    (delayed CGS2, one reduction per step; preconditioned solves with restart <= 64) or
    "cgs2"."""
    if len(args) == 3:
        M, b, x = args
    elif len(args) == 2:
        M = None
        b, x = args
    else:
        raise TypeError("gmres(A, [M,] b, x, ...)")
    if orthogonalization not in ORTHOGONALIZATION:
        raise ValueError(f"orthogonalization must be one of {sorted(ORTHOGONALIZATION)}")
    opts = _GmresOptions(float(tolerance), int(max_iterations), int(restart),
                         ORTHOGONALIZATION[orthogonalization])
    res = _SolverResult()
    mh = M._h if M is not None else None
    if isinstance(b, DeviceVector):
        check(lib().mk_gmres_device(A.Q._h, A.shift(), mh, b.ptr, x.ptr, C.byref(opts),
                                    C.byref(res)))
    else:
        ba = np.ascontiguousarray(b, np.float64)
        if x.dtype != np.float64 or not x.flags.c_contiguous:
            raise TypeError("x must be a contiguous float64 array")
        check(lib().mk_gmres(A.Q._h, A.shift(), mh, _p(ba), _p(x), C.byref(opts), C.byref(res)))
    return SolverResult(res.iterations, res.residual, bool(res.converged))


@dataclass
class ExpvResult:
    steps: int
    rejected: int
    error: float


def expv(Q: Generator, t: float, v, krylov_dim: int = 30, tolerance: float = 1e-8):
    """num::expv(t, Q.matrix(), v, krylov_dim, tolerance) on the device (fsp.hpp:152, :225):
    w = exp(t Q) v by Sidje's Krylov expv for a host vector v; returns (w, ExpvResult)."""
    va = np.ascontiguousarray(v, np.float64)
    if va.shape != (Q.rows(),):
        raise ValueError("v must have one entry per state")
    w = np.zeros_like(va)
    o = _ExpvOptions(int(krylov_dim), float(tolerance))
    r = _ExpvResult()
    check(lib().mk_expv(Q._h, float(t), _p(va), _p(w), C.byref(o), C.byref(r)))
    return w, ExpvResult(r.steps, r.rejected, r.error)


def expv_device(Q: Generator, t: float, v: "DeviceVector", w: "DeviceVector",
                krylov_dim: int = 30, tolerance: float = 1e-8) -> ExpvResult:
    """w = exp(t Q) v with both vectors in HBM (they may be the same vector)."""
    o = _ExpvOptions(int(krylov_dim), float(tolerance))
    r = _ExpvResult()
    check(lib().mk_expv_device(Q._h, float(t), v.ptr, w.ptr, C.byref(o), C.byref(r)))
    return ExpvResult(r.steps, r.rejected, r.error)


def expv_batch(Q: Generator, ts, V, krylov_dim: int = 30, tolerance: float = 1e-8):
    """w_j = exp(t_j Q) v_j for the rows v_j of V (count x states) advanced together on the
    device (mk_expv_batch): one Arnoldi process per substep for all vectors, Expokit's step
    control per vector. Returns (W, [ExpvResult])."""
    Va = np.ascontiguousarray(np.atleast_2d(V), np.float64)
    ta = np.ascontiguousarray(np.atleast_1d(ts), np.float64)
    if Va.shape[1] != Q.rows() or ta.shape != (Va.shape[0],):
        raise ValueError("V must hold one vector per time, each with one entry per state")
    W = np.zeros_like(Va)
    o = _ExpvOptions(int(krylov_dim), float(tolerance))
    r = (_ExpvResult * max(1, len(ta)))()
    check(lib().mk_expv_batch(Q._h, len(ta), _p(ta), _p(Va), _p(W), C.byref(o), r))
    return W, [ExpvResult(x.steps, x.rejected, x.error) for x in r[:len(ta)]]


class BatchSolver:
    """Batched resolvent solves (s_j I - Q) x_j = b_j, j < len(shifts), over one population box:
    the generator and the sector multigrid hierarchy are shared, the terminal blocks are
    re-inverted per distinct shift, and every system runs flexible GMRES with delayed CGS2,
    stopping on its own tolerance (mk_batch_create / mk_batch_solve). The multigrid's
    precision selects the preconditioner; its own shift is not used."""

    def __init__(self, Q: Generator, M: SectorMultigrid, shifts: Sequence[float], rows=None,
                 block: int = 1):
        """rows (optional): per system, the leading rows it has (the end of a sector of the
        box; for an absorbing one-species box, the size of the box of a smaller cap).
        block = 4: every 4 consecutive systems (one shift, one size) share a block Krylov basis
        for their first restart cycle, projected by block Gram-Schmidt on matrix cores."""
        self.shifts = np.ascontiguousarray(shifts, np.float64)
        self._Q, self._M = Q, M  # the batch reads both handles
        self._h = C.c_void_p()
        self.rows = None if rows is None else np.ascontiguousarray(rows, np.int64)
        if self.rows is not None and self.rows.shape != self.shifts.shape:
            raise ValueError("one row count per system")
        check(lib().mk_batch_create_sized(Q._h, M._h, len(self.shifts), _p(self.shifts),
                                          _p(self.rows) if self.rows is not None else None,
                                          C.byref(self._h)))
        if block != 1:
            check(lib().mk_batch_set_block(self._h, int(block)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mk_batch_destroy(self._h)
            self._h = None

    def __len__(self) -> int:
        return len(self.shifts)

    def solve(self, B, X=None, *, tolerance: float = 1e-10, tolerances=None,
              max_iterations: int = 1000, restart: int = 50):
        """B: (n_systems, rows) right-hand sides; X: initial guesses of the same shape
        (zeros if None), overwritten with the solutions. tolerance / tolerances: absolute
        residual norm for all / each system. Returns (X, [SolverResult per system])."""
        n = self._Q.rows()
        k = len(self.shifts)
        Ba = np.ascontiguousarray(B, np.float64)
        if Ba.shape != (k, n):
            raise ValueError(f"B must have shape ({k}, {n})")
        if X is None:
            X = np.zeros((k, n))
        if X.dtype != np.float64 or not X.flags.c_contiguous or X.shape != (k, n):
            raise TypeError("X must be a contiguous float64 array of B's shape")
        tol = None
        if tolerances is not None:
            tol = np.ascontiguousarray(tolerances, np.float64)
            if tol.shape != (k,):
                raise ValueError("one tolerance per system")
        opts = _GmresOptions(float(tolerance), int(max_iterations), int(restart), 0)
        res = (_SolverResult * k)()
        check(lib().mk_batch_solve(self._h, _p(Ba), _p(X), _p(tol) if tol is not None else None,
                                   C.byref(opts), res))
        return X, [SolverResult(r.iterations, r.residual, bool(r.converged)) for r in res]


@dataclass
class ServedResult:
    kind: str          # "resolvent" or "expv"
    solve: Optional[SolverResult]
    expv: Optional[ExpvResult]
    queued_s: float
    run_s: float


class Server:
    """Request serving (mk_server_*). Resolvent families of at most batch_rows states are
    served by continuous batching: a resident batch engine of batch_slots systems admits queued
    requests at every restart boundary; Krylov exponentials over a generator of at most
    batch_rows states join a batched expv engine of that generator at substep boundaries.
    Larger families (a generator and a pool of Jacobi or red-black hierarchies, re-targeted to
    each request's shift) and larger exponentials run on worker threads with their own HIP
    streams, admitted first come first served by free worker, HBM footprint and free hierarchy.
    batch_rows = 0 turns batching off."""

    def __init__(self, workers: int = 4, hbm_budget: int = 64 << 30, batch_rows: int = None,
                 batch_slots: int = 64):
        self._h = C.c_void_p()
        check(lib().mk_server_create(int(workers), int(hbm_budget), C.byref(self._h)))
        if batch_rows is not None:
            check(lib().mk_server_batching(self._h, int(batch_rows), int(batch_slots)))
        self._keep = []
        self._rows = {}
        self._family_rows = []

    def close(self):
        if getattr(self, "_h", None):
            lib().mk_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_family(self, Q: Generator, hierarchies: Sequence[SectorMultigrid]) -> int:
        arr = (C.c_void_p * len(hierarchies))(*[m._h.value for m in hierarchies])
        fam = C.c_int32()
        check(lib().mk_server_add_family(self._h, Q._h, len(hierarchies), arr, C.byref(fam)))
        self._keep.append((Q, list(hierarchies)))
        self._family_rows.append(Q.rows())
        return fam.value

    def submit_resolvent(self, family: int, shift: float, b, tolerance: float,
                         restart: int = 50, max_iterations: int = 1000) -> int:
        """(shift I - Q) x = b; b has one entry per state of the family's box, or fewer: the
        leading rows of whole sectors (for an absorbing one-species box, a box of a smaller
        cap; batched families only)."""
        ba = np.ascontiguousarray(b, np.float64)
        if ba.ndim != 1 or not 0 < len(ba) <= self._family_rows[family]:
            raise ValueError("b must have at most one entry per state of the family")
        o = _GmresOptions(0.0, int(max_iterations), int(restart), 0)
        rid = C.c_int64()
        check(lib().mk_server_submit_resolvent_sized(self._h, int(family), float(shift), _p(ba),
                                                     len(ba), float(tolerance), C.byref(o),
                                                     C.byref(rid)))
        self._rows[rid.value] = len(ba)
        return rid.value

    def submit_expv(self, Q: Generator, t: float, v, krylov_dim: int = 30,
                    tolerance: float = 1e-8) -> int:
        va = np.ascontiguousarray(v, np.float64)
        o = _ExpvOptions(int(krylov_dim), float(tolerance))
        rid = C.c_int64()
        check(lib().mk_server_submit_expv(self._h, Q._h, float(t), _p(va), C.byref(o),
                                          C.byref(rid)))
        self._keep.append((Q,))
        self._rows[rid.value] = len(va)
        return rid.value

    def wait(self, rid: int):
        """Block until request rid is done; returns (solution, ServedResult)."""
        out = np.zeros(self._rows.pop(rid))
        r = _ServerResult()
        check(lib().mk_server_wait(self._h, int(rid), _p(out), C.byref(r)))
        if r.kind == 0:
            return out, ServedResult("resolvent", SolverResult(r.solve.iterations, r.solve.residual,
                                                               bool(r.solve.converged)), None,
                                     r.queued_s, r.run_s)
        return out, ServedResult("expv", None, ExpvResult(r.expv.steps, r.expv.rejected,
                                                          r.expv.error), r.queued_s, r.run_s)


def batch_graphs_enable(on: bool = True) -> None:
    """Graph replay of the batched Arnoldi steps (default on)."""
    check(lib().mk_batch_graphs_enable(1 if on else 0))


def graphs_enable(on: bool = True) -> None:
    """Replay repeated solver iterations and expv steps as captured HIP graphs (default off)."""
    check(lib().mk_graphs_enable(1 if on else 0))


def device_setup_enable(on: bool = True) -> None:
    """Assemble RDME generators on the device (default) or on the host; both give the same
    operator bit for bit."""
    check(lib().mk_device_setup_enable(1 if on else 0))


def tail_lds_enable(on: bool = True) -> None:
    """Keep the fused tail's levels in LDS where they fit (default on); bitwise the same."""
    check(lib().mk_tail_lds_enable(1 if on else 0))


def fused_first_enable(on: bool = True) -> None:
    """Fuse a red-black level's first half-sweep into the kernel that makes its right-hand side
    (default on); bitwise the same."""
    check(lib().mk_fused_first_enable(1 if on else 0))


def coarse_fuse_enable(on: bool = True) -> None:
    """Solve a direct coarse level and prolong it into the level above in one launch (default
    on); bitwise the same."""
    check(lib().mk_coarse_fuse_enable(1 if on else 0))


def tiny_gmres_enable(on: bool = True) -> None:
    """Run a whole solve in one workgroup when the problem has at most 1024 states and its
    sector hierarchy fits the LDS tail (default on); the same iteration, sums in another order."""
    check(lib().mk_tiny_gmres_enable(1 if on else 0))


def fused_tail_enable(on: bool = True) -> None:
    """Run the small multigrid levels in one workgroup-wide kernel (default on)."""
    check(lib().mk_fused_tail_enable(1 if on else 0))


# the multigrid classes before gmres_dot count the finest level's launches; coarse_levels
# counts every launch below it
KERNEL_CLASSES = ("shifted_spmv", "smooth_first", "smooth_sweep", "residual", "restrict",
                  "prolong", "terminal", "gmres_dot", "gmres_update", "gmres_other",
                  "coarse_levels")


def profile_enable(on=True):
    """Time kernel launches with HIP events: True = every class, False = off, or an iterable
    of class names."""
    if on is True:
        mask = -1
    elif not on:
        mask = 0
    else:
        mask = 0
        for name in on:
            mask |= 1 << KERNEL_CLASSES.index(name)
    check(lib().mk_profile_enable(mask))


def profile_reset():
    check(lib().mk_profile_reset())


def profile_read():
    out = {}
    for k in KERNEL_CLASSES:
        n, ms, by = C.c_int64(), C.c_double(), C.c_double()
        check(lib().mk_profile_read(k.encode(), C.byref(n), C.byref(ms), C.byref(by)))
        out[k] = dict(launches=n.value, ms=ms.value, bytes=by.value)
    return out


def synchronize():
    check(lib().mk_synchronize())
