"""The RDME resolvent workloads measured and tested by this package.

Each builder returns a ``Workload``: the model, the box caps, the shift(s) and right-hand
side(s). Models follow the reference's benchmark harness (birth ``0.05 * volume`` and death
``0.02 * n`` per voxel, examples/rdme_multigrid_benchmark.cpp:65-70) and its two-species
annihilation example (examples/rdme_two_particle_annihilation.cpp:61-65).

Random right-hand sides are drawn with numpy's PCG64 generator (``default_rng(seed).random``)
in state-index order, then normalised to unit L1 norm.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

from .model import Grid, LocalReaction, ReactionDiffusion, VoxelDomain, birth_death, box_size


@dataclass
class Workload:
    name: str
    model: ReactionDiffusion
    max_totals: List[int]
    shifts: List[float]
    rhs_kind: str               # "empty" (delta at the empty state) or "random"
    seeds: List[int] = field(default_factory=lambda: [0])

    @property
    def n_voxels(self) -> int:
        return self.model.n_voxels()

    @property
    def states(self) -> int:
        return box_size(self.n_voxels, self.max_totals)

    def rhs(self, seed: int = 0, n: int = None) -> np.ndarray:
        n = self.states if n is None else n
        if self.rhs_kind == "empty":
            b = np.zeros(n)
            b[0] = 1.0
            return b
        b = np.random.default_rng(seed).random(n)
        return b / b.sum()


def birth_death_1d(voxels: int, cap: int, shift: float = 0.5, rhs: str = "empty",
                   diffusivity: float = 1.0, birth: float = 0.05, death: float = 0.02,
                   name: str = None) -> Workload:
    """1D RDME on `voxels` voxels of width 1/voxels, one species, birth/death per voxel."""
    model = ReactionDiffusion(VoxelDomain(Grid([voxels], [1.0 / voxels]), [diffusivity]),
                              birth_death(birth, death))
    return Workload(name or f"rdme1d_v{voxels}_cap{cap}", model, [cap], [shift], rhs)


def annihilation_2d(cells: int = 8, caps: Sequence[int] = (2, 2), shift: float = 1.0,
                    birth: float = 0.1, k: float = 1.0, diffusivity: float = 1.0) -> Workload:
    """2D two-species RDME: births of A and B at birth*volume, A + B -> 0 at k nA nB / volume."""
    h = 1.0 / cells
    reactions = [
        LocalReaction([1, 0], lambda n, vol: birth * vol),
        LocalReaction([0, 1], lambda n, vol: birth * vol),
        LocalReaction([-1, -1], lambda n, vol: k * n[0] * n[1] / vol),
    ]
    model = ReactionDiffusion(VoxelDomain(Grid([cells, cells], [h, h]), [diffusivity, diffusivity]),
                              reactions)
    return Workload(f"rdme2d_{cells}x{cells}_caps{caps[0]}{caps[1]}", model, list(caps), [shift],
                    "empty")


def config_small() -> Workload:
    """1D, 16 voxels, cap 3 (969 states), s = 0.5, b = e_0."""
    return birth_death_1d(16, 3, 0.5, "empty", name="rdme1d_v16_cap3")


def config_large() -> Workload:
    """1D, 64 voxels, cap 5 (11,238,513 states), s = 0.5, b ~ U[0,1) seed 0, L1-normalised."""
    return birth_death_1d(64, 5, 0.5, "random", name="rdme1d_v64_cap5")


def config_2d() -> Workload:
    """2D 8x8 two-species annihilation, caps {2,2} (4,601,025 states), s = 1.0, b = e_0."""
    return annihilation_2d(8, (2, 2), 1.0)


def config_multishift() -> Workload:
    """1D, 32 voxels, cap 4 (58,905 states), 64 shifts 10^(-2 + 4i/63) x 4 random rhs."""
    w = birth_death_1d(32, 4, 0.5, "random", name="rdme1d_v32_cap4_multishift")
    w.shifts = [10.0 ** (-2.0 + 4.0 * i / 63.0) for i in range(64)]
    w.seeds = [0, 1, 2, 3]
    return w


def lumped_sector_sums(shift: float, beta: float, delta: float, cap: int,
                       b_sector: np.ndarray) -> np.ndarray:
    """Sector sums X_k of the solution of (sI - Q) x = b for a linear birth-death RDME in a
    box of cap `cap`: diffusion conserves the total N and the births (total rate beta) and
    deaths (delta per molecule) depend on N only, so X solves the tridiagonal system
    (s + beta + delta k) X_k - beta X_{k-1} - delta (k+1) X_{k+1} = B_k, births out of the
    top sector being absorbed (no X_{cap+1} term)."""
    K = cap + 1
    M = np.zeros((K, K))
    for k in range(K):
        M[k, k] = shift + beta + delta * k
        if k > 0:
            M[k, k - 1] = -beta
        if k + 1 < K:
            M[k, k + 1] = -delta * (k + 1)
    return np.linalg.solve(M, np.asarray(b_sector, np.float64))
