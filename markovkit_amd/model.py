"""Host-side description of an RDME problem, mirroring MarkovKit's model types.

Reference types mirrored (paths relative to the markovkit repository root):

* ``Grid``             num::structured_grid_graph<D> as used by ``grid<D>``
                       (include/markovkit/rdme/multigrid.hpp:20-21): axis 0 fastest,
                       non-periodic, neighbour weight = face area / spacing, cell volume =
                       product of spacings (pinned by tests/test_rdme.cpp:22-33).
* ``Graph``            an explicit voxel graph (num::graph, tests/test_rdme.cpp:35-43).
* ``VoxelDomain``      voxel_domain (include/markovkit/rdme/reaction_diffusion.hpp:23-56).
* ``LocalReaction``    local_reaction (reaction_diffusion.hpp:63-67): ``change`` per species and
                       ``rate(counts, volume)``.
* ``ReactionDiffusion`` reaction_diffusion (reaction_diffusion.hpp:71-119).

The HIP library receives local reaction rates as tables evaluated on the host:
``rates[r, v, t]`` is ``reactions[r].rate(counts, volumes[v])`` for the per-voxel count tuple
``t`` (mixed radix ``max_totals[s] + 1``, species 0 most significant). A state of a
``population_box`` never holds more than ``max_totals[s]`` molecules of species ``s`` in a voxel,
so the table holds every value the reference's ``std::function`` rates are asked for.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np


class Grid:
    """Structured voxel grid: ``dims`` cells per axis, ``spacing`` per axis."""

    def __init__(self, dims: Sequence[int], spacing: Optional[Sequence[float]] = None):
        self.dims = [int(d) for d in dims]
        if spacing is None:
            spacing = [1.0] * len(self.dims)
        self.spacing = [float(h) for h in spacing]
        if len(self.spacing) != len(self.dims) or not 1 <= len(self.dims) <= 3:
            raise ValueError("Grid: 1 to 3 axes, one spacing per axis")
        if any(d < 1 for d in self.dims):
            raise ValueError("Grid: every axis needs a cell")

    @property
    def dim(self) -> int:
        return len(self.dims)

    def n_vertices(self) -> int:
        n = 1
        for d in self.dims:
            n *= d
        return n

    def cell_volume(self) -> float:
        v = 1.0
        for h in self.spacing:
            v *= h
        return v

    def face_weight(self, axis: int) -> float:
        area = 1.0
        for b, h in enumerate(self.spacing):
            if b != axis:
                area *= h
        return area / self.spacing[axis]

    def coordinate(self, v: int) -> List[int]:
        c = []
        for d in self.dims:
            c.append(v % d)
            v //= d
        return c

    def neighbors(self, v: int):
        """[(to, weight)], axis by axis, lower neighbour first."""
        c = self.coordinate(v)
        out = []
        stride = 1
        for a, d in enumerate(self.dims):
            w = self.face_weight(a)
            if c[a] > 0:
                out.append((v - stride, w))
            if c[a] + 1 < d:
                out.append((v + stride, w))
            stride *= d
        return out

    def adjacency(self):
        offsets = [0]
        to, weight = [], []
        for v in range(self.n_vertices()):
            for t, w in self.neighbors(v):
                to.append(t)
                weight.append(w)
            offsets.append(len(to))
        return (np.asarray(offsets, np.int64), np.asarray(to, np.int64),
                np.asarray(weight, np.float64))


class Graph:
    """Explicit undirected voxel graph: ``add_edge(u, v, w)`` adds both directions."""

    def __init__(self, n_vertices: int):
        self.adj = [[] for _ in range(n_vertices)]

    def add_edge(self, u: int, v: int, w: float):
        self.adj[u].append((v, float(w)))
        self.adj[v].append((u, float(w)))

    def n_vertices(self) -> int:
        return len(self.adj)

    def neighbors(self, v: int):
        return list(self.adj[v])

    def adjacency(self):
        offsets = [0]
        to, weight = [], []
        for v in range(self.n_vertices()):
            for t, w in self.adj[v]:
                to.append(t)
                weight.append(w)
            offsets.append(len(to))
        return (np.asarray(offsets, np.int64), np.asarray(to, np.int64),
                np.asarray(weight, np.float64))


class VoxelDomain:
    """voxel_domain: graph, per-voxel volumes, per-species diffusivities."""

    def __init__(self, graph, *args):
        self.graph = graph
        if len(args) == 1:
            if not isinstance(graph, Grid):
                raise ValueError("voxel_domain: uniform volumes need a grid")
            self.volumes = np.full(graph.n_vertices(), graph.cell_volume())
            self.diffusivities = np.asarray(args[0], np.float64)
        elif len(args) == 2:
            self.volumes = np.asarray(args[0], np.float64)
            self.diffusivities = np.asarray(args[1], np.float64)
        else:
            raise TypeError("VoxelDomain(graph, [volumes,] diffusivities)")
        if self.volumes.shape != (graph.n_vertices(),):
            raise ValueError("voxel_domain: one volume per voxel")
        if not np.all(np.isfinite(self.volumes)) or not np.all(self.volumes > 0):
            raise ValueError("voxel_domain: volumes must be positive")

    def n_voxels(self) -> int:
        return self.graph.n_vertices()

    def n_species(self) -> int:
        return len(self.diffusivities)

    def jump_rate(self, species: int, frm: int, weight: float) -> float:
        return self.diffusivities[species] * weight / self.volumes[frm]


@dataclass
class LocalReaction:
    change: Sequence[int]
    rate: Callable[[Sequence[int], float], float]


@dataclass
class ReactionDiffusion:
    domain: VoxelDomain
    reactions: List[LocalReaction] = field(default_factory=list)

    def n_voxels(self) -> int:
        return self.domain.n_voxels()

    def n_species(self) -> int:
        return self.domain.n_species()

    def index(self, voxel: int, species: int) -> int:
        return voxel * self.n_species() + species

    def rate_tables(self, max_totals: Sequence[int]) -> np.ndarray:
        """rates[r, v, t] for every count tuple t with n_s <= max_totals[s]."""
        S = self.n_species()
        V = self.n_voxels()
        radix = [int(c) + 1 for c in max_totals]
        T = int(np.prod(radix))
        out = np.zeros((len(self.reactions), V, T), np.float64)
        tuples = list(itertools.product(*[range(r) for r in radix]))  # species 0 most significant
        cache = {}
        for r, reaction in enumerate(self.reactions):
            for v in range(V):
                vol = float(self.domain.volumes[v])
                key = (r, vol)
                if key not in cache:
                    cache[key] = np.array([float(reaction.rate(list(t), vol)) for t in tuples])
                out[r, v] = cache[key]
        return out

    def changes(self) -> np.ndarray:
        S = self.n_species()
        ch = np.zeros((len(self.reactions), S), np.int32)
        for r, reaction in enumerate(self.reactions):
            if len(reaction.change) != S:
                raise ValueError("local_reaction: one change per species")
            ch[r] = reaction.change
        return ch


def birth_death(birth: float, death: float) -> List[LocalReaction]:
    """A birth at rate birth*volume and a death at rate death*count in every voxel
    (tests/test_rdme.cpp:16-20, examples/rdme_multigrid_benchmark.cpp:65-70)."""
    return [LocalReaction([1], lambda n, vol: birth * vol),
            LocalReaction([-1], lambda n, vol: death * n[0])]


def compositions(voxels: int, total: int) -> int:
    from math import comb
    return comb(total + voxels - 1, voxels - 1)


def box_size(voxels: int, max_totals: Sequence[int]) -> int:
    n = 0
    for totals in itertools.product(*[range(c + 1) for c in max_totals]):
        m = 1
        for t in totals:
            m *= compositions(voxels, t)
        n += m
    return n
