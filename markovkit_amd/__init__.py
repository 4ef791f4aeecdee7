"""markovkit_amd: MarkovKit's RDME resolvent solve (GMRES on shift I - Q over a
population_box, preconditioned by sector_multigrid) as hand-written HIP kernels for MI355X.

The model description (``model``) is plain host logic; the solver API (``api``) calls the
in-tree HIP library through its C ABI and has no CPU fallback.
"""
from .model import (Graph, Grid, LocalReaction, ReactionDiffusion, VoxelDomain, birth_death,
                    box_size, compositions)
from .api import (ABSORBING, AGGREGATE_BLOCK, FP32, FP64, ILU0, JACOBI, RED_BLACK, REFLECTING, BatchSolver, batch_graphs_enable,
                  DeviceVector, Server,
                  Generator, MultigridOptions, PopulationBox, PopulationSector, SectorMultigrid, ShiftedGenerator,
                  SolverResult, device_setup_enable, expv, expv_batch, expv_device, fused_first_enable, coarse_fuse_enable, tiny_gmres_enable, fused_tail_enable, tail_lds_enable, gmres, graphs_enable, profile_enable, profile_read,
                  profile_reset, synchronize)
from . import boxmg
from .boxmg import BoxMultigrid

__all__ = [
    "Graph", "Grid", "LocalReaction", "ReactionDiffusion", "VoxelDomain", "birth_death",
    "box_size", "compositions", "ABSORBING", "AGGREGATE_BLOCK", "FP32", "FP64", "ILU0", "JACOBI", "RED_BLACK",
    "REFLECTING", "BatchSolver", "batch_graphs_enable", "DeviceVector", "Server", "Generator", "MultigridOptions", "PopulationBox", "PopulationSector",
    "SectorMultigrid", "ShiftedGenerator", "SolverResult", "device_setup_enable", "expv", "expv_batch", "expv_device", "fused_first_enable", "coarse_fuse_enable", "tiny_gmres_enable", "fused_tail_enable", "tail_lds_enable", "gmres", "graphs_enable",
    "profile_enable",
    "profile_read", "profile_reset", "synchronize", "boxmg", "BoxMultigrid",
]
