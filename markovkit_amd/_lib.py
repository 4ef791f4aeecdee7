"""Loads ``libmarkovkit_amd.so`` (built in-tree) and declares its C ABI for ctypes.

There is no fallback: if the library is missing or cannot load, importing the solver API
raises. Status codes are mapped back to the reference's exception types: invalid argument ->
ValueError (std::invalid_argument), overflow -> OverflowError (std::overflow_error), runtime
and device failures -> RuntimeError.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmarkovkit_amd.so")

# exported entry points, with their ctypes signatures (matching include/markovkit_amd.h)
P = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double
pi64 = C.POINTER(C.c_int64)
pf64 = C.POINTER(C.c_double)


class MultigridOptions(C.Structure):
    _fields_ = [("pre_sweeps", i32), ("post_sweeps", i32), ("jacobi_weight", f64),
                ("coarsest_size", i64), ("max_levels", i32), ("coarse_cycles", i32),
                ("sweep_growth", i32), ("smoother", i32), ("precision", i32),
                ("cycle_levels", i32)]


class GmresOptions(C.Structure):
    _fields_ = [("tolerance", f64), ("max_iterations", i64), ("restart", i64),
                ("orthogonalization", i32)]


class SolverResult(C.Structure):
    _fields_ = [("iterations", i64), ("residual", f64), ("converged", i32)]


class ExpvOptions(C.Structure):
    _fields_ = [("krylov_dim", i64), ("tolerance", f64)]


class ExpvResult(C.Structure):
    _fields_ = [("steps", i64), ("rejected", i64), ("error", f64)]


class BoxMgParams(C.Structure):
    _fields_ = [("dimension", i32), ("cells", i32), ("cap", i32), ("diffusion", f64),
                ("shift", f64), ("c1", f64), ("c2", f64), ("c3", f64), ("c4", f64),
                ("smoother", i32), ("sweeps", i32), ("cycles", i32), ("cycle_type", i32),
                ("terminal_rows", i64)]


class BoxMgResult(C.Structure):
    _fields_ = [("iterations", i32), ("converged", i32), ("error", f64), ("history", i64)]


class ServerResult(C.Structure):
    _fields_ = [("kind", i32), ("status", i32), ("solve", SolverResult), ("expv", ExpvResult),
                ("queued_s", f64), ("run_s", f64)]


SIGNATURES = {
    "mk_last_error": ([], C.c_char_p),
    "mk_default_multigrid_options": ([C.POINTER(MultigridOptions)], None),
    "mk_box_create": ([i64, i32, P, C.POINTER(P)], i32),
    "mk_box_destroy": ([P], i32),
    "mk_box_size": ([P, pi64], i32),
    "mk_box_n_sectors": ([P, pi64], i32),
    "mk_box_sector": ([P, i64, pi64, pi64, P], i32),
    "mk_box_state": ([P, i64, P], i32),
    "mk_box_find": ([P, P, pi64], i32),
    "mk_sector_space_create": ([i64, i32, P, C.POINTER(P)], i32),
    "mk_domain_create_grid": ([i32, P, P, P, i32, P, C.POINTER(P)], i32),
    "mk_domain_create_graph": ([i64, P, P, P, P, i32, P, C.POINTER(P)], i32),
    "mk_domain_destroy": ([P], i32),
    "mk_generator_create_rdme": ([P, P, i32, P, P, i32, C.POINTER(P)], i32),
    "mk_generator_create_jumps": ([i64, P, P, P, P, P, C.POINTER(P)], i32),
    "mk_generator_destroy": ([P], i32),
    "mk_generator_size": ([P, pi64, pi64], i32),
    "mk_generator_rates": ([P, P, P], i32),
    "mk_generator_entries": ([P, pi64, P, P, P], i32),
    "mk_generator_apply": ([P, P, P], i32),
    "mk_shifted_apply": ([P, f64, P, P], i32),
    "mk_shifted_apply_device": ([P, f64, P, P], i32),
    "mk_sector_mg_create": ([P, P, f64, C.POINTER(MultigridOptions), C.POINTER(P)], i32),
    "mk_sector_mg_destroy": ([P], i32),
    "mk_sector_mg_info": ([P, pi64, pi64, pi64], i32),
    "mk_sector_mg_apply": ([P, P, P], i32),
    "mk_sector_mg_level": ([P, i32, pi64, pi64, C.POINTER(i32)], i32),
    "mk_sector_mg_transfer": ([P, i32, P, P], i32),
    "mk_sector_mg_relax": ([P, P, P, i32, i32], i32),
    "mk_sector_mg_relax_device": ([P, P, P, i32, i32], i32),
    "mk_sector_mg_apply_device": ([P, P, P], i32),
    "mk_gmres": ([P, f64, P, P, P, C.POINTER(GmresOptions), C.POINTER(SolverResult)], i32),
    "mk_gmres_device": ([P, f64, P, P, P, C.POINTER(GmresOptions), C.POINTER(SolverResult)], i32),
    "mk_gmres_device_noalloc": ([P, f64, P, P, P, C.POINTER(GmresOptions), C.POINTER(SolverResult)], i32),
    "mk_sector_mg_set_shift": ([P, f64], i32),
    "mk_expv": ([P, f64, P, P, C.POINTER(ExpvOptions), C.POINTER(ExpvResult)], i32),
    "mk_expv_device": ([P, f64, P, P, C.POINTER(ExpvOptions), C.POINTER(ExpvResult)], i32),
    "mk_expv_batch": ([P, i64, P, P, P, C.POINTER(ExpvOptions), P], i32),
    "mk_server_create": ([i32, i64, C.POINTER(P)], i32),
    "mk_server_destroy": ([P], i32),
    "mk_server_add_family": ([P, P, i32, P, C.POINTER(i32)], i32),
    "mk_server_submit_resolvent": ([P, i32, f64, P, f64, C.POINTER(GmresOptions), C.POINTER(i64)], i32),
    "mk_server_submit_resolvent_sized": ([P, i32, f64, P, i64, f64, C.POINTER(GmresOptions), pi64], i32),
    "mk_server_batching": ([P, i64, i32], i32),
    "mk_server_submit_expv": ([P, P, f64, P, C.POINTER(ExpvOptions), C.POINTER(i64)], i32),
    "mk_server_wait": ([P, i64, P, C.POINTER(ServerResult)], i32),
    "mk_batch_create": ([P, P, i64, P, C.POINTER(P)], i32),
    "mk_batch_create_sized": ([P, P, i64, P, P, C.POINTER(P)], i32),
    "mk_batch_graphs_enable": ([i32], i32),
    "mk_batch_set_block": ([P, i32], i32),
    "mk_batch_destroy": ([P], i32),
    "mk_batch_solve": ([P, P, P, P, C.POINTER(GmresOptions), C.POINTER(SolverResult)], i32),
    "mk_batch_solve_device": ([P, P, P, P, C.POINTER(GmresOptions), C.POINTER(SolverResult)], i32),
    "mk_box_mg_create": ([C.POINTER(BoxMgParams), C.POINTER(P)], i32),
    "mk_box_mg_destroy": ([P], i32),
    "mk_box_mg_info": ([P, pi64, C.POINTER(i32)], i32),
    "mk_box_mg_level": ([P, i32, pi64, pi64, pi64, C.POINTER(i32)], i32),
    "mk_box_mg_generator": ([P, C.POINTER(P)], i32),
    "mk_box_mg_cycle_device": ([P, P, P], i32),
    "mk_box_mg_apply_device": ([P, P, P], i32),
    "mk_box_mg_solve": ([P, P, P, f64, i32, P, i64, C.POINTER(BoxMgResult)], i32),
    "mk_box_mg_solve_device": ([P, P, P, f64, i32, P, i64, C.POINTER(BoxMgResult)], i32),
    "mk_box_mg_step_device": ([P, P, i32, f64, i32, C.POINTER(BoxMgResult)], i32),
    "mk_gmres_box": ([P, f64, P, P, P, C.POINTER(GmresOptions), C.POINTER(SolverResult)], i32),
    "mk_gmres_box_device": ([P, f64, P, P, P, C.POINTER(GmresOptions), C.POINTER(SolverResult)], i32),
    "mk_box_mg_graphs_enable": ([i32], i32),
    "mk_box_mg_small_enable": ([i32], i32),
    "mk_device_alloc": ([i64, C.POINTER(P)], i32),
    "mk_device_free": ([P], i32),
    "mk_copy_to_device": ([P, P, i64], i32),
    "mk_copy_to_host": ([P, P, i64], i32),
    "mk_device_zero": ([P, i64], i32),
    "mk_device_copy": ([P, P, i64], i32),
    "mk_synchronize": ([], i32),
    "mk_graphs_enable": ([i32], i32),
    "mk_fused_tail_enable": ([i32], i32),
    "mk_device_setup_enable": ([i32], i32),
    "mk_tail_lds_enable": ([i32], i32),
    "mk_fused_first_enable": ([i32], i32),
    "mk_coarse_fuse_enable": ([i32], i32),
    "mk_tiny_gmres_enable": ([i32], i32),
    "mk_set_xcd_chunk": ([i32], i32),
    "mk_get_xcd_chunk": ([C.POINTER(i32)], i32),
    "mk_profile_enable": ([i32], i32),
    "mk_profile_reset": ([], i32),
    "mk_profile_read": ([C.c_char_p, pi64, pf64, pf64], i32),
}

_lib = None


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m markovkit_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(status: int):
    if status == 0:
        return
    msg = lib().mk_last_error().decode()
    if status == 1:
        raise ValueError(msg)
    if status == 2:
        raise OverflowError(msg)
    raise RuntimeError(msg)
