"""ctypes access to tests/native/libsetup_probe.so: the HIP library's host setup code
(state space, gather-form generator, multigrid hierarchy) compiled for CPU-only tests."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "native", "setup_probe.cpp"),
       os.path.join(ROOT, "markovkit_amd", "csrc", "setup.cpp")]
LIB = os.path.join(HERE, "native", "libsetup_probe.so")

_lib = None


def build():
    tmp = LIB + ".tmp"
    # the library's host flags: no contraction of products into sums
    subprocess.run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-pthread",
                    "-march=x86-64-v3", "-ffp-contract=off", *SRC, "-o", tmp], check=True)
    os.replace(tmp, LIB)


def lib():
    global _lib
    if _lib is None:
        stale = (not os.path.exists(LIB) or
                 any(os.path.getmtime(s) > os.path.getmtime(LIB) for s in SRC
                     + [os.path.join(ROOT, "markovkit_amd", "csrc", "setup.hpp")]))
        if stale:
            build()
        L = C.CDLL(LIB)
        L.probe_error.restype = C.c_char_p
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(st):
    if st != 0:
        msg = lib().probe_error().decode()
        raise {1: ValueError, 2: OverflowError}.get(st, RuntimeError)(msg)


def _domain_args(model):
    from markovkit_amd.model import Grid
    dom = model.domain
    vols = np.ascontiguousarray(dom.volumes, np.float64)
    diff = np.ascontiguousarray(dom.diffusivities, np.float64)
    if isinstance(dom.graph, Grid):
        dims = np.asarray(dom.graph.dims, np.int64)
        sp = np.asarray(dom.graph.spacing, np.float64)
        keep = [dims, sp, vols, diff]
        return keep, [dom.graph.dim, _p(dims), _p(sp), C.c_int64(dom.n_voxels()), None, None,
                      None, _p(vols), dom.n_species(), _p(diff)]
    off, to, w = dom.graph.adjacency()
    keep = [off, to, w, vols, diff]
    return keep, [0, None, None, C.c_int64(dom.n_voxels()), _p(off), _p(to), _p(w), _p(vols),
                  dom.n_species(), _p(diff)]


def box_states(V, caps):
    caps = np.asarray(caps, np.int32)
    n, k = C.c_int64(), C.c_int64()
    _check(lib().probe_box(C.c_int64(V), len(caps), _p(caps), C.byref(n), C.byref(k)))
    states = np.zeros(n.value * V * len(caps), np.int32)
    begin = np.zeros(k.value + 1, np.int64)
    found = np.zeros(n.value, np.int64)
    _check(lib().probe_box_states(C.c_int64(V), len(caps), _p(caps), _p(states), _p(begin),
                                  _p(found)))
    return states.reshape(n.value, -1), begin, found


def sliced(model, caps, absorbing=True):
    """the coded sliced layout of the generator, or None when it does not fit"""
    caps = np.asarray(caps, np.int32)
    keep, args = _domain_args(model)
    changes = np.ascontiguousarray(model.changes(), np.int32)
    rates = np.ascontiguousarray(model.rate_tables(caps), np.float64)
    h = C.c_void_p()
    _check(lib().probe_generator(*args, _p(caps), len(model.reactions), _p(changes), _p(rates),
                                 1 if absorbing else 0, C.byref(h)))
    from markovkit_amd.model import box_size
    return _sliced_from(h, box_size(model.n_voxels(), list(caps)))


def sliced_csr(ptr, col, val, exit_):
    """the coded sliced layout of an arbitrary CSR (rows x any column range), or None"""
    n = len(ptr) - 1
    ptr = np.ascontiguousarray(ptr, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    val = np.ascontiguousarray(val, np.float64)
    exit_ = np.ascontiguousarray(exit_, np.float64)
    lib().probe_csr.restype = C.c_void_p
    h = C.c_void_p(lib().probe_csr(C.c_int64(n), _p(ptr), _p(col), _p(val), _p(exit_)))
    return _sliced_from(h, n)


def _sliced_from(h, rows):
    ns, padded, nv, ne = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    ok = lib().probe_sliced_sizes(h, C.byref(ns), C.byref(padded), C.byref(nv), C.byref(ne))
    if not ok:
        lib().probe_generator_free(h)
        return None
    soff = np.zeros(ns.value + 1, np.uint32)
    ent = np.zeros(padded.value, np.uint32)
    rlen = np.zeros(rows, np.uint8)
    rcode = np.zeros(rows, np.uint8)
    vdict = np.zeros(nv.value)
    edict = np.zeros(ne.value)
    lib().probe_sliced_arrays(h, _p(soff), _p(ent), _p(rlen), _p(rcode), _p(vdict), _p(edict))
    lib().probe_generator_free(h)
    return dict(soff=soff, ent=ent, rlen=rlen, rcode=rcode, vdict=vdict, edict=edict, n=rows)


def sliced_apply(S, shift, x, gather_sum=False):
    """(shift I - C) x evaluated with the device kernel's index arithmetic: slice s = row // 64,
    lane l = row % 64, entry k at soff[s] + 64 k + l, column = low 24 bits, code = high 8;
    gather_sum returns the plain row sums of rate * x[column] instead"""
    n = S["n"]
    y = np.zeros(n)
    for r in range(n):
        s, l = divmod(r, 64)
        base = int(S["soff"][s]) + l
        acc = 0.0
        for k in range(int(S["rlen"][r])):
            idx = base + 64 * k
            assert idx < int(S["soff"][s + 1])
            e = int(S["ent"][idx])
            c, code = e & 0xFFFFFF, e >> 24
            assert c < len(x) and code < len(S["vdict"])
            acc += S["vdict"][code] * x[c]
        y[r] = acc if gather_sum else (shift + S["edict"][S["rcode"][r]]) * x[r] - acc
    return y


def generator(model, caps, absorbing=True):
    """(ptr, col, val, exit, leak, n_jumps) of the gather-form generator."""
    caps = np.asarray(caps, np.int32)
    keep, args = _domain_args(model)
    changes = np.ascontiguousarray(model.changes(), np.int32)
    rates = np.ascontiguousarray(model.rate_tables(caps), np.float64)
    h = C.c_void_p()
    _check(lib().probe_generator(*args, _p(caps), len(model.reactions), _p(changes), _p(rates),
                                 1 if absorbing else 0, C.byref(h)))
    n, nnz, nj = C.c_int64(), C.c_int64(), C.c_int64()
    lib().probe_generator_sizes(h, C.byref(n), C.byref(nnz), C.byref(nj))
    ptr = np.zeros(n.value + 1, np.int64)
    col = np.zeros(nnz.value, np.int64)
    val = np.zeros(nnz.value)
    ex = np.zeros(n.value)
    lk = np.zeros(n.value)
    lib().probe_generator_arrays(h, _p(ptr), _p(col), _p(val), _p(ex), _p(lk))
    lib().probe_generator_free(h)
    return ptr, col, val, ex, lk, nj.value


class Hierarchy:
    def __init__(self, model, caps, shift, pre=2, post=2, omega=2.0 / 3.0, coarsest=4,
                 max_levels=32, cycles=1, growth=0):
        caps = np.asarray(caps, np.int32)
        keep, args = _domain_args(model)
        dim, dims, sp, _, _, _, _, vols, S, diff = args
        self.h = C.c_void_p()
        _check(lib().probe_hierarchy(dim, dims, sp, vols, S, diff, _p(caps), C.c_double(shift),
                                     pre, post, C.c_double(omega), C.c_int64(coarsest),
                                     max_levels, cycles, growth, C.byref(self.h)))
        self._keep = keep
        lib().probe_hierarchy_levels.restype = C.c_int64

    def levels(self, k):
        return lib().probe_hierarchy_levels(self.h, C.c_int64(k))

    def info(self, k, l):
        rows, term, nnz = C.c_int64(), C.c_int(), C.c_int64()
        lib().probe_level_info(self.h, C.c_int64(k), C.c_int64(l), C.byref(rows), C.byref(term),
                               C.byref(nnz))
        return rows.value, bool(term.value), nnz.value

    def transfer(self, k, l):
        rows, _, _ = self.info(k, l)
        co = np.zeros(rows, np.int64)
        w = np.zeros(rows)
        lib().probe_level_transfer(self.h, C.c_int64(k), C.c_int64(l), _p(co), _p(w))
        return co, w

    def operator(self, k, l):
        rows, _, nnz = self.info(k, l)
        ptr = np.zeros(rows + 1, np.int64)
        col = np.zeros(nnz, np.int64)
        val = np.zeros(nnz)
        ex = np.zeros(rows)
        lib().probe_level_operator(self.h, C.c_int64(k), C.c_int64(l), _p(ptr), _p(col), _p(val),
                                   _p(ex))
        return ptr, col, val, ex

    def inverse(self, k, l):
        rows, term, _ = self.info(k, l)
        assert term
        inv = np.zeros(rows * rows)
        lib().probe_level_inverse(self.h, C.c_int64(k), C.c_int64(l), _p(inv))
        return inv.reshape(rows, rows)

    def __del__(self):
        if getattr(self, "h", None):
            lib().probe_hierarchy_free(self.h)
            self.h = None
