"""Brute-force generator for tiny RDME boxes, written independently of the oracle.

States are enumerated with itertools and ordered by an explicit sort key (sector totals in
mixed radix with species 0 most significant, then the counts of species 0 voxel by voxel, then
species 1, ...: the order population_box documents at rdme/population_space.hpp:34-36 and
:178-189); jumps come from an explicit list of moves and a dictionary lookup, not from
arithmetic rank. Used to pin the oracle's ordering, rates, exit and leak rates.
"""
from __future__ import annotations

import itertools

import numpy as np


def box_states(V, caps):
    S = len(caps)
    states = []
    for flat in itertools.product(*[range(max(caps) + 1)] * (V * S)):
        x = np.array(flat).reshape(V, S)
        totals = x.sum(axis=0)
        if np.all(totals <= np.array(caps)):
            states.append(tuple(flat))

    def key(flat):
        x = np.array(flat).reshape(V, S)
        totals = tuple(int(t) for t in x.sum(axis=0))
        per_species = tuple(int(x[v, s]) for s in range(S) for v in range(V))
        return totals, per_species

    states.sort(key=key)
    return states


def dense_generator(model, caps, absorbing=True):
    """(Q dense, exit, leak, states): Q[i, j] = rate j -> i, Q[j, j] = -exit_j."""
    V, S = model.n_voxels(), model.n_species()
    states = box_states(V, caps)
    index = {s: i for i, s in enumerate(states)}
    n = len(states)
    Q = np.zeros((n, n))
    exit_ = np.zeros(n)
    leak = np.zeros(n)
    dom = model.domain
    for j, flat in enumerate(states):
        x = list(flat)
        moves = []
        for v in range(V):
            for to, w in dom.graph.neighbors(v):
                for s in range(S):
                    cnt = x[v * S + s]
                    if cnt == 0:
                        continue
                    y = list(x)
                    y[v * S + s] -= 1
                    y[to * S + s] += 1
                    moves.append((tuple(y), cnt * dom.jump_rate(s, v, w)))
        for v in range(V):
            counts = x[v * S:(v + 1) * S]
            for r in model.reactions:
                rate = r.rate(list(counts), float(dom.volumes[v]))
                if not rate > 0:
                    continue
                y = list(x)
                ok = True
                for s in range(S):
                    y[v * S + s] += r.change[s]
                    ok = ok and y[v * S + s] >= 0
                if ok:
                    moves.append((tuple(y), rate))
        for y, rate in moves:
            i = index.get(y)
            if i is None:
                leak[j] += rate
            else:
                Q[i, j] += rate
                exit_[j] += rate
        if absorbing:
            exit_[j] += leak[j]
        Q[j, j] -= exit_[j]
    return Q, exit_, leak, states
