"""Writes the golden fixtures under tests/golden/ that are not copied reference data.

rdme1d_v16_cap3_direct.npy: the solution of (0.5 I - Q) x = e_0 for the 1D, 16-voxel, cap-3
RDME (969 states; births 0.05 * volume, deaths 0.02 * n, D = 1, h = 1/16, absorbing), with Q
assembled by the independent brute-force enumeration in tests/bruteforce.py (not by the
oracle) and solved by scipy's sparse LU.

The transfer CSVs beside it are data files copied from the reference
(rdme/paper/data/transfer_sparsity/schlogl_k40_level1.csv -> transfer_2to1_cap40.csv,
schlogl_k40_population4.csv -> transfer_4to2_population4.csv).

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from markovkit_amd.workloads import config_small  # noqa: E402
from tests import bruteforce  # noqa: E402


def main():
    w = config_small()
    # brute force enumerates (cap+1)^V tuples; the 16-voxel box is built from its sectors'
    # explicit compositions instead (same sort key as bruteforce.box_states)
    V, caps = w.n_voxels, w.max_totals
    import itertools
    states = []
    for t in range(caps[0] + 1):
        for cuts in itertools.combinations_with_replacement(range(V), t):
            x = [0] * V
            for c in cuts:
                x[c] += 1
            states.append(tuple(x))
    states.sort(key=lambda s: (sum(s), s))
    index = {s: i for i, s in enumerate(states)}
    n = len(states)
    assert n == 969
    dom = w.model.domain
    rows, cols, vals = [], [], []
    exit_ = np.zeros(n)
    for j, s in enumerate(states):
        x = list(s)
        moves = []
        for v in range(V):
            for to, wt in dom.graph.neighbors(v):
                if x[v]:
                    y = list(x)
                    y[v] -= 1
                    y[to] += 1
                    moves.append((tuple(y), x[v] * dom.jump_rate(0, v, wt)))
        for v in range(V):
            for r in w.model.reactions:
                rate = r.rate([x[v]], float(dom.volumes[v]))
                if rate > 0:
                    y = list(x)
                    y[v] += r.change[0]
                    if y[v] >= 0:
                        moves.append((tuple(y), rate))
        for y, rate in moves:
            exit_[j] += rate
            i = index.get(y)
            if i is not None:
                rows.append(i)
                cols.append(j)
                vals.append(-rate)
    rows += list(range(n))
    cols += list(range(n))
    vals += list(0.5 + exit_)
    A = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    b = np.zeros(n)
    b[0] = 1.0
    x = spl.spsolve(A, b)
    np.save(os.path.join(HERE, "rdme1d_v16_cap3_direct.npy"), x)
    print("wrote rdme1d_v16_cap3_direct.npy", n, x.sum())


if __name__ == "__main__":
    main()
