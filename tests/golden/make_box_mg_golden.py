"""Collects the recorded runs of the reference's whole-box multigrid driver into one fixture,
tests/golden/box_mg_recorded.json.

Source: the reference's data files rdme/paper/data/full_multigrid/<name>.json (the command
line of each run), <name>.csv (its summary row) and, where present, <name>_history.csv (the
l1 error bound before every cycle). Only numbers are kept: the run's parameters, iteration
count, final error bound, mass, minimum probability, and the history. Nothing of the
reference's code is read.

Run from the repository root with the reference checked out at /root/reference:
python tests/golden/make_box_mg_golden.py
"""
import csv
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = "/root/reference/rdme/paper/data/full_multigrid"

RUNS = [
    "graph_gap_cap2_w", "voxels_1d_64_mg", "voxels_1d_64_kcycle", "voxels_2d_4_kcycle",
    "voxels_2d_8_mg", "voxels_3d_2_kcycle", "voxels_3d_4_mg", "voxels_1d_128_mg",
    "block_v16_k4_gs_s1", "block_v16_k4_gs_s2", "block_v16_k4_block-ilu_s1",
    "block_v16_k4_block-ilu_s2", "block_schlogl_v4_k40_gs_s1",
    "block_schlogl_v4_k40_block-ilu_s1_kcycle", "block_schlogl_v4_k40_gs_s1_kcycle",
    "block_v16_k8_gs_s2",
    # headline timings of the paper (recorded on an Apple M1 Pro, one thread)
    "voxels_1d_1024_mg", "voxels_1d_1024_kcycle", "voxels_3d_8_mg", "block_v64_k4_gs_s1",
    "block_v64_k4_block-ilu_s1",
]

FLAGS = {"--check", "--export-only", "--profile", "--adaptive-fsp"}


def params(command):
    p = {"dimension": 1, "cells": 4, "cap": 40, "sweeps": 2, "cycles": 1, "initial": 8,
         "diffusion": 1.0, "shift": 0.05, "tolerance": 1e-9, "max_iterations": 1000,
         "rates": [0.0133, 0.0006, 0.6924125, 0.114725], "smoother": "gs", "kcycle": False}
    a = command[1:]
    i = 0
    while i < len(a):
        k = a[i]
        if k in FLAGS:
            i += 1
            continue
        v = a[i + 1]
        i += 2
        key = k[2:].replace("-", "_")
        if key in ("dimension", "cells", "cap", "sweeps", "cycles", "initial", "max_iterations"):
            p[key] = int(v)
        elif key in ("diffusion", "shift", "tolerance"):
            p[key] = float(v)
        elif key == "linear":
            s = float(v)
            p["rates"] = [0.0, 0.0, 2 * s, s]
        elif key == "smoother":
            p["smoother"] = v
        elif key == "cycle_type":
            p["kcycle"] = v == "k"
        elif key == "method":
            assert v == "mg", v
    return p


def main():
    out = {}
    for name in RUNS:
        with open(os.path.join(DATA, name + ".json")) as f:
            cmd = json.load(f)["command"]
        with open(os.path.join(DATA, name + ".csv")) as f:
            row = list(csv.DictReader(f))[-1]
        rec = {"params": params(cmd),
               "states": int(row["states"]),
               "iterations": int(row["iterations"]),
               "error": float(row["residual_l1_bound"]),
               "mass": float(row["mass"]),
               "min_probability": float(row["min_probability"]),
               "setup_s": float(row["setup_s"]),
               "solve_s": float(row["solve_s"])}
        hp = os.path.join(DATA, name + "_history.csv")
        if os.path.exists(hp):
            with open(hp) as f:
                rec["history"] = [float(r["residual_l1_bound"]) for r in csv.DictReader(f)]
        out[name] = rec
    with open(os.path.join(HERE, "box_mg_recorded.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
