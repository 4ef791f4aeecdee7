"""Collects the recorded runs of the reference's whole-box multigrid driver into fixtures:
tests/golden/box_mg_recorded.json (standalone solves), box_mg_steps.json (time-stepping runs:
per-solve rows and population marginals) and box_mg_adaptive.json (adaptive FSP attempts).

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

STEP_RUNS = ["time_check_s10", "transient_k80_s20"]
ADAPTIVE = "/root/reference/rdme/paper/data/adaptive_population_fsp"
ADAPTIVE_RUNS = ["linear_mid", "linear_tight", "linear_start8", "linear_loose", "cycle_2d",
                 "cycle_v", "schlogl_refined", "diffusion"]


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
        elif key in ("steps",):
            p[key] = int(v)
        elif key in ("integrator",):
            p[key] = v
        elif key == "fsp_tolerance":
            p[key] = float(v)
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

    steps = {}
    for name in STEP_RUNS:
        with open(os.path.join(DATA, name + ".json")) as f:
            cmd = json.load(f)["command"]
        with open(os.path.join(DATA, name + ".csv")) as f:
            rows = list(csv.DictReader(f))
        rec = {"params": params(cmd), "states": int(rows[0]["states"]),
               "solves": [{"step": int(r["step"]), "iterations": int(r["iterations"]),
                           "error": float(r["residual_l1_bound"]), "mass": float(r["mass"]),
                           "high_mass": float(r["high_mass"]),
                           "solve_s": float(r["solve_s"])} for r in rows],
               "wall_s": json.load(open(os.path.join(DATA, name + ".json")))["wall_seconds"]}
        mp = os.path.join(DATA, name + "_marginal.csv")
        if os.path.exists(mp):
            with open(mp) as f:
                m = list(csv.DictReader(f))
            times = sorted({float(r["time"]) for r in m})
            rec["marginal_times"] = times
            rec["marginals"] = [[float(r["probability"]) for r in m if float(r["time"]) == t]
                                for t in times]
        steps[name] = rec
    with open(os.path.join(HERE, "box_mg_steps.json"), "w") as f:
        json.dump(steps, f)

    adaptive = {}
    for name in ADAPTIVE_RUNS:
        with open(os.path.join(ADAPTIVE, name + ".json")) as f:
            cmd = json.load(f)["command"]
        with open(os.path.join(ADAPTIVE, name + ".csv")) as f:
            rows = list(csv.DictReader(f))
        p = params(cmd)
        p["cap_given"] = "--cap" in cmd
        adaptive[name] = {"params": p, "attempts": [
            {"attempt": int(r["attempt"]), "step": int(r["step"]), "cap": int(r["cap"]),
             "states": int(r["states"]), "accepted": int(r["accepted"]),
             "iterations": int(r["iterations"]), "error": float(r["residual_l1_bound"]),
             "missing_mass": float(r["missing_mass"])} for r in rows]}
    with open(os.path.join(HERE, "box_mg_adaptive.json"), "w") as f:
        json.dump(adaptive, f, indent=0)


if __name__ == "__main__":
    main()
