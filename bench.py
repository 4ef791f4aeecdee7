#!/usr/bin/env python3
"""RDME resolvent-solve benchmark on one MI355X.

Workload (default): 1D RDME on 64 voxels (h = 1/64, D = 1, births 0.05 * volume, deaths
0.02 * n), box cap 5 = 11,238,513 states, one shift s = 0.5, b ~ U[0,1) (numpy PCG64, seed 0)
normalised to unit L1 norm. One step = one resolvent solve (sI - Q) x = b from x = 0 to a
true FP64 relative residual <= 1e-10 by flexible GMRES(50) with delayed CGS2, preconditioned by
sector_multigrid (red-black Gauss-Seidel, 1 + 2 sweeps, W-cycles on the top 2 levels, sectors of
at most 800 states solved densely). b, x, the operator and the hierarchy are resident in HBM
before the timed region starts.

value = solves/s of the mixed path (FP32 multigrid inside FP64 FGMRES); the plain FP64 path
(FP64 multigrid) is timed beside it. The roofline of the dominant kernel class comes from HIP
events recorded around that class's launches inside the timed mixed region, its HBM traffic from
two rocprofv3 counter passes started by this script; cpu_baseline is one complete solve of the
same system by the FP64 CPU oracle on this host (the reference's method, 1 core), whose solution
also gives the per-state L1 distance. The default run adds the other configurations as fields:
the reference's method on the GPU (jacobi), the multi-shift batch (multishift), the paper
driver's transient (box_mg), the 2D two-species box (c3), the serving mix (serving), the batched
exponentials (expv) and the 969-state case (c1).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
EPS = float(np.finfo(np.float64).eps)
METRIC = "RDME resolvent solves/s + p99 latency at ≤1e-10 residual, mixed vs FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q))


def run_solves(mk, A, M, b_dev, x_dev, tol, steps, restart, profile_classes=None):
    lat = []
    iters = []
    if profile_classes:
        mk.profile_reset()
        mk.profile_enable(profile_classes)
    mk.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        x_dev.zero()
        res = mk.gmres(A, M, b_dev, x_dev, tolerance=tol, max_iterations=1000, restart=restart)
        lat.append(time.perf_counter() - t)
        iters.append(res.iterations)
        if not res.converged:
            raise RuntimeError(f"solve did not converge: {res}")
    mk.synchronize()
    total = time.perf_counter() - t0
    prof = None
    if profile_classes:
        prof = mk.profile_read()
        mk.profile_enable(False)
    return total, lat, iters, res, prof


SMOOTHERS = {"jacobi": "JACOBI", "red_black": "RED_BLACK", "ilu0": "ILU0",
             "aggregate_block": "AGGREGATE_BLOCK"}


def host_info():
    """CPU model and core counts of this host (the CPU baseline's hardware)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cores": usable}


def cpu_baseline(w, b, x_gpu, Q, tol, restart):
    """One complete FP64 CPU oracle solve of the same system with the reference's method
    (scatter apply_shifted, per-sector weighted-Jacobi V-cycle, restarted GMRES), on one core,
    timed; the oracle's setup (state space, generator, hierarchy) is timed separately and
    excluded. Its solution also gives the per-state parity of the GPU solution:
    ||x_gpu - x_oracle||_1 <= (||r_gpu||_1 + ||r_oracle||_1 + rounding) / s, where both
    residuals are evaluated in FP64 and rounding = 16 eps sum_j (s + 2 exit_j - leak_j)
    (|x_gpu,j| + |x_oracle,j|) bounds their evaluation error."""
    from oracle.oracle import Problem, SectorMultigrid
    shift = w.shifts[0]
    t = time.perf_counter()
    P = Problem(w.model, w.max_totals)
    Mo = SectorMultigrid(P, shift)
    setup = time.perf_counter() - t
    log(f"[cpu] oracle setup {setup:.1f}s; one complete FP64 solve on one core")
    t = time.perf_counter()
    x_ref, info = P.gmres(shift, b, mg=Mo, tol=tol, max_iterations=1000, restart=restart)
    solve_s = time.perf_counter() - t
    log(f"[cpu] oracle solve {solve_s:.1f}s, {info['iterations']} iterations, "
        f"converged {info['converged']}")
    r_gpu = b - P.apply_shifted(shift, x_gpu)
    r_ref = b - P.apply_shifted(shift, x_ref)
    absA = shift + 2 * Q.exit_rates() - Q.leak_rates()
    rounding = 16 * EPS * float((absA * (np.abs(x_gpu) + np.abs(x_ref))).sum())
    l1 = float(np.abs(x_gpu - x_ref).sum())
    bound = float((np.abs(r_gpu).sum() + np.abs(r_ref).sum() + rounding) / shift)
    hi = host_info()
    out = {
        "value": 1.0 / solve_s,
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"one complete FP64 CPU oracle solve (scatter apply_shifted, per-sector "
                   f"weighted-Jacobi V-cycle 2+2, GMRES({restart}) with MGS) of the same "
                   f"{P.n}-state system to the same tolerance: {info['iterations']} iterations, "
                   f"{solve_s:.1f}s on one core of {hi['cpu_model']}; setup {setup:.1f}s excluded"),
        "seconds_per_solve": solve_s,
        "setup_s": setup,
        "iterations": info["iterations"],
        "converged": bool(info["converged"]),
        "relative_residual": float(np.linalg.norm(r_ref) / np.linalg.norm(b)),
        **hi,
    }
    parity = {"l1_vs_oracle": l1, "l1_vs_oracle_bound": bound, "within_bound": l1 <= bound,
              "oracle_relative_residual": out["relative_residual"]}
    return out, parity, P


def box_mg(mk, cpu=True, solves=None):
    """The paper driver's whole-box multigrid on its recorded transient (transient_k80_s20:
    Schlogl rates, 4 voxels, population cap 80, 1,929,501 states, shift 0.4, W-cycles with two
    sweeps, a Rannacher start and Crank-Nicolson steps: 21 solves, each to an l1 error bound of
    1e-9). Device-resident: the distribution stays in HBM across the steps. The solver is the
    library's (four-colour Gauss-Seidel; the driver used lexicographic Gauss-Seidel). Checked
    against the recorded final population marginal; the recorded driver wall time (Apple M1 Pro,
    one thread) is context. CPU baseline: the oracle's restatement of the same algorithm on a
    bounded sample (cycles of the first solve), scaled to the device's cycle count."""
    from markovkit_amd import boxmg
    with open(os.path.join(ROOT, "tests", "golden", "box_mg_steps.json")) as f:
        rec = json.load(f)["transient_k80_s20"]
    p = rec["params"]
    solves = p["steps"] + 1 if solves is None else min(solves, p["steps"] + 1)
    t = time.perf_counter()
    M = mk.BoxMultigrid(p["dimension"], p["cells"], p["cap"], p["diffusion"], p["shift"],
                        tuple(p["rates"]), "colour-gs", p["sweeps"], p["cycles"], p["kcycle"])
    mk.synchronize()
    setup = time.perf_counter() - t
    _, p0 = M.initial_problem(p["initial"])
    dp = mk.DeviceVector(M.rows(), p0)

    def transient():
        cycles = 0
        for step in range(1, solves + 1):
            it, err, ok = M.step_device(dp, step > 2, p["tolerance"], p["max_iterations"])
            if not ok:
                raise RuntimeError(f"box multigrid step {step} did not converge ({err})")
            cycles += it
        return cycles

    transient()  # warm-up: captures the cycle graph
    dp.upload(p0)
    mk.synchronize()
    t = time.perf_counter()
    cycles = transient()
    mk.synchronize()
    wall = time.perf_counter() - t
    final = boxmg.marginal(dp.download(), M.voxels, M.cap)
    dev = float(np.max(np.abs(final - np.asarray(rec["marginals"][solves - 1]))))
    # dominant kernel: the finest level's colour passes, timed over one solve
    dp.upload(p0)
    mk.profile_enable(["smooth_sweep"])
    mk.profile_reset()
    M.step_device(dp, False, p["tolerance"], p["max_iterations"])
    mk.synchronize()
    sw = mk.profile_read()["smooth_sweep"]
    mk.profile_enable(False)
    achieved = sw["bytes"] / sw["ms"] / 1e6
    out = {
        "workload": (f"paper transient {M.rows()} states ({p['cells']} voxels, cap {p['cap']}, "
                     f"Schlogl), W-cycle 2+2 four-colour GS, {solves} solves (BE start + CN) to "
                     f"l1 bound {p['tolerance']}"),
        "solves": solves, "cycles": cycles, "seconds": wall, "solves_per_s": solves / wall,
        "setup_s": setup, "levels": M.levels(),
        "final_marginal_max_abs_dev_vs_recorded": dev,
        "recorded_driver": {"wall_s": rec["wall_s"],
                            "cycles": sum(r["iterations"] for r in rec["solves"]),
                            "hardware": "Apple M1 Pro, one thread (lexicographic Gauss-Seidel)"},
        "speedup_vs_recorded_wall": rec["wall_s"] / wall,
        "roofline": {"kernel": "smooth_sweep (finest level, one colour per launch)",
                     "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "launches": sw["launches"],
                     "avg_launch_ms": sw["ms"] / sw["launches"],
                     "algorithmic_bytes_per_launch": sw["bytes"] / sw["launches"]},
        "cpu_baseline": None,
    }
    log(f"[box] {solves} solves, {cycles} cycles in {wall:.2f}s (recorded driver {rec['wall_s']:.1f}s); "
        f"sweep {achieved:.0f} GB/s; final marginal dev {dev:.1e}")
    if cpu:
        from oracle.oracle import BoxMultigrid as Oracle
        t = time.perf_counter()
        O = Oracle(p["dimension"], p["cells"], p["cap"], p["diffusion"], p["shift"],
                   tuple(p["rates"]), "colour-gs", p["sweeps"], p["cycles"], p["kcycle"])
        osetup = time.perf_counter() - t
        b = p["shift"] * p0
        k = 3
        t = time.perf_counter()
        O.solve(b, p0, 0.0, k)
        per_cycle = (time.perf_counter() - t) / k
        est = per_cycle * cycles
        out["cpu_baseline"] = {
            "value": solves / est, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"FP64 CPU oracle (box_mg.hpp, the same four-colour W-cycle) on the same "
                       f"{O.n}-state operator: {k} cycles of the first solve ({per_cycle:.2f}s per "
                       f"cycle, error bound included) scaled to the device's {cycles} cycles; "
                       f"setup {osetup:.1f}s excluded"),
            "seconds": est}
        out["gpu_over_cpu"] = out["solves_per_s"] / out["cpu_baseline"]["value"]
        log(f"[box cpu] {per_cycle:.2f}s per cycle -> {est:.0f}s for the transient")
    return out


def multishift(mk, workloads, steps, restart, sequential_sample=16, tune=True):
    """C4: 64 log-spaced shifts x 4 random L1-normalised rhs over the 58,905-state 1D box, one
    batch of 256 systems (FP32 sector multigrid, per-system FGMRES with delayed CGS2), each to
    1e-10 relative residual, with the sweep counts picked by the autotuner (trial batches,
    untimed); beside it the same systems solved one at a time on the GPU (a bounded sample of
    them, each with its own red-black sector multigrid)."""
    from markovkit_amd import autotune
    w = workloads.config_multishift()
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    n = box.size()
    rhs = []
    for seed in w.seeds:
        b = np.random.default_rng(seed).random(n)
        rhs.append(b / b.sum())
    shifts = [s for s in w.shifts for _ in rhs]
    B = np.array([b for _ in w.shifts for b in rhs])
    tol = 1e-10 * np.linalg.norm(B, axis=1)
    tuned = None
    if tune:
        t = time.perf_counter()
        tuned = autotune.tune_batch(Q, w.model.domain, box, shifts, B, restart=restart)
        tuned_s = time.perf_counter() - t
        names = {mk.JACOBI: "jacobi", mk.RED_BLACK: "red_black"}
        log(f"[multishift] autotuned {names.get(tuned.smoother)} {tuned.pre}/{tuned.post} in "
            f"{tuned_s:.1f}s: " + ", ".join(f"{names.get(tr.smoother)} {tr.pre}/{tr.post}="
                                            f"{tr.seconds * 1e3:.1f}ms" for tr in tuned.trials))
    t = time.perf_counter()
    M = mk.SectorMultigrid(w.model.domain, box, 1.0,
                           tuned.options(precision=mk.FP32) if tuned else mk.MultigridOptions(precision=mk.FP32))
    batch = mk.BatchSolver(Q, M, shifts)
    setup = time.perf_counter() - t
    batch.solve(B, tolerances=tol, restart=restart)  # warm-up
    mk.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        X, res = batch.solve(B, tolerances=tol, restart=restart)
        lat.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    if not all(r.converged for r in res):
        raise RuntimeError("multishift batch: a system did not converge")
    worst = max(float(np.linalg.norm(B[j] - Q.apply_shifted(shifts[j], X[j])) / np.linalg.norm(B[j]))
                for j in range(len(shifts)))
    its = [r.iterations for r in res]
    # sequential single-system solves on the GPU, a sample spread over the shifts, each system
    # solved once as a caller would
    pick = np.linspace(0, len(shifts) - 1, sequential_sample).round().astype(int)
    seq = 0.0
    for j in pick:
        Ms = mk.SectorMultigrid(w.model.domain, box, shifts[j],
                                mk.MultigridOptions(precision=mk.FP32, smoother=mk.RED_BLACK))
        A = mk.ShiftedGenerator(Q, shifts[j])
        x = np.zeros(n)
        mk.gmres(A, Ms, B[j], x, tolerance=tol[j], restart=restart)
        x[:] = 0.0
        t = time.perf_counter()
        r1 = mk.gmres(A, Ms, B[j], x, tolerance=tol[j], restart=restart)
        seq += time.perf_counter() - t
        if not r1.converged:
            raise RuntimeError("multishift: a single solve did not converge")
    seq_rate = len(pick) / seq
    # block Krylov: the 4 right-hand sides of every shift share one basis for their first restart
    # cycle, block Gram-Schmidt projections on FP64 matrix cores
    bsolver = mk.BatchSolver(Q, M, shifts, block=4)
    bsolver.solve(B, tolerances=tol, restart=restart)  # warm-up
    mk.synchronize()
    blat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        Xb, resb = bsolver.solve(B, tolerances=tol, restart=restart)
        blat.append(time.perf_counter() - t)
    btotal = time.perf_counter() - t0
    if not all(r.converged for r in resb):
        raise RuntimeError("multishift block batch: a system did not converge")
    bworst = max(float(np.linalg.norm(B[j] - Q.apply_shifted(shifts[j], Xb[j])) / np.linalg.norm(B[j]))
                 for j in range(len(shifts)))
    bits = [r.iterations for r in resb]
    block = {"value": len(shifts) * steps / btotal, "batch_ms": 1e3 * btotal / steps,
             "batch_p50_ms": 1e3 * percentile(blat, 50), "iterations_min": min(bits),
             "iterations_max": max(bits), "worst_relative_residual": bworst,
             "method": "block FGMRES(restart) per shift (4 rhs, shared basis, block CGS2 with "
                       "V^T W on v_mfma_f64_16x16x4f64, Cholesky QR twice), then per-system restarts"}
    log(f"[multishift] block: {block['value']:.1f} solves/s, iterations {min(bits)}-{max(bits)}")
    return {
        "workload": f"{w.name}: {n} states, {len(shifts)} systems (64 shifts 1e-2..1e2 x 4 rhs), "
                    f"FP32 sector multigrid, FGMRES({restart}) + delayed CGS2, tol 1e-10 relative",
        "value": len(shifts) * steps / total,
        "unit": "solves/s",
        "batch_ms": 1e3 * total / steps,
        "batch_p50_ms": 1e3 * percentile(lat, 50),
        "batch_p99_ms": 1e3 * percentile(lat, 99),
        "iterations_min": min(its),
        "iterations_max": max(its),
        "worst_relative_residual": worst,
        "setup_s": setup,
        "sequential_gpu_solves_per_s": seq_rate,
        "batched_over_sequential": (len(shifts) * steps / total) / seq_rate,
        "block": block,
        "autotuned_sweeps": [tuned.pre, tuned.post] if tuned else None,
        "autotuned_smoother": ({mk.JACOBI: "jacobi", mk.RED_BLACK: "red_black"}.get(tuned.smoother)
                               if tuned else None),
        "autotune_trials": ([{"smoother": {mk.JACOBI: "jacobi", mk.RED_BLACK: "red_black"}.get(tr.smoother),
                              "sweeps": [tr.pre, tr.post], "ms": 1e3 * tr.seconds,
                              "max_iterations": tr.iterations}
                             for tr in tuned.trials] if tuned else None),
    }


def expv_leg(mk, vectors=64, seed=2):
    """The serving mix's exponentials alone: `vectors` toggle-switch point masses (0..200 x 0..200,
    absorbing box, states drawn with numpy PCG64 seed 2) advanced to t = 5 at tolerance 1e-8 with
    Krylov dimension 30, as one batched call (delayed CGS2 Arnoldi, Expokit step control per vector),
    and one vector alone through the single-vector expv."""
    from markovkit_amd.fsp import box_generator, toggle_switch
    F = mk.Generator.from_jumps(*box_generator(toggle_switch(), [200, 200], absorbing=True))
    n = F.rows()
    rng = np.random.default_rng(seed)
    V = np.zeros((vectors, n))
    V[np.arange(vectors), rng.integers(0, n, vectors)] = 1.0
    mk.expv_batch(F, np.full(vectors, 0.1), V)  # warm-up
    mk.synchronize()
    t0 = time.perf_counter()
    W, res = mk.expv_batch(F, np.full(vectors, 5.0), V, krylov_dim=30, tolerance=1e-8)
    batch_s = time.perf_counter() - t0
    mk.expv(F, 0.1, V[0])
    t0 = time.perf_counter()
    w1, r1 = mk.expv(F, 5.0, V[0], krylov_dim=30, tolerance=1e-8)
    single_s = time.perf_counter() - t0
    steps = [r.steps for r in res]
    return {"workload": f"toggle switch, {n} states (absorbing 0..200 x 0..200), {vectors} point masses "
                        f"to t = 5 at 1e-8, Krylov dimension 30",
            "vectors_per_s": vectors / batch_s, "batch_ms": 1e3 * batch_s,
            "ms_per_vector": 1e3 * batch_s / vectors, "steps_min": min(steps), "steps_max": max(steps),
            "single_vector_ms": 1e3 * single_s, "single_steps": r1.steps,
            "max_l1_batch0_vs_single": float(np.abs(W[0] - w1).sum()),
            "min_mass": float(W.sum(axis=1).min())}


def serving(mk, workloads, n_requests=120, workers=4, seed=1, rate=200.0, smoother_name="red_black"):
    """C5: Poisson arrivals at `rate` requests/s (numpy PCG64, seed 1). 70% are RDME resolvent
    requests of the 1D family (voxels in {16, 32, 48, 64}, caps in {3, 4, 5}, shift ~ U[0.1, 10],
    b ~ U[0,1) L1-normalised, tolerance 1e-10 relative, FGMRES(50) with an FP32 sector
    multigrid), 30% toggle-switch FSP exponentials (0..200 x 0..200 = 40,401 states, absorbing
    box, p0 at a random state, exp(5 Q) p0 at tolerance 1e-8). Families are set up before the
    arrivals start and stay resident. Continuous batching: on 16 and 32 voxels one family on the
    cap-5 box serves all three caps (a request of cap k is the box's leading sectors 0..k), and
    every family of at most 524,288 states runs in a batch engine that admits requests at restart
    boundaries; the larger families run on `workers` worker streams with two hierarchies each
    (above a million states: 1 + 2 sweeps, W-cycles on the top 3 levels, dense sectors of at most
    100 states, whose terminal inverses a new shift recomputes quickly). Requests are submitted
    at their arrival times."""
    from markovkit_amd.fsp import box_generator, toggle_switch
    from markovkit_amd.model import box_size
    smoother = getattr(mk, SMOOTHERS[smoother_name])
    rng = np.random.default_rng(seed)
    kinds = rng.random(n_requests) < 0.7
    voxels = rng.choice([16, 32, 48, 64], n_requests)
    caps = rng.choice([3, 4, 5], n_requests)
    shifts = rng.uniform(0.1, 10.0, n_requests)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    shared = (16, 32)  # grids whose cap-5 box serves every cap
    t = time.perf_counter()
    built = {}
    for v, c in sorted({(int(v), 5 if int(v) in shared else int(c))
                        for v, c, k in zip(voxels, caps, kinds) if k}):
        w = workloads.birth_death_1d(v, c, 1.0, "random")
        box = mk.PopulationBox(w.n_voxels, w.max_totals)
        Q = mk.Generator(box, w.model)
        big = box.size() > 1_000_000
        opts = mk.MultigridOptions(precision=mk.FP32, smoother=smoother,
                                   **(dict(pre_sweeps=1, post_sweeps=2, coarse_cycles=2,
                                           cycle_levels=3, coarsest_size=100) if big else {}))
        mgs = [mk.SectorMultigrid(w.model.domain, box, 1.0, opts) for _ in range(2)]
        built[(v, c)] = (Q, mgs)
    J = box_generator(toggle_switch(), [200, 200], absorbing=True)
    F = mk.Generator.from_jumps(*J)
    setup = time.perf_counter() - t
    srv = mk.Server(workers=workers, hbm_budget=96 << 30)
    fid = {key: srv.add_family(Q, mgs) for key, (Q, mgs) in built.items()}
    batched = {key for key, (Q, _) in built.items() if Q.rows() <= 524288}
    log(f"[serving] {len(built)} resolvent families ({len(batched)} batched) and the toggle switch "
        f"set up in {setup:.1f}s")

    def route(v, c):
        key = (v, 5) if v in shared else (v, c)
        return key, box_size(v, [c])

    payload = []
    for i in range(n_requests):
        if kinds[i]:
            _, rows = route(int(voxels[i]), int(caps[i]))
            b = rng.random(rows)
            payload.append(b / b.sum())
        else:
            v = np.zeros(F.rows())
            v[int(rng.integers(0, F.rows()))] = 1.0
            payload.append(v)
    # warm-up: one request of every family and an exponential
    warm = [srv.submit_resolvent(fid[key], 1.0, np.full(Q.rows(), 1.0 / Q.rows()), 1e-8)
            for key, (Q, _) in built.items()]
    v0 = np.zeros(F.rows())
    v0[0] = 1.0
    warm.append(srv.submit_expv(F, 0.1, v0))
    for r in warm:
        srv.wait(r)
    ids, cls = [], []
    t0 = time.perf_counter()
    for i in range(n_requests):
        delay = t0 + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        if kinds[i]:
            b = payload[i]
            key, _ = route(int(voxels[i]), int(caps[i]))
            ids.append(srv.submit_resolvent(fid[key], float(shifts[i]), b,
                                            1e-10 * float(np.linalg.norm(b))))
            cls.append("resolvent_batched" if key in batched else "resolvent_worker")
        else:
            ids.append(srv.submit_expv(F, 5.0, payload[i], tolerance=1e-8))
            cls.append("expv")
    lat, run, by = [], [], {}
    for i, r in enumerate(ids):
        _, info = srv.wait(r)
        if info.kind == "resolvent" and not info.solve.converged:
            raise RuntimeError("serving: a resolvent request did not converge")
        latency = info.queued_s + info.run_s
        lat.append(latency)
        run.append(info.run_s)
        by.setdefault(cls[i], []).append((latency, info.run_s))
    total = time.perf_counter() - t0
    srv.close()
    per_kind = {k: {"requests": len(v), "p50_ms": 1e3 * percentile([a for a, _ in v], 50),
                    "p99_ms": 1e3 * percentile([a for a, _ in v], 99),
                    "mean_run_ms": 1e3 * float(np.mean([b for _, b in v]))} for k, v in by.items()}
    return {
        "workload": (f"{n_requests} requests, Poisson {rate:.0f}/s (seed {seed}): 70% 1D RDME "
                     f"resolvents (voxels 16-64, caps 3-5, s ~ U[0.1,10], 1e-10 relative), 30% "
                     f"toggle-switch FSP expv (40,401 states, t = 5, 1e-8); continuous batching for "
                     f"families of <= 524,288 states, {workers} worker streams for the rest; "
                     f"{smoother_name} smoother"),
        "value": n_requests / total,
        "unit": "requests/s",
        "offered_rate": rate,
        "p50_ms": 1e3 * percentile(lat, 50),
        "p99_ms": 1e3 * percentile(lat, 99),
        "mean_run_ms": 1e3 * float(np.mean(run)),
        "per_kind": per_kind,
        "makespan_s": total,
        "setup_s": setup,
        "families": len(built),
        "batched_families": len(batched),
    }


def measure_traffic(args, kernel_class):
    """HBM bytes per launch of the finest level's sweep kernel, measured in this run: this script
    runs again in --traffic-probe mode (setup, a 256 MiB copy with the library's plain copy
    kernel, two solves) under `rocprofv3 --pmc FETCH_SIZE`, then under `--pmc WRITE_SIZE`, one
    counter per pass; each counter is scaled to bytes by what it reads for the copy (256 MiB read,
    256 MiB written). Returns None with the reason when rocprofv3 is absent or a pass fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof_bin = shutil.which("rocprofv3")
    if prof_bin is None:
        return {"bytes_per_launch": None, "error": "rocprofv3 not found"}
    mode = "csr_mode)5" if args.smoother == "red_black" else "csr_mode)2"
    copy_bytes = 256 << 20
    passthrough = ["--workload", args.workload, "--pre", str(args.pre), "--post", str(args.post),
                   "--cycles", str(args.cycles), "--cycle-levels", str(args.cycle_levels),
                   "--coarsest", str(args.coarsest), "--smoother", args.smoother,
                   "--restart", str(args.restart)]
    out = {"kernel_class": kernel_class}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="mk_pmc_", dir="/tmp")
        cmd = [prof_bin, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "t", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe", *passthrough]
        env = dict(os.environ, TMPDIR="/tmp")
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=240)
        except subprocess.TimeoutExpired:
            return {"bytes_per_launch": None, "error": f"{counter} pass timed out"}
        path = os.path.join(d, "t_counter_collection.csv")
        if r.returncode != 0 or not os.path.exists(path):
            return {"bytes_per_launch": None, "error": f"{counter} pass failed ({r.returncode})"}
        copy, sweep = [], {}
        with open(path) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] != counter:
                    continue
                name, grid, v = row["Kernel_Name"], int(row["Grid_Size"]), float(row["Counter_Value"])
                if "k_copy16" in name:
                    copy.append(v)
                elif "k_csr_sliced<" in name and mode in name:
                    sweep.setdefault(grid, []).append(v)
        if not copy or not sweep:
            return {"bytes_per_launch": None, "error": f"{counter} pass: kernels not found"}
        scale = copy_bytes / (float(np.mean(copy)) * 1024.0)
        fine = max(sweep)  # the finest level's grid
        key = "fetch" if counter == "FETCH_SIZE" else "write"
        out[f"{key}_bytes"] = float(np.mean(sweep[fine])) * 1024.0 * scale
        out[f"{key}_scale"] = scale
        out["dispatches"] = len(sweep[fine])
        out["grid"] = fine
    out["bytes_per_launch"] = out["fetch_bytes"] + out["write_bytes"]
    out["source"] = ("rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, in this run (two solves of the "
                     "same hierarchy), each scaled by its reading for a 256 MiB copy")
    return out


def c3(mk, workloads, steps, restart):
    """2D two-species annihilation on 8 x 8 voxels (h = 1/8, D = 1, births 0.1 * volume each,
    A + B -> 0 at n_A n_B / volume), caps {2, 2}: 4,601,025 states, s = 1, b = e_0. FGMRES(restart)
    preconditioned by the red-black sector multigrid (V-cycle, 2 + 2 sweeps), mixed (FP32
    multigrid) and plain FP64, each to 1e-10 relative residual. Checked in the run: true
    residual, mass balance 1^T b = sum (s + leak) x, x >= -bound, and the mixed solution within
    the L1 bound of the FP64 one."""
    w = workloads.config_2d()
    shift = w.shifts[0]
    t = time.perf_counter()
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    n = box.size()
    mk.synchronize()
    t_gen = time.perf_counter() - t
    b = w.rhs(0, n)
    tol = 1e-10 * float(np.linalg.norm(b))
    A = mk.ShiftedGenerator(Q, shift)
    b_dev, x_dev = mk.DeviceVector(n, b), mk.DeviceVector(n)
    abs_A = shift + 2 * Q.exit_rates() - Q.leak_rates()
    out = {"workload": f"{w.name}: 2D two-species annihilation, {n} states, s={shift}, b=e_0, "
                       f"FGMRES({restart}) + red-black sector multigrid (V, 2+2), tol 1e-10 relative",
           "states": n, "setup_s": {"generator": t_gen}}
    xs = {}
    for name, prec in (("mixed", mk.FP32), ("fp64", mk.FP64)):
        t = time.perf_counter()
        M = mk.SectorMultigrid(w.model.domain, box, shift,
                               mk.MultigridOptions(precision=prec, smoother=mk.RED_BLACK))
        mk.synchronize()
        out["setup_s"][f"multigrid_{name}"] = time.perf_counter() - t
        run_solves(mk, A, M, b_dev, x_dev, tol, 2, restart)
        total, lat, iters, res, _ = run_solves(mk, A, M, b_dev, x_dev, tol, steps, restart)
        x = x_dev.download()
        r = b - Q.apply_shifted(shift, x)
        bound = float((np.abs(r).sum() + 16 * EPS * (abs_A * np.abs(x)).sum()) / shift)
        # 1^T b - sum (s + leak) x = 1^T r exactly, so it is at most ||r||_1 plus rounding
        mass = abs(float(b.sum()) - float(((shift + Q.leak_rates()) * x).sum()))
        mass_bound = float(np.abs(r).sum() + 64 * EPS * ((shift + Q.leak_rates()) * np.abs(x)).sum())
        rel = float(np.linalg.norm(r) / np.linalg.norm(b))
        if not (rel <= 1.0001e-10 and float(x.min()) >= -bound and mass <= mass_bound):
            raise RuntimeError(f"C3 {name}: residual {rel:.3e}, min x {x.min():.3e}, mass {mass:.3e}")
        xs[name] = (x, r)
        rec = {"value": steps / total, "p50_ms": 1e3 * percentile(lat, 50),
               "p99_ms": 1e3 * percentile(lat, 99), "iterations": iters[-1],
               "relative_residual": rel, "l1_bound": bound, "mass_balance_error": mass,
               "min_x": float(x.min())}
        if name == "mixed":
            out.update(rec)
        else:
            out["fp64"] = rec
        del M
    (xm, rm), (x64, r64) = xs["mixed"], xs["fp64"]
    out["l1_mixed_vs_fp64"] = float(np.abs(xm - x64).sum())
    out["l1_mixed_vs_fp64_bound"] = float((np.abs(rm).sum() + np.abs(r64).sum() + 16 * EPS *
                                           (abs_A * (np.abs(xm) + np.abs(x64))).sum()) / shift)
    if out["l1_mixed_vs_fp64"] > out["l1_mixed_vs_fp64_bound"]:
        raise RuntimeError(f"C3: mixed and FP64 solutions differ by {out['l1_mixed_vs_fp64']:.3e}")
    out["mixed_over_fp64"] = out["value"] / out["fp64"]["value"]
    return out


def c1(mk, workloads, restart, systems=256, single_steps=200):
    """The reference's CPU-runnable case: 1D, 16 voxels, cap 3 (969 states), s = 0.5. A single
    solve of (0.5 I - Q) x = e_0 (mixed and plain FP64, the reference's weighted-Jacobi sector
    multigrid 2+2, FGMRES; and mixed with the red-black smoother) runs as one kernel launch in
    one workgroup, bound by the latency of its dependent barrier phases, so the case is also run
    as one batch of `systems` systems of the same shift: e_0 and random L1-normalised right-hand
    sides (seeds 1..systems-1), each to 1e-10 relative residual. The CPU oracle (the reference's
    method: weighted-Jacobi V-cycle, GMRES with MGS; one core) solves the same systems one by
    one, timed beside it."""
    from oracle.oracle import Problem, SectorMultigrid as OracleMG
    w = workloads.config_small()
    shift = w.shifts[0]
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    n = box.size()
    A = mk.ShiftedGenerator(Q, shift)
    b = w.rhs(0, n)
    tol = 1e-10 * float(np.linalg.norm(b))
    b_dev, x_dev = mk.DeviceVector(n, b), mk.DeviceVector(n)
    out = {"workload": f"{w.name}: {n} states, s={shift}; single solve of b=e_0, and a batch of "
                       f"{systems} systems (e_0 and random L1-normalised rhs, seeds 1..{systems - 1}); "
                       f"FGMRES({restart}) + red-black sector multigrid (2+2), tol 1e-10 relative"}
    for name, prec, smoother in (("mixed", mk.FP32, mk.JACOBI), ("fp64", mk.FP64, mk.JACOBI),
                                 ("mixed_red_black", mk.FP32, mk.RED_BLACK)):
        M = mk.SectorMultigrid(w.model.domain, box, shift,
                               mk.MultigridOptions(precision=prec, smoother=smoother))
        run_solves(mk, A, M, b_dev, x_dev, tol, 100, restart)
        total, lat, iters, _, _ = run_solves(mk, A, M, b_dev, x_dev, tol, single_steps, restart)
        x = x_dev.download()
        rel = float(np.linalg.norm(b - Q.apply_shifted(shift, x)) / np.linalg.norm(b))
        if rel > 1.0001e-10:
            raise RuntimeError(f"C1 single {name}: relative residual {rel:.3e}")
        out[f"single_{name}"] = {"value": single_steps / total, "p50_ms": 1e3 * percentile(lat, 50),
                                 "iterations": iters[-1], "relative_residual": rel}
        del M
    B = np.empty((systems, n))
    B[0] = b
    for j in range(1, systems):
        r = np.random.default_rng(j).random(n)
        B[j] = r / r.sum()
    tols = 1e-10 * np.linalg.norm(B, axis=1)
    M = mk.SectorMultigrid(w.model.domain, box, shift,
                           mk.MultigridOptions(precision=mk.FP32, smoother=mk.RED_BLACK))
    batch = mk.BatchSolver(Q, M, [shift] * systems)
    batch.solve(B, tolerances=tols, restart=restart)  # warm-up
    mk.synchronize()
    steps = 5
    t0 = time.perf_counter()
    for _ in range(steps):
        X, res = batch.solve(B, tolerances=tols, restart=restart)
    total = time.perf_counter() - t0
    if not all(r.converged for r in res):
        raise RuntimeError("C1 batch: a system did not converge")
    worst = max(float(np.linalg.norm(B[j] - Q.apply_shifted(shift, X[j])) / np.linalg.norm(B[j]))
                for j in range(systems))
    its = [r.iterations for r in res]
    out["batched"] = {"value": systems * steps / total, "batch_ms": 1e3 * total / steps,
                      "iterations_min": min(its), "iterations_max": max(its),
                      "worst_relative_residual": worst}
    # the same batch with the reference's weighted-Jacobi hierarchy
    Mj = mk.SectorMultigrid(w.model.domain, box, shift, mk.MultigridOptions(precision=mk.FP32))
    bj = mk.BatchSolver(Q, Mj, [shift] * systems)
    bj.solve(B, tolerances=tols, restart=restart)  # warm-up
    mk.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        Xj, resj = bj.solve(B, tolerances=tols, restart=restart)
    tj = time.perf_counter() - t0
    if not all(r.converged for r in resj):
        raise RuntimeError("C1 Jacobi batch: a system did not converge")
    worst_j = max(float(np.linalg.norm(B[j] - Q.apply_shifted(shift, Xj[j])) / np.linalg.norm(B[j]))
                  for j in range(systems))
    itsj = [r.iterations for r in resj]
    out["batched_jacobi"] = {"value": systems * steps / tj, "batch_ms": 1e3 * tj / steps,
                             "iterations_min": min(itsj), "iterations_max": max(itsj),
                             "worst_relative_residual": worst_j}
    del bj, Mj
    # the CPU oracle on the same systems, one core
    P = Problem(w.model, w.max_totals)
    Mo = OracleMG(P, shift)
    P.gmres(shift, b, mg=Mo, tol=tol, restart=restart)  # warm-up
    l1 = 0.0
    t0 = time.perf_counter()
    for j in range(systems):
        xo, info = P.gmres(shift, B[j], mg=Mo, tol=float(tols[j]), restart=restart)
        if not info["converged"]:
            raise RuntimeError("C1: an oracle solve did not converge")
        l1 = max(l1, float(np.abs(xo - X[j]).sum()))
    cpu = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": systems / cpu, "unit": "solves/s", "cores": 1, "kind": "port",
                           "sample": f"the {systems} systems of the batch solved one by one by the "
                                     f"FP64 CPU oracle (weighted-Jacobi V-cycle 2+2, GMRES({restart}) "
                                     f"with MGS), one core"}
    out["max_l1_batch_vs_oracle"] = l1
    out["batched_over_cpu"] = out["batched"]["value"] / out["cpu_baseline"]["value"]
    out["single_mixed_over_cpu"] = out["single_mixed"]["value"] / out["cpu_baseline"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large",
                    choices=["large", "small", "2d", "multishift", "serving", "box", "expv", "c1"])
    ap.add_argument("--requests", type=int, default=120, help="serving: number of requests")
    ap.add_argument("--workers", type=int, default=4, help="serving: worker streams")
    ap.add_argument("--restart", type=int, default=50)
    ap.add_argument("--pre", type=int, default=None,
                    help="pre-smoothing sweeps (multigrid_options; default: 1 on the large "
                         "workload, else the reference's 2)")
    ap.add_argument("--post", type=int, default=None,
                    help="post-smoothing sweeps (multigrid_options; default: the reference's 2)")
    ap.add_argument("--cycles", type=int, default=None,
                    help="coarse_cycles of the timed hierarchy (1 V-cycle, 2 W-cycle; default: 2 "
                         "on the large workload, the reference's 1 on the others)")
    ap.add_argument("--cycle-levels", type=int, default=None,
                    help="levels that take --cycles coarse visits (0: all; the ones below take "
                         "one; default: 2 on the large workload, else 0)")
    ap.add_argument("--coarsest", type=int, default=None,
                    help="coarsest_size of the timed hierarchy (sectors of at most this many "
                         "states are solved densely; default: 800 on the large workload, else "
                         "the reference's default)")
    ap.add_argument("--smoother", default="red_black", choices=list(SMOOTHERS),
                    help="sector multigrid smoother of the timed and served solves (red_black: "
                         "multicolour Gauss-Seidel; jacobi: the reference's default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched multi-shift solves with the default sweep counts")
    ap.add_argument("--no-jacobi", action="store_true",
                    help="skip the comparison solve with the reference's Jacobi smoother")
    ap.add_argument("--no-multishift", action="store_true",
                    help="skip the batched multi-shift measurement after the headline")
    ap.add_argument("--box-solves", type=int, default=None,
                    help="box workload: time only the first N solves of the transient (profiling)")
    ap.add_argument("--no-box", action="store_true",
                    help="skip the whole-box multigrid transient after the headline")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving mix after the headline")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the 2D two-species annihilation solves after the headline")
    ap.add_argument("--no-expv", action="store_true",
                    help="skip the batched toggle-switch exponentials after the headline")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 969-state single, batched and CPU-oracle solves after the headline")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the HBM-counter passes (rocprofv3 --pmc) of the roofline traffic")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="internal: setup, a 256 MiB copy and two solves, for the counter passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (each solve "
                          "runs on one MI355X; sharding is described in DESIGN.md)"}))
        sys.exit(2)

    import markovkit_amd as mk
    from markovkit_amd import workloads

    if args.workload == "expv":
        ev = expv_leg(mk)
        out = {"metric": "batched toggle-switch exponentials (t = 5, 1e-8) per second",
               "value": ev["vectors_per_s"], "unit": "vectors/s", "n_gpus": 1, "steps": 1, "warmup": 1,
               "ms_per_step": ev["batch_ms"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": {"workload": ev["workload"]}, "expv": ev}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "serving":
        sv = serving(mk, workloads, args.requests, args.workers, smoother_name=args.smoother)
        out = {"metric": "served RDME resolvent + FSP requests/s with p50/p99 latency",
               "value": sv["value"], "unit": "requests/s", "n_gpus": 1, "steps": args.requests,
               "warmup": 1, "ms_per_step": 1e3 * sv["makespan_s"] / args.requests,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f64 (f32 multigrid)", "data": "synthetic",
               "config": {"workload": sv["workload"]}, "serving": sv}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "c1":
        c = c1(mk, workloads, args.restart)
        out = {"metric": "single RDME resolvent solves/s at <=1e-10 residual (969-state case)",
               "value": c["single_mixed"]["value"], "unit": "solves/s", "n_gpus": 1,
               "steps": 200, "warmup": 100, "ms_per_step": 1e3 / c["single_mixed"]["value"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f64 (f32 multigrid)", "data": "synthetic",
               "config": {"workload": c["workload"]}, "cpu_baseline": c["cpu_baseline"], "c1": c}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "box":
        bx = box_mg(mk, cpu=not args.no_cpu, solves=args.box_solves)
        out = {"metric": "whole-box multigrid time-step solves/s (paper driver's CN transient)",
               "value": bx["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": bx["solves"],
               "warmup": bx["solves"], "ms_per_step": 1e3 * bx["seconds"] / bx["solves"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": "synthetic (the driver's initial state)",
               "config": {"workload": bx["workload"]}, "roofline": bx["roofline"],
               "cpu_baseline": bx["cpu_baseline"], "box_mg": bx}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "multishift":
        ms = multishift(mk, workloads, args.steps, args.restart, tune=not args.no_autotune)
        out = {"metric": "batched multi-shift RDME resolvent solves/s at <=1e-10 residual",
               "value": ms["value"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
               "warmup": 1, "ms_per_step": ms["batch_ms"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64 (f32 multigrid)",
               "data": "synthetic", "config": {"workload": ms["workload"]}, "multishift": ms}
        print(json.dumps(out), flush=True)
        return

    w = {"large": workloads.config_large, "small": workloads.config_small,
         "2d": workloads.config_2d}[args.workload]()
    shift = w.shifts[0]
    log(f"[setup] {w.name}: {w.states} states")
    t = time.perf_counter()
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    t_gen = time.perf_counter() - t
    log(f"[setup] generator {t_gen:.1f}s, {Q.n_jumps()} jumps")
    t = time.perf_counter()
    smoother = getattr(mk, SMOOTHERS[args.smoother])
    # the large workload's hierarchy: 1 + 2 sweeps, W-cycles on the top 2 levels, dense sectors
    # of <= 800 states (the 8-voxel level: blocks of up to 792 states, a wave per row)
    ref = mk.MultigridOptions()
    tuned_large = args.workload == "large"
    if args.pre is None:
        args.pre = 1 if tuned_large else ref.pre_sweeps
    if args.post is None:
        args.post = ref.post_sweeps
    if args.cycles is None:
        args.cycles = 2 if tuned_large else ref.coarse_cycles
    if args.cycle_levels is None:
        args.cycle_levels = 2 if tuned_large else ref.cycle_levels
    if args.coarsest is None:
        args.coarsest = 800 if tuned_large else ref.coarsest_size
    opts = dict(pre_sweeps=args.pre, post_sweeps=args.post, coarse_cycles=args.cycles,
                cycle_levels=args.cycle_levels, coarsest_size=args.coarsest)
    M32 = mk.SectorMultigrid(w.model.domain, box, shift,
                             mk.MultigridOptions(precision=mk.FP32, smoother=smoother, **opts))
    t_mg = time.perf_counter() - t
    log(f"[setup] FP32 multigrid {t_mg:.1f}s, {M32.max_levels()} levels")
    A = mk.ShiftedGenerator(Q, shift)
    n = box.size()
    b = w.rhs(0, n)
    tol = 1e-10 * float(np.linalg.norm(b))
    b_dev = mk.DeviceVector(n, b)
    x_dev = mk.DeviceVector(n)
    if args.traffic_probe:
        from markovkit_amd.calibrate import copy_bandwidth
        copy_bandwidth(256 << 20, 1)
        run_solves(mk, A, M32, b_dev, x_dev, tol, 2, args.restart)
        return

    # warm-up, and a fully profiled solve (the per-class breakdown)
    # solves of about a millisecond need more warm-up before the clocks settle (both legs)
    warm = max(args.warmup, 1) if n > 100_000 else max(args.warmup, 100)
    run_solves(mk, A, M32, b_dev, x_dev, tol, warm, args.restart)
    _, _, _, _, prof_all = run_solves(mk, A, M32, b_dev, x_dev, tol, 1, args.restart,
                                      profile_classes=True)
    # the roofline kernel: the finest level's smoothing sweep, the path's dominant HBM stream
    # (the coarse levels' many small launches are latency-bound, and events around each of them
    # would distort both their share and the timed region)
    dominant = "smooth_sweep"
    log("[profile] " + ", ".join(f"{k}={v['ms']:.2f}ms/{v['launches']}" for k, v in
                                 prof_all.items() if v["launches"]))

    # timed mixed region, events only around the dominant class
    total, lat, iters, res, prof = run_solves(mk, A, M32, b_dev, x_dev, tol, args.steps,
                                              args.restart, profile_classes=[dominant])
    x_mixed = x_dev.download()
    log(f"[mixed] {args.steps} solves in {total:.3f}s, {iters[-1]} iterations, "
        f"residual {res.residual:.3e}")
    r = b - Q.apply_shifted(shift, x_mixed)
    rel_res = float(np.linalg.norm(r) / np.linalg.norm(b))
    # ||x - x*||_1 <= ||b - A x||_1 / s, with the FP64 evaluation error of b - A x (at most
    # 16 eps || |A| |x| ||_1, where || |A| |x| ||_1 = sum_j (s + 2 exit_j - leak_j) |x_j|)
    abs_A = shift + 2 * Q.exit_rates() - Q.leak_rates()
    l1_bound = float((np.abs(r).sum() + 16 * EPS * (abs_A * np.abs(x_mixed)).sum()) / shift)

    fp64 = None
    if not args.no_fp64:
        t = time.perf_counter()
        M64 = mk.SectorMultigrid(w.model.domain, box, shift,
                                 mk.MultigridOptions(precision=mk.FP64, smoother=smoother,
                                                     **opts))
        log(f"[setup] FP64 multigrid {time.perf_counter() - t:.1f}s")
        run_solves(mk, A, M64, b_dev, x_dev, tol, warm, args.restart)
        # the same events as the mixed timed region (around the dominant class), so that both
        # legs carry the same timing cost
        total64, lat64, iters64, res64, _ = run_solves(mk, A, M64, b_dev, x_dev, tol, args.steps,
                                                       args.restart, profile_classes=[dominant])
        x64 = x_dev.download()
        log(f"[fp64] {args.steps} solves in {total64:.3f}s, {iters64[-1]} iterations")
        r64 = b - Q.apply_shifted(shift, x64)
        fp64 = {
            "value": args.steps / total64,
            "p50_ms": 1e3 * percentile(lat64, 50),
            "p99_ms": 1e3 * percentile(lat64, 99),
            "iterations": iters64[-1],
            "relative_residual": float(np.linalg.norm(r64) / np.linalg.norm(b)),
            "l1_mixed_vs_fp64": float(np.abs(x_mixed - x64).sum()),
            "l1_bound": float((np.abs(r).sum() + np.abs(r64).sum() +
                               16 * EPS * (abs_A * (np.abs(x_mixed) + np.abs(x64))).sum()) / shift),
        }
        del M64

    # sector sums against the lumped closed form (1D linear birth-death boxes)
    lumped = None
    if len(w.max_totals) == 1:
        K = box.n_sectors()
        begins = [box.sector_begin(k) for k in range(K)] + [n]
        Bk = np.array([b[begins[k]:begins[k + 1]].sum() for k in range(K)])
        Xk = np.array([x_mixed[begins[k]:begins[k + 1]].sum() for k in range(K)])
        X = workloads.lumped_sector_sums(shift, 0.05, 0.02, w.max_totals[0], Bk)
        lumped = float(np.abs(Xk - X).sum())
        if lumped > l1_bound:
            raise RuntimeError(f"sector sums off the closed form by {lumped:.3e}, above the L1 "
                               f"bound {l1_bound:.3e} of the solution")

    # a problem whose whole hierarchy fits one workgroup is solved by one launch per solve: the
    # per-class breakdown and the roofline kernel then come from the host-driven loop, profiled
    # on its own (the timed solves above stay one launch each)
    roofline_path = "the timed solves"
    if prof[dominant]["launches"] == 0:
        mk.tiny_gmres_enable(False)
        try:
            run_solves(mk, A, M32, b_dev, x_dev, tol, warm, args.restart)
            _, _, _, _, prof_all = run_solves(mk, A, M32, b_dev, x_dev, tol, 1, args.restart,
                                              profile_classes=True)
            _, _, _, _, prof = run_solves(mk, A, M32, b_dev, x_dev, tol, args.steps, args.restart,
                                          profile_classes=[dominant])
        finally:
            mk.tiny_gmres_enable(True)
        roofline_path = ("the host-driven Krylov loop, profiled separately: the timed solves run "
                         "as one workgroup launch each")
    d = prof[dominant]
    launches = d["launches"]
    avg_ms = d["ms"] / launches
    bytes_per_launch = d["bytes"] / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    measured = None
    if args.workload == "large" and not args.no_traffic:
        log("[traffic] FETCH_SIZE and WRITE_SIZE passes (rocprofv3 --pmc)")
        measured = measure_traffic(args, dominant)
        log(f"[traffic] {measured}")
    traffic = measured.get("bytes_per_launch") if measured else None

    out = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (f32 multigrid)",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: {len(w.model.domain.graph.dims)}D RDME resolvent, "
                        f"{len(w.max_totals)} species, {n} states, s={shift}, "
                        f"{'b~U[0,1) seed 0, L1-normalised' if w.rhs_kind == 'random' else 'b=e_0'}, "
                        f"FGMRES({args.restart}) + sector multigrid, tol 1e-10 relative",
            "states": n,
            "jumps": Q.n_jumps(),
            "multigrid_levels": M32.max_levels(),
            "smoother": args.smoother,
            "sweeps": [args.pre, args.post],
            "coarse_cycles": args.cycles,
            "cycle_levels": args.cycle_levels,
            "coarsest_size": args.coarsest,
        },
        "p50_ms": 1e3 * percentile(lat, 50),
        "p99_ms": 1e3 * percentile(lat, 99),
        "iterations": iters[-1],
        "relative_residual": rel_res,
        "l1_bound": l1_bound,
        "lumped_sector_sum_error": lumped,
        "fp64": fp64,
        "mixed_over_fp64": (args.steps / total) / fp64["value"] if fp64 else None,
        "setup_s": {"generator": t_gen, "multigrid_fp32": t_mg},
        "roofline": {
            "kernel": dominant,
            "path": roofline_path,
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "traffic_measurement": measured,
            "launches": launches,
            "avg_launch_ms": avg_ms,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "kernel_ms_per_solve": {k: v["ms"] for k, v in prof_all.items() if v["launches"]},
        # achieved algorithmic GB/s of every kernel class over the fully profiled solve
        "kernel_gbs": {k: v["bytes"] / v["ms"] / 1e6 for k, v in prof_all.items()
                       if v["launches"] and v["bytes"] > 0},
        "cpu_baseline": None,
    }
    if args.workload == "large" and not args.no_multishift:
        log("[multishift] 256-system batch")
        out["multishift"] = multishift(mk, workloads, 3, args.restart, sequential_sample=8,
                                       tune=not args.no_autotune)
        log(f"[multishift] {out['multishift']['value']:.1f} solves/s")
    if args.workload == "large" and not args.no_box:
        log("[box] whole-box multigrid transient")
        out["box_mg"] = box_mg(mk, cpu=False)
    ref_iters = iters[-1]
    if args.smoother != "jacobi" and not args.no_jacobi:
        # the reference's own method (weighted Jacobi V-cycle) on the GPU, the method the CPU
        # oracle solve below runs
        del M32
        Mj = mk.SectorMultigrid(w.model.domain, box, shift, mk.MultigridOptions(precision=mk.FP32))
        run_solves(mk, A, Mj, b_dev, x_dev, tol, 1, args.restart)
        tj, latj, itj, _, _ = run_solves(mk, A, Mj, b_dev, x_dev, tol, max(3, args.steps // 4),
                                         args.restart)
        ref_iters = itj[-1]
        out["jacobi"] = {"value": len(latj) / tj, "p50_ms": 1e3 * percentile(latj, 50),
                         "iterations": ref_iters}
        log(f"[jacobi] {out['jacobi']['value']:.2f} solves/s, {ref_iters} iterations")
        del Mj
    if args.workload == "large" and not args.no_c3:
        log("[c3] 2D two-species annihilation, 4,601,025 states")
        out["c3"] = c3(mk, workloads, max(10, args.steps), args.restart)
        log(f"[c3] {out['c3']['value']:.1f} solves/s mixed, {out['c3']['fp64']['value']:.1f} FP64")
    if args.workload == "large" and not args.no_expv:
        log("[expv] batched toggle-switch exponentials")
        out["expv"] = expv_leg(mk)
    if args.workload == "large" and not args.no_c1:
        log("[c1] 1D 16 voxels cap 3, 969 states: single solves, a 256-system batch, the CPU oracle")
        out["c1"] = c1(mk, workloads, args.restart)
        log(f"[c1] single {out['c1']['single_mixed']['value']:.0f} solves/s, batched "
            f"{out['c1']['batched']['value']:.0f} solves/s, CPU oracle "
            f"{out['c1']['cpu_baseline']['value']:.0f} solves/s")
    if args.workload == "large" and not args.no_serving:
        log("[serving] serving mix")
        out["serving"] = serving(mk, workloads, args.requests, args.workers, smoother_name=args.smoother)
        log(f"[serving] {out['serving']['value']:.1f} requests/s, p50 {out['serving']['p50_ms']:.0f} ms, "
            f"p99 {out['serving']['p99_ms']:.0f} ms")
    from markovkit_amd.calibrate import copy_bandwidth
    out["stream_copy_gbs"] = copy_bandwidth()
    out["roofline"]["frac_of_stream_copy"] = achieved / out["stream_copy_gbs"]
    log(f"[copy] plain copy kernel {out['stream_copy_gbs']:.0f} GB/s read+write")
    if not args.no_cpu:
        out["cpu_baseline"], out["parity_vs_oracle"], _ = cpu_baseline(
            w, b, x_mixed, Q, tol, args.restart)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        if not out["parity_vs_oracle"]["within_bound"]:
            raise RuntimeError(f"GPU solution off the oracle's: {out['parity_vs_oracle']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
