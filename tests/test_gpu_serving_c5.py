"""Serving at the serving mix's full sizes: the twelve 1D birth-death resolvent families (voxels
{16, 32, 48, 64} x caps {3, 4, 5}, 969 to 11,238,513 states; births 0.05 * volume, deaths
0.02 * n, D = 1, h = 1 / voxels, examples/rdme_multigrid_benchmark.cpp:65-70) with FP32 red-black
hierarchies kept resident, and the toggle switch on its 0..200 x 0..200 box (40,401 states,
exp(5 Q) p0 at 1e-8).

Checks:
- every request served in the mix (concurrent, several families and exponentials at once) is
  bitwise the same as the same request served alone;
- every resolvent reaches the 1e-10 relative residual and its sector sums equal the lumped
  closed form (the chain is lumpable by total population for every family) within the L1 bound
  ||x - x*||_1 <= (||b - A x||_1 + 16 eps sum_j (s + 2 exit_j - leak_j) |x_j|) / s;
- resolvents of the families the oracle solves in about a second (<= 58,905 states) lie within
  the L1 bound of the oracle's FP64 solution, state by state;
- every exponential lies within 1e-7 in L1 of the oracle's expv (both carry local errors below
  1e-8 per unit time).
"""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.fsp import box_generator, toggle_switch
from markovkit_amd.workloads import birth_death_1d, lumped_sector_sums
from oracle.oracle import Problem, SectorMultigrid as OracleMG, expv_jumps

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
VOXELS = (16, 32, 48, 64)
CAPS = (3, 4, 5)


@pytest.fixture(scope="module")
def mix():
    fams = {}
    for v in VOXELS:
        for c in CAPS:
            w = birth_death_1d(v, c, 1.0, "random")
            box = mk.PopulationBox(w.n_voxels, w.max_totals)
            Q = mk.Generator(box, w.model)
            mgs = [mk.SectorMultigrid(w.model.domain, box, 1.0,
                                      mk.MultigridOptions(precision=mk.FP32, smoother=mk.RED_BLACK))
                   for _ in range(2)]
            fams[(v, c)] = (w, box, Q, mgs)
            print(f"[c5] family v={v} cap={c}: {box.size()} states", flush=True)
    J = box_generator(toggle_switch(), [200, 200], absorbing=True)
    F = mk.Generator.from_jumps(*J)
    assert F.rows() == 40401
    rng = np.random.default_rng(1)
    reqs = []
    for key in fams:
        for _ in range(2):
            n = fams[key][2].rows()
            b = rng.random(n)
            reqs.append(("resolvent", key, float(rng.uniform(0.1, 10.0)), b / b.sum()))
    for _ in range(6):
        p0 = np.zeros(F.rows())
        p0[int(rng.integers(0, F.rows()))] = 1.0
        reqs.append(("expv", None, 5.0, p0))
    order = rng.permutation(len(reqs))
    reqs = [reqs[i] for i in order]
    return fams, F, J, reqs


def serve(fams, F, reqs, workers, together):
    srv = mk.Server(workers=workers, hbm_budget=96 << 30)
    fid = {key: srv.add_family(f[2], f[3]) for key, f in fams.items()}
    out = []

    def submit(req):
        kind, key, p, vec = req
        if kind == "expv":
            return srv.submit_expv(F, p, vec, tolerance=1e-8)
        return srv.submit_resolvent(fid[key], p, vec, 1e-10 * np.linalg.norm(vec))

    if together:
        rids = [submit(r) for r in reqs]
        out = [srv.wait(r) for r in rids]
    else:
        out = [srv.wait(submit(r)) for r in reqs]
    srv.close()
    return out


def test_serving_mix_full_sizes(mix):
    fams, F, J, reqs = mix
    print(f"[c5] {len(reqs)} requests in the mix", flush=True)
    in_mix = serve(fams, F, reqs, 4, True)
    print("[c5] the same requests one at a time", flush=True)
    solo = serve(fams, F, reqs, 4, False)
    oracle_problems = {}
    for req, (x, info), (xa, _) in zip(reqs, in_mix, solo):
        kind, key, p, vec = req
        assert info.kind == kind
        assert np.array_equal(x, xa), (kind, key, p)
        if kind == "expv":
            wo, _, _ = expv_jumps(*J[:4], p, vec, tol=1e-8)
            assert np.abs(x - wo).sum() <= 1e-7, np.abs(x - wo).sum()
            continue
        w, box, Q, _ = fams[key]
        assert info.solve.converged
        r = vec - Q.apply_shifted(p, x)
        assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(vec)
        abs_A = p + 2 * Q.exit_rates() - Q.leak_rates()
        bound = (np.abs(r).sum() + 16 * EPS * (abs_A * np.abs(x)).sum()) / p
        K = box.n_sectors()
        begins = [box.sector_begin(k) for k in range(K)] + [box.size()]
        Bk = np.array([vec[begins[k]:begins[k + 1]].sum() for k in range(K)])
        Xk = np.array([x[begins[k]:begins[k + 1]].sum() for k in range(K)])
        X = lumped_sector_sums(p, 0.05, 0.02, key[1], Bk)
        assert np.abs(Xk - X).sum() <= bound, (key, np.abs(Xk - X).sum(), bound)
        if box.size() <= 58905:
            if key not in oracle_problems:
                oracle_problems[key] = Problem(w.model, w.max_totals)
            P = oracle_problems[key]
            x_ref, oinfo = P.gmres(p, vec, mg=OracleMG(P, p), tol=1e-10 * np.linalg.norm(vec),
                                   restart=50)
            assert oinfo["converged"]
            r_ref = vec - P.apply_shifted(p, x_ref)
            obound = bound + (np.abs(r_ref).sum() + 16 * EPS * (abs_A * np.abs(x_ref)).sum()) / p
            assert np.abs(x - x_ref).sum() <= obound, (key, p)
    run = {}
    for req, (_, info) in zip(reqs, in_mix):
        run.setdefault(req[0], []).append(info.run_s)
    print("[c5] mean run ms: " + ", ".join(f"{k} {1e3 * np.mean(v):.1f}" for k, v in run.items()))
