"""GPU tests of request serving (mk_server_*).

Worker path (batching off): every request's result is bitwise the same as the same request
solved alone by the library (each request computes on its own stream with fixed-order
reductions, and a pooled hierarchy re-targeted to a shift holds exactly the terminal inverses a
new hierarchy for that shift would); admission by HBM budget serialises requests without losing
any.

Continuous batching (the default for small families): requests join a resident batch engine at
restart boundaries while others run. Every result is bitwise the same as the same request served
alone (a lane's arithmetic does not depend on the other lanes), and within the L1 bound
||x - x'||_1 <= (||b - A x||_1 + ||b - A x'||_1) / s of the library's single-system solve.
Requests of a smaller cap on the same grid (a leading block of whole sectors) share the engine
and match their own box's solve and the lumped closed form of their sector sums.
"""
import time

import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.fsp import box_generator, toggle_switch
from markovkit_amd.workloads import birth_death_1d, lumped_sector_sums

pytestmark = pytest.mark.gpu


def family(voxels, cap, smoother, pool=2):
    w = birth_death_1d(voxels, cap, 0.5, "random")
    box = mk.PopulationBox(w.n_voxels, w.max_totals)
    Q = mk.Generator(box, w.model)
    mgs = [mk.SectorMultigrid(w.model.domain, box, 1.0,
                              mk.MultigridOptions(precision=mk.FP32, smoother=smoother))
           for _ in range(pool)]
    return w, box, Q, mgs


@pytest.fixture(scope="module", params=[mk.JACOBI, mk.RED_BLACK], ids=["jacobi", "red_black"])
def setup(request):
    fams = [family(16, 3, request.param), family(32, 3, request.param)]
    J = box_generator(toggle_switch(), [60, 60], absorbing=True)
    F = mk.Generator.from_jumps(*J)
    return fams, F, request.param


def requests(fams, F):
    rng = np.random.default_rng(7)
    out = []
    for k in range(10):
        f = k % 2
        n = fams[f][2].rows()
        b = rng.random(n)
        b /= b.sum()
        out.append(("resolvent", f, float(rng.uniform(0.1, 10.0)), b))
        if k % 3 == 0:
            v = np.zeros(F.rows())
            v[int(rng.integers(0, F.rows()))] = 1.0
            out.append(("expv", None, float(rng.uniform(0.5, 3.0)), v))
    return out


def alone(fams, F, req, smoother):
    """the library's own single-request answer"""
    kind, f, p, vec = req
    if kind == "expv":
        w, _ = mk.expv(F, p, vec, tolerance=1e-8)
        return w
    w_, box, Q, _ = fams[f]
    M = mk.SectorMultigrid(w_.model.domain, box, p,
                           mk.MultigridOptions(precision=mk.FP32, smoother=smoother))
    x = np.zeros(Q.rows())
    res = mk.gmres(mk.ShiftedGenerator(Q, p), M, vec, x, tolerance=1e-10 * np.linalg.norm(vec),
                   restart=50)
    assert res.converged
    return x


def serve(fams, F, reqs, workers, budget, batch_rows, together=True, stagger=0.0):
    srv = mk.Server(workers=workers, hbm_budget=budget, batch_rows=batch_rows)
    ids = [srv.add_family(Q, mgs) for (_, _, Q, mgs) in fams]
    assert ids == list(range(len(fams)))

    def submit(req):
        kind, f, p, vec = req
        if kind == "expv":
            return srv.submit_expv(F, p, vec, tolerance=1e-8)
        return srv.submit_resolvent(f, p, vec, 1e-10 * np.linalg.norm(vec))

    if together:
        rids = []
        for q in reqs:
            rids.append(submit(q))
            if stagger:
                time.sleep(stagger)
        out = [srv.wait(r) for r in rids]
    else:
        out = [srv.wait(submit(q)) for q in reqs]
    srv.close()
    return out


@pytest.mark.parametrize("workers,budget", [(3, 64 << 30), (2, 1)], ids=["concurrent", "serialised"])
def test_worker_path_alone_and_in_the_mix(setup, workers, budget):
    fams, F, smoother = setup
    reqs = requests(fams, F)
    served = serve(fams, F, reqs, workers, budget, batch_rows=0)
    for req, (x, info) in zip(reqs, served):
        ref = alone(fams, F, req, smoother)
        assert info.kind == req[0]
        if info.kind == "resolvent":
            assert info.solve.converged
        assert np.array_equal(x, ref), (req[0], req[2], np.abs(x - ref).max())
    run = [info.run_s for _, info in served]
    print(f"[serving workers={workers}] {len(reqs)} requests, mean run {np.mean(run) * 1e3:.2f} ms")


def test_continuous_batching_alone_and_in_the_mix(setup):
    """Requests arrive 2 ms apart while earlier ones run, so later ones join the engines at
    restart boundaries; each result equals the same request served alone bit for bit."""
    fams, F, smoother = setup
    reqs = requests(fams, F)
    mixed = serve(fams, F, reqs, 2, 64 << 30, batch_rows=None, stagger=0.002)
    solo = serve(fams, F, reqs, 2, 64 << 30, batch_rows=None, together=False)
    for req, (x, info), (xa, ia) in zip(reqs, mixed, solo):
        assert info.kind == req[0]
        assert np.array_equal(x, xa), (req[0], req[2])
        if info.kind != "resolvent":
            continue
        assert info.solve.converged and info.solve.iterations == ia.solve.iterations
        _, _, Q, _ = fams[req[1]]
        p, b = req[2], req[3]
        r = b - Q.apply_shifted(p, x)
        assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b)
        ref = alone(fams, F, req, smoother)
        bound = (np.abs(r).sum() + np.abs(b - Q.apply_shifted(p, ref)).sum()) / p
        assert np.abs(x - ref).sum() <= bound
    run = [info.run_s for _, info in mixed]
    print(f"[continuous batching] {len(reqs)} requests, mean run {np.mean(run) * 1e3:.2f} ms")


def test_smaller_caps_share_a_family(setup):
    """The cap-3 family on 32 voxels serves requests of caps 0 to 3 (leading sectors): each
    equals its own box's solve within the L1 bound and its sector sums the lumped closed form."""
    fams, _, smoother = setup
    w, box, Q, mgs = fams[1]
    srv = mk.Server(workers=1)
    fam = srv.add_family(Q, mgs)
    rng = np.random.default_rng(3)
    reqs = []
    for cap in (1, 3, 2, 0, 3, 1):
        rows = mk.PopulationBox(32, [cap]).size()
        b = rng.random(rows)
        b /= b.sum()
        s = float(rng.uniform(0.1, 10.0))
        reqs.append((cap, s, b, srv.submit_resolvent(fam, s, b, 1e-10 * np.linalg.norm(b))))
    for cap, s, b, rid in reqs:
        x, info = srv.wait(rid)
        assert info.solve.converged and len(x) == len(b)
        wc = birth_death_1d(32, cap, s, "random")
        bc = mk.PopulationBox(32, [cap])
        Qc = mk.Generator(bc, wc.model)
        r = b - Qc.apply_shifted(s, x)
        assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b)
        Mc = mk.SectorMultigrid(wc.model.domain, bc, s,
                                mk.MultigridOptions(precision=mk.FP32, smoother=smoother))
        x1 = np.zeros(len(b))
        assert mk.gmres(mk.ShiftedGenerator(Qc, s), Mc, b, x1, tolerance=1e-10 * np.linalg.norm(b)).converged
        bound = (np.abs(r).sum() + np.abs(b - Qc.apply_shifted(s, x1)).sum()) / s
        assert np.abs(x - x1).sum() <= bound
        begins = [bc.sector_begin(k) for k in range(cap + 1)] + [len(b)]
        Bk = np.array([b[begins[k]:begins[k + 1]].sum() for k in range(cap + 1)])
        Xk = np.array([x[begins[k]:begins[k + 1]].sum() for k in range(cap + 1)])
        absA = s + 2 * Qc.exit_rates() - Qc.leak_rates()
        bound = (np.abs(r).sum() + 16 * np.finfo(float).eps * (absA * np.abs(x)).sum()) / s
        assert np.abs(Xk - lumped_sector_sums(s, 0.05, 0.02, cap, Bk)).sum() <= bound
    srv.close()


def test_unknown_family_is_rejected(setup):
    fams, F, _ = setup
    srv = mk.Server(workers=1)
    srv.add_family(fams[0][2], fams[0][3])
    with pytest.raises(Exception):
        srv.submit_resolvent(3, 1.0, np.ones(fams[0][2].rows()), 1e-10)
    srv.close()
