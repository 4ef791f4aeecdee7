"""GPU tests of request serving (mk_server_*): every request's result is the same when it runs
alone and when it runs in a mix of concurrent requests on several workers (bitwise: each
request computes on its own stream with fixed-order reductions, and a pooled hierarchy
re-targeted to a shift holds exactly the terminal inverses a new hierarchy for that shift
would); admission by HBM budget serialises requests without losing any.
"""
import numpy as np
import pytest

import markovkit_amd as mk
from markovkit_amd.fsp import box_generator, toggle_switch
from markovkit_amd.workloads import birth_death_1d

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


@pytest.mark.parametrize("workers,budget", [(3, 64 << 30), (2, 1)], ids=["concurrent", "serialised"])
def test_results_alone_and_in_the_mix(setup, workers, budget):
    fams, F, smoother = setup
    reqs = requests(fams, F)
    srv = mk.Server(workers=workers, hbm_budget=budget)
    ids = [srv.add_family(Q, mgs) for (_, _, Q, mgs) in fams]
    assert ids == [0, 1]
    rids = []
    for kind, f, p, vec in reqs:
        if kind == "expv":
            rids.append(srv.submit_expv(F, p, vec, tolerance=1e-8))
        else:
            rids.append(srv.submit_resolvent(f, p, vec, 1e-10 * np.linalg.norm(vec)))
    served = [srv.wait(r) for r in rids]
    srv.close()
    for req, (x, info) in zip(reqs, served):
        ref = alone(fams, F, req, smoother)
        assert info.kind == req[0]
        if info.kind == "resolvent":
            assert info.solve.converged
        assert np.array_equal(x, ref), (req[0], req[2], np.abs(x - ref).max())
    run = [info.run_s for _, info in served]
    print(f"[serving workers={workers}] {len(reqs)} requests, mean run {np.mean(run) * 1e3:.2f} ms")


def test_unknown_family_is_rejected(setup):
    fams, F, _ = setup
    srv = mk.Server(workers=1)
    srv.add_family(fams[0][2], fams[0][3])
    with pytest.raises(Exception):
        srv.submit_resolvent(3, 1.0, np.ones(fams[0][2].rows()), 1e-10)
    srv.close()
