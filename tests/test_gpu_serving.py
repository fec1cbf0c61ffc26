"""Serving loop: requests admitted by HBM footprint, condensed together in
rounds of one batch graph each on several streams; every result is the same
alone and in the mix (bit for bit: the kernels are cell-independent), and
matches the FP64 oracle at 1e-10 relative."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SPECS = [  # (physics, D, k, n)
    (0, 2, 1, 6), (0, 2, 2, 8), (0, 2, 3, 5), (0, 3, 1, 3), (0, 3, 2, 3),
    (1, 3, 1, 2), (1, 3, 2, 2), (0, 2, 2, 7), (0, 2, 1, 9), (1, 3, 1, 3),
]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _oracle(req):
    rng = np.random.default_rng(req.seed)
    nc = req.n ** req.dim
    if req.physics == 0:
        prob = O.Problem.darcy(req.dim, req.order, mms=O.MMS_SINPROD)
        return O.hdg_solve(prob, (req.n,) * req.dim, coef=np.exp(rng.standard_normal(nc)))
    prob = O.Problem.elasticity(3, req.order, tau=float(req.n), nu=0.3, mms=O.ELAS_MMS_SIN)
    return O.hdg_solve(prob, (req.n,) * 3, coef=rng.uniform(0.2, 0.4, nc))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_results_alone_equal_results_in_the_mix(ctx, precision):
    import gridaphybrid as gh
    from gridaphybrid import serving as sv
    print(f"serving mix vs alone ({precision})", flush=True)
    prec = gh.MIXED if precision == "mixed" else gh.FP64
    # two waves: the second arrives while the first is being served
    reqs = [sv.SolveRequest(i, 0.0 if i < 5 else 1e-4, p, D, k, n, 100 + i)
            for i, (p, D, k, n) in enumerate(SPECS)]
    srv = sv.Server(gh, hbm_budget=16 << 30, precision=prec, streams=2, skeleton="spsolve")
    srv.run(reqs)
    assert all(r.done is not None for r in reqs)
    assert max(srv.rounds) > 1, srv.rounds          # requests were condensed together
    for r in reqs:
        alone = sv.SolveRequest(r.rid, 0.0, r.physics, r.dim, r.order, r.n, r.seed)
        srv.run([alone])
        assert srv.rounds == [1]
        assert np.array_equal(alone.result, r.result), r
    for r in reqs[:6]:
        ref = _oracle(r)
        assert rel(r.result, ref.free_values) < 1e-10, (r, rel(r.result, ref.free_values))
    srv.close()


def test_admission_respects_the_budget(ctx):
    """A budget that holds about one request at a time serialises the rounds;
    a request larger than the whole budget is rejected."""
    import gridaphybrid as gh
    from gridaphybrid import serving as sv
    print("serving admission", flush=True)
    reqs = [sv.SolveRequest(i, 0.0, 0, 2, 2, 16, 7 + i) for i in range(4)]
    fp = sv.footprint(reqs[0], gh.MIXED)
    big = sv.SolveRequest(9, 0.0, 0, 3, 2, 24, 3)
    srv = sv.Server(gh, hbm_budget=int(fp * 1.5), precision=gh.MIXED, streams=2)
    srv.run(reqs + [big])
    assert all(r.done is not None for r in reqs)
    assert big.done is None and srv.rejected == [big]
    assert max(srv.rounds) == 1
    srv.close()
