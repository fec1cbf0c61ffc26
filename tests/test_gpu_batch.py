"""Uneven batches (hdg_batch_*): many independent meshes of different local
system sizes condensed and assembled by one graph replay give the same results
as each mesh condensed on its own, and as the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (physics, D, k, n, seed): two members share each of the first classes so the
# per-class slicing is exercised; the last two use the blocked kernel
MEMBERS = [
    (0, 2, 1, 5, 1), (0, 2, 1, 3, 2), (0, 2, 2, 4, 3), (0, 2, 3, 3, 4),
    (0, 3, 1, 2, 5), (0, 3, 1, 3, 6), (0, 3, 2, 2, 7),
    (1, 3, 1, 2, 8), (1, 3, 2, 1, 9), (1, 3, 1, 1, 10),
    (0, 3, 3, 1, 11), (1, 3, 3, 1, 12),
]


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape
    if b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _build(ctx, spec, assemble):
    import gridaphybrid as gh
    phys, D, k, n, seed = spec
    model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
    nc = n ** D
    if phys == 0:
        X = gh.HDGSpaces.darcy_baseline(model, k)
        wf = gh.DarcyHDG(tau=1.0, kappa=O.lognormal_kappa(nc, seed=seed),
                         source=O.MMS_SINPROD)
    else:
        X = gh.HDGSpaces.elasticity(model, k)
        nu = np.random.default_rng(seed).uniform(0.2, 0.4, nc)
        wf = gh.ElasticityHDG(tau=float(n), nu=nu, E=1.0, source=gh._lib.SOURCE_ELAS_SIN)
    return gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx, assemble=assemble)


def _state(op):
    S, g, cd = op.cell_blocks()
    U, Y = op.local_factors()
    K, rhs = op.skeleton_system()
    lam = op.interpolate_trace()
    fv = op.recover(lam).cpu().numpy()
    return dict(S=S, g=g, cd=cd, U=U, Y=Y, nz=K.data, rv=K.indices, cp=K.indptr, rhs=rhs, fv=fv,
                lam=lam)


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_matches_single_meshes(ctx, autotune):
    import gridaphybrid as gh
    print(f"batch autotune={autotune}", flush=True)
    alone = []
    for spec in MEMBERS:
        op = _build(ctx, spec, True)
        alone.append(_state(op))
        op.close()
    ops = [_build(ctx, spec, False) for spec in MEMBERS]
    batch = gh.HybridOperatorBatch(ops, autotune=autotune)
    cls = batch.classes()
    assert len(cls) == 9
    assert sum(c["nmesh"] for c in cls) == len(MEMBERS)
    assert sum(c["ncells"] for c in cls) == sum(s[3] ** s[1] for s in MEMBERS)
    for rep in range(3):   # replays of the same graph give the same values
        batch.assemble()
        for i, (op, ref) in enumerate(zip(ops, alone)):
            got = _state(op)
            for key in ("cd", "rv", "cp"):
                assert np.array_equal(got[key], ref[key]), (i, key)
            if not autotune:
                # the default kernel family for every class: bit for bit
                for key in ("S", "g", "U", "Y", "nz", "rhs", "fv"):
                    assert np.array_equal(got[key], ref[key]), (rep, i, key)
            else:
                # another kernel family may round differently; k=3 elasticity
                # has cond(A_ii) ~ 3e8
                tol = 1e-9 if MEMBERS[i][0] == 1 and MEMBERS[i][2] == 3 else 1e-11
                for key in ("S", "g", "nz", "rhs", "fv"):
                    assert rel(got[key], ref[key]) < tol, (rep, i, key)
    # a batched mesh cannot be destroyed before its batch
    from gridaphybrid import _lib
    assert _lib.lib().hdg_mesh_destroy(ops[0].mesh) == _lib.HDG_ERR_STATE
    batch.close()
    for op in ops:
        op.close()


def test_batch_recover_and_oracle(ctx):
    """hdg_batch_recover for every member at once, checked against the oracle."""
    import torch
    import gridaphybrid as gh
    print("batch recover vs oracle", flush=True)
    specs = [MEMBERS[0], MEMBERS[2], MEMBERS[4], MEMBERS[7]]
    ops = [_build(ctx, spec, False) for spec in specs]
    batch = gh.HybridOperatorBatch(ops, autotune=True)
    batch.assemble()
    refs, lams, outs = [], [], []
    for spec, op in zip(specs, ops):
        phys, D, k, n, seed = spec
        nc = n ** D
        if phys == 0:
            prob = O.Problem.darcy(D, k, mms=O.MMS_SINPROD)
            ref = O.hdg_solve(prob, (n,) * D, coef=O.lognormal_kappa(nc, seed=seed))
        else:
            nu = np.random.default_rng(seed).uniform(0.2, 0.4, nc)
            prob = O.Problem.elasticity(D, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN)
            ref = O.hdg_solve(prob, (n,) * D, coef=nu)
        refs.append(ref)
        lams.append(torch.tensor(ref.lam, dtype=torch.float64, device="cuda"))
        outs.append(torch.empty(op.info.nfree_total, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    batch.recover(lams, outs)
    ctx.synchronize()
    for op, ref, out in zip(ops, refs, outs):
        S, g, cd = op.cell_blocks()
        assert np.array_equal(cd, ref.cell_dofs)
        assert rel(S, ref.S) < 1e-12 and rel(g, ref.g) < 1e-12
        assert rel(out.cpu().numpy(), ref.free_values) < 1e-11
    batch.close()
    for op in ops:
        op.close()
