"""Mixed precision (HDG_MIXED): FP32 Gauss-Jordan inverse of A_ii, FP64
refinement of W = A_ii^-1 [A_ib b_i], FP64 S, g and recovery.  Every result is
checked against the FP64 oracle at relative 1e-10 (max-norm error over the
max-norm of the oracle value)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape
    if b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _darcy(ctx, D, k, n, seed, precision):
    import gridaphybrid as gh
    model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
    X = gh.HDGSpaces.darcy_baseline(model, k)
    kappa = O.lognormal_kappa(n ** D, seed=seed)
    wf = gh.DarcyHDG(tau=1.0, kappa=kappa, source=O.MMS_SINPROD)
    op = gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx, precision=precision)
    ref = O.hdg_solve(O.Problem.darcy(D, k, mms=O.MMS_SINPROD), (n,) * D, coef=kappa)
    return op, ref


def _elasticity(ctx, k, n, seed, precision):
    import gridaphybrid as gh
    model = gh.CartesianDiscreteModel((0.0, 1.0) * 3, (n,) * 3)
    X = gh.HDGSpaces.elasticity(model, k)
    nu = np.random.default_rng(seed).uniform(0.2, 0.4, n ** 3)
    wf = gh.ElasticityHDG(tau=float(n), nu=nu, E=1.0, source=gh._lib.SOURCE_ELAS_SIN)
    op = gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx, precision=precision)
    prob = O.Problem.elasticity(3, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN)
    ref = O.hdg_solve(prob, (n,) * 3, coef=nu)
    return op, ref


def _check(op, ref, mixed_expected=None):
    st = op.refinement_stats()
    if mixed_expected is None:   # the mixed path serves 64 <= ni <= 128
        mixed_expected = 64 <= op.info.ni <= 128
    assert st["mixed"] == mixed_expected, st
    assert st["unconverged"] == 0 and not st["fp64_fallback"], st
    S, g, cd = op.cell_blocks()
    assert np.array_equal(cd, ref.cell_dofs)
    errs = {"S": rel(S, ref.S), "g": rel(g, ref.g)}
    K, rhs = op.skeleton_system()
    assert np.array_equal(K.indptr, ref.colptr) and np.array_equal(K.indices, ref.rowval)
    errs["nzval"] = rel(K.data, ref.nzval)
    errs["rhs"] = rel(rhs, ref.rhs)
    fv = op.recover(ref.lam).cpu().numpy()
    na, nc = op.info.n_flux, op.info.ncells
    errs["q"] = rel(fv[:nc * na], ref.free_values[:nc * na])
    errs["u"] = rel(fv[nc * na:nc * op.info.ni], ref.free_values[nc * na:nc * op.info.ni])
    print("errors vs FP64 oracle:", {k: f"{v:.1e}" for k, v in errs.items()}, st, flush=True)
    for key, e in errs.items():
        assert e < RTOL, (key, e)
    return st


@pytest.mark.parametrize("D,k,n", [(2, 1, 8), (2, 2, 16), (2, 3, 6), (3, 1, 4), (3, 2, 4),
                                   (3, 2, 6)])
def test_mixed_darcy(ctx, D, k, n):
    from gridaphybrid import MIXED
    print(f"mixed darcy D={D} k={k} n={n}", flush=True)
    op, ref = _darcy(ctx, D, k, n, seed=D * 10 + k, precision=MIXED)
    st = _check(op, ref)
    if not st["mixed"]:
        op.close()
        return
    assert 1 <= st["max_sweeps"] <= 10
    U, W = op.local_factors()
    assert U is None
    # W solves A_ii W = [A_ib b_i] on the oracle's cell systems
    ni = op.info.ni
    B = np.concatenate([ref.A[:, :ni, ni:], ref.b[:, :ni, None]], axis=2)
    resid = np.einsum("cij,cjk->cik", ref.A[:, :ni, :ni], W) - B
    assert np.max(np.abs(resid)) <= 1e-10 * np.max(np.abs(B))
    op.close()


@pytest.mark.parametrize("k,n", [(1, 2), (2, 2)])
def test_mixed_elasticity(ctx, k, n):
    from gridaphybrid import MIXED
    print(f"mixed elasticity k={k} n={n}", flush=True)
    op, ref = _elasticity(ctx, k, n, seed=40 + k, precision=MIXED)
    _check(op, ref)
    op.close()


def test_mixed_falls_back_to_fp64_for_large_interiors(ctx):
    """ni = 256 (3D Darcy k=3) is beyond the FP32 inverse: the FP64 path runs."""
    from gridaphybrid import MIXED
    print("mixed darcy 3D k=3 (FP64 path)", flush=True)
    op, ref = _darcy(ctx, 3, 3, 2, seed=33, precision=MIXED)
    _check(op, ref, mixed_expected=False)
    op.close()


def test_mixed_batch_matches_single(ctx):
    """The batch graph in mixed precision gives the single-mesh results."""
    import gridaphybrid as gh
    from gridaphybrid import MIXED
    print("mixed batch", flush=True)
    specs = [(2, 2, 8, 1), (3, 2, 3, 2), (3, 1, 3, 3), (2, 2, 5, 4)]
    alone = []
    for D, k, n, seed in specs:
        op, ref = _darcy(ctx, D, k, n, seed, MIXED)
        S, g, _ = op.cell_blocks()
        alone.append((S, g, ref))
        op.close()
    ops = []
    for D, k, n, seed in specs:
        model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
        X = gh.HDGSpaces.darcy_baseline(model, k)
        wf = gh.DarcyHDG(tau=1.0, kappa=O.lognormal_kappa(n ** D, seed=seed),
                         source=O.MMS_SINPROD)
        ops.append(gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx, assemble=False))
    batch = gh.HybridOperatorBatch(ops, autotune=False)
    batch.assemble(MIXED)
    for op, (S0, g0, ref) in zip(ops, alone):
        S, g, _ = op.cell_blocks()
        assert np.array_equal(S, S0) and np.array_equal(g, g0)
        assert op.refinement_stats()["mixed"] == (op.info.ni >= 64)
        fv = op.recover(ref.lam).cpu().numpy()
        assert rel(fv, ref.free_values) < RTOL
    batch.close()
    for op in ops:
        op.close()
