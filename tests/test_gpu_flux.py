"""Flux-block Schur complement (kernel family 8): the flux unknowns are
eliminated with A_qq^-1 = C^-1 (x) M^-1 (per-cell coefficient matrix C, fixed
scalar mass matrix M) and the reduced system is condensed by an elimination
kernel.  Checked against the FP64 oracle (relative 1e-12 for S, g and the
skeleton system, or for elasticity ten times the oracle's own variation under
a reordering of the interior unknowns when that is larger, as in the dense
elasticity parity test; recovery as in the dense parity tests), against the
dense getrf-order families on the same cells, and for its failure report."""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import _device_cell_systems, check_local_factors, rel

pytestmark = pytest.mark.gpu

RTOL = 1e-12
FLUX = 8
FUSED = 9
FLUX_FAMS = (FLUX, FUSED)


def _darcy(ctx, k, n, seed, family=None, kappa=None):
    import gridaphybrid as gh
    model = gh.CartesianDiscreteModel((0.0, 1.0) * 3, (n,) * 3)
    X = gh.HDGSpaces.darcy_baseline(model, k)
    kappa = O.lognormal_kappa(n ** 3, seed=seed) if kappa is None else kappa
    wf = gh.DarcyHDG(tau=1.0, kappa=kappa, source=O.MMS_SINPROD)
    return gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx, family=family), kappa


def _elasticity(ctx, k, n, seed, family=None):
    import gridaphybrid as gh
    model = gh.CartesianDiscreteModel((0.0, 1.0) * 3, (n,) * 3)
    X = gh.HDGSpaces.elasticity(model, k)
    nu = np.random.default_rng(seed).uniform(0.2, 0.4, n ** 3)
    wf = gh.ElasticityHDG(tau=float(n), nu=nu, E=1.0, source=gh._lib.SOURCE_ELAS_SIN)
    return gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx, family=family), nu


def _oracle_self_variation(op, ref):
    """How far the oracle's own S and g (and the assembled matrix and rhs,
    which add the Dirichlet lifting) move when only the order of the interior
    unknowns changes (same LU, another pivot sequence): the rounding floor of
    the problem, which is above 1e-12 for the worse-conditioned elasticity
    cells."""
    ni, nb = op.info.ni, op.info.nb
    perm = np.r_[np.random.default_rng(1).permutation(ni), ni:ni + nb]
    S2, g2, _ = O.condense(ref.A[:, perm][:, :, perm], ref.b[:, perm], ni, nb)
    _, _, nz2, rhs2 = O.assemble(S2, g2, ref.cell_dofs, ref.nfree, ref.lam_dir)
    return (max(rel(S2, ref.S), rel(nz2, ref.nzval)), max(rel(g2, ref.g), rel(rhs2, ref.rhs)))


def _check_vs_oracle(op, ref, tol_fv, adaptive=False):
    S, g, cd = op.cell_blocks()
    assert np.array_equal(cd, ref.cell_dofs)
    errs = {"S": rel(S, ref.S), "g": rel(g, ref.g)}
    K, rhs = op.skeleton_system()
    assert np.array_equal(K.indptr, ref.colptr) and np.array_equal(K.indices, ref.rowval)
    errs["nzval"], errs["rhs"] = rel(K.data, ref.nzval), rel(rhs, ref.rhs)
    fv = op.recover(ref.lam).cpu().numpy()
    errs["fv"] = rel(fv, ref.free_values)
    tol_S = tol_g = RTOL
    if adaptive:
        vS, vg = _oracle_self_variation(op, ref)
        tol_S, tol_g = max(RTOL, 10 * vS), max(RTOL, 10 * vg)
    print("errors vs FP64 oracle:", {k: f"{v:.1e}" for k, v in errs.items()},
          f"tolerances S {tol_S:.1e} g {tol_g:.1e}", flush=True)
    for key, tol in (("S", tol_S), ("g", tol_g), ("nzval", tol_S), ("rhs", tol_g)):
        assert errs[key] < tol, (key, errs[key], tol)
    assert errs["fv"] < tol_fv
    return S, g, fv


@pytest.mark.parametrize("k,n", [(2, 3), (2, 4)])
def test_flux_darcy_vs_oracle_and_dense(ctx, k, n):
    print(f"flux darcy 3D k={k} n={n}", flush=True)
    op, kappa = _darcy(ctx, k, n, seed=40 + n)
    rows, fam = op.factor_rows()
    assert fam in FLUX_FAMS
    assert rows == (op.info.ni if fam == FUSED else op.info.ni - op.info.n_flux)
    ref = O.hdg_solve(O.Problem.darcy(3, k, mms=O.MMS_SINPROD), (n,) * 3, coef=kappa)
    S, g, fv = _check_vs_oracle(op, ref, 1e-11)
    A, b = _device_cell_systems(op, 0, min(op.info.ncells, 16))
    U, Y = op.local_factors()
    check_local_factors(op, A, b, U, Y, min(op.info.ncells, 16))
    # the dense getrf-order elimination on the same cells
    dense, _ = _darcy(ctx, k, n, seed=40 + n, family=3)
    assert dense.factor_rows() == (dense.info.ni, 3)
    S2, g2, _ = dense.cell_blocks()
    assert rel(S, S2) < RTOL and rel(g, g2) < RTOL
    fv2 = dense.recover(ref.lam).cpu().numpy()
    assert rel(fv, fv2) < 1e-11
    dense.close()
    op.close()


@pytest.mark.parametrize("k,n", [(1, 2), (2, 2)])
def test_flux_elasticity_vs_oracle(ctx, k, n):
    """Stress block C (x) M with the per-cell 6x6 compliance C (nu per cell)."""
    print(f"flux elasticity k={k} n={n}", flush=True)
    op, nu = _elasticity(ctx, k, n, seed=50 + k)
    assert op.factor_rows()[1] == FLUX
    prob = O.Problem.elasticity(3, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN)
    ref = O.hdg_solve(prob, (n,) * 3, coef=nu)
    _check_vs_oracle(op, ref, 1e-11, adaptive=True)
    A, b = _device_cell_systems(op, 0, op.info.ncells)
    U, Y = op.local_factors()
    check_local_factors(op, A, b, U, Y, op.info.ncells)
    op.close()


def test_flux_family_selection(ctx):
    """Explicit family control: flux_schur where the class has the flux block,
    refused where the family cannot run; -1 restores the default."""
    import gridaphybrid as gh
    op, _ = _darcy(ctx, 2, 2, seed=3, family=FLUX)
    assert op.factor_rows()[1] == FLUX
    op.set_family(3)
    op.assemble()
    assert op.factor_rows()[1] == 3
    op.set_family(None)
    op.assemble()
    assert op.factor_rows()[1] == FLUX    # the default for the class
    op.close()
    model = gh.CartesianDiscreteModel((0.0, 1.0) * 2, (4, 4))
    X = gh.HDGSpaces.darcy_baseline(model, 2)
    small = gh.HybridAffineFEOperator(gh.DarcyHDG(), X, X, [1, 2], [3], ctx=ctx)
    assert small.factor_rows()[1] == 0   # the wavefront kernel keeps small classes
    with pytest.raises(gh.HDGError):
        small.set_family(3)              # a 32x16 grid cannot hold a 39-row system
    small.close()


def test_flux_singular_coefficient_reported(ctx):
    """kappa = inf makes C = |det J| / kappa = 0 in one cell: that cell is
    reported singular (the flux block A_qq is zero, so A_ii is singular)."""
    import gridaphybrid as gh
    kappa = np.ones(8)
    kappa[5] = np.inf
    with pytest.raises(gh.SingularError):
        _darcy(ctx, 2, 2, seed=0, kappa=kappa)


@pytest.mark.parametrize("k,n", [(2, 4)])
def test_fused_matches_flux_schur(ctx, k, n):
    """Family 9 (fused) against family 8 (flux_schur + reduced elimination) on
    the same cells: S, g to 1e-12 and the recovered values to 1e-11."""
    print(f"fused vs flux_schur darcy 3D k={k} n={n}", flush=True)
    a, kappa = _darcy(ctx, k, n, seed=70, family=FUSED)
    b, _ = _darcy(ctx, k, n, seed=70, family=FLUX)
    assert a.factor_rows()[1] == FUSED and b.factor_rows()[1] == FLUX
    Sa, ga, _ = a.cell_blocks()
    Sb, gb, _ = b.cell_blocks()
    assert rel(Sa, Sb) < RTOL and rel(ga, gb) < RTOL
    lam = a.interpolate_trace()
    assert rel(a.recover(lam).cpu().numpy(), b.recover(lam).cpu().numpy()) < 1e-11
    a.close()
    b.close()


@pytest.mark.parametrize("phys,k,n", [(0, 2, 3), (1, 1, 2)])
def test_fused_family_vs_oracle(ctx, phys, k, n):
    """Family 9 (the fused flux-block kernel, an autotuner candidate) against
    the oracle, with its W rows checked as A_ii^-1 [A_ib b_i]."""
    print(f"fused family vs oracle phys={phys} k={k} n={n}", flush=True)
    if phys == 0:
        op, kappa = _darcy(ctx, k, n, seed=80, family=FUSED)
        ref = O.hdg_solve(O.Problem.darcy(3, k, mms=O.MMS_SINPROD), (n,) * 3, coef=kappa)
        _check_vs_oracle(op, ref, 1e-11)
    else:
        op, nu = _elasticity(ctx, k, n, seed=81, family=FUSED)
        prob = O.Problem.elasticity(3, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN)
        ref = O.hdg_solve(prob, (n,) * 3, coef=nu)
        _check_vs_oracle(op, ref, 1e-11, adaptive=True)
    assert op.factor_rows()[1] == FUSED
    A, b = _device_cell_systems(op, 0, min(op.info.ncells, 16))
    U, Y = op.local_factors()
    assert U is None
    check_local_factors(op, A, b, U, Y, min(op.info.ncells, 16))
    op.close()
