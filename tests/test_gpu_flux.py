"""Flux-block Schur complement (kernel families 8 and 10): the flux unknowns are
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
REF = 10
FLUX_FAMS = (FLUX, REF)


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
    assert rows == op.info.ni - op.info.n_flux
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
    assert op.factor_rows()[1] == REF     # the default for the Darcy class
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
def test_ref_matches_flux_schur(ctx, k, n):
    """Family 10 (the reduced system from the reference decomposition of the
    flux Schur complement) against family 8 (two MFMA products per cell) on the
    same cells: S, g to 1e-12, the recovered values to 1e-11."""
    print(f"flux_ref vs flux_schur darcy 3D k={k} n={n}", flush=True)
    a, kappa = _darcy(ctx, k, n, seed=71, family=REF)
    b, _ = _darcy(ctx, k, n, seed=71, family=FLUX)
    assert a.factor_rows() == (a.info.ni - a.info.n_flux, REF)
    assert b.factor_rows()[1] == FLUX
    Sa, ga, _ = a.cell_blocks()
    Sb, gb, _ = b.cell_blocks()
    assert rel(Sa, Sb) < RTOL and rel(ga, gb) < RTOL
    lam = a.interpolate_trace()
    assert rel(a.recover(lam).cpu().numpy(), b.recover(lam).cpu().numpy()) < 1e-11
    ref = O.hdg_solve(O.Problem.darcy(3, k, mms=O.MMS_SINPROD), (n,) * 3, coef=kappa)
    _check_vs_oracle(a, ref, 1e-11)
    A, bb = _device_cell_systems(a, 0, 16)
    U, Y = a.local_factors()
    check_local_factors(a, A, bb, U, Y, 16)
    a.close()
    b.close()


@pytest.mark.parametrize("frac", [0.0, 0.7])
def test_ref_general_affine_cells(ctx, frac):
    """The reference decomposition on sheared cells (non-diagonal Jacobians),
    some re-parametrised by rotations of the reference cell (oriented facets):
    family 10 is the default and matches the oracle."""
    import gridaphybrid as gh
    import meshgen
    print(f"flux_ref on sheared cells, rotated fraction {frac}", flush=True)
    n, k = 3, 2
    x0, J, cf, nf, perm, rot = meshgen.sheared_rotated_mesh(3, n, seed=12, frac=frac)
    fperm = perm if perm.any() else None
    kappa = O.lognormal_kappa(n ** 3, seed=13)
    model = gh.AffineDiscreteModel(x0=x0, J=J, cell_facets=cf, nfacets=nf, facet_perm=fperm)
    X = gh.HDGSpaces.darcy_baseline(model, k)
    op = gh.HybridAffineFEOperator(gh.DarcyHDG(tau=1.0, kappa=kappa, source=O.MMS_SINPROD),
                                   X, X, [1, 2], [3], ctx=ctx)
    assert op.factor_rows()[1] == REF
    ref = O.hdg_solve_mesh(O.Problem.darcy(3, k, mms=O.MMS_SINPROD),
                           meshgen.oracle_mesh(x0, J, cf, nf, fperm), coef=kappa)
    _check_vs_oracle(op, ref, 1e-11)
    op.close()


def test_ref_needs_positive_orientation(ctx):
    """A cell with det J < 0 (a reflected parametrisation) falls outside the
    cofactor identity of the reference decomposition: the class then keeps
    family 8, and the result still matches the oracle."""
    import gridaphybrid as gh
    import meshgen
    n, k = 2, 2
    x0, J, cf, nf, perm, rot = meshgen.sheared_rotated_mesh(3, n, seed=14, frac=0.0)
    # reflect cell 3 in its first reference direction: xi_0 -> 1 - xi_0, which
    # swaps its local facets 4 and 5 (the ones fixing direction 0) and flips
    # the first coordinate of the facets that fix directions 1 and 2
    x0[3] = x0[3] + J[3][:, 0]
    J[3][:, 0] = -J[3][:, 0]
    cf[3][[4, 5]] = cf[3][[5, 4]]
    perm[:] = 0
    perm[3] = [2, 2, 2, 2, 0, 0]
    kappa = O.lognormal_kappa(n ** 3, seed=15)
    model = gh.AffineDiscreteModel(x0=x0, J=J, cell_facets=cf, nfacets=nf, facet_perm=perm)
    X = gh.HDGSpaces.darcy_baseline(model, k)
    op = gh.HybridAffineFEOperator(gh.DarcyHDG(tau=1.0, kappa=kappa, source=O.MMS_SINPROD),
                                   X, X, [1, 2], [3], ctx=ctx)
    assert op.factor_rows()[1] == FLUX
    with pytest.raises(gh.HDGError):
        op.set_family(REF)
    ref = O.hdg_solve_mesh(O.Problem.darcy(3, k, mms=O.MMS_SINPROD),
                           meshgen.oracle_mesh(x0, J, cf, nf, perm), coef=kappa)
    _check_vs_oracle(op, ref, 1e-11)
    # the reflected cell leaves the discrete problem unchanged: the skeleton
    # solution equals the one of the unreflected mesh
    x0u, Ju, cfu, _, _, _ = meshgen.sheared_rotated_mesh(3, n, seed=14, frac=0.0)
    ref_u = O.hdg_solve_mesh(O.Problem.darcy(3, k, mms=O.MMS_SINPROD),
                             meshgen.oracle_mesh(x0u, Ju, cfu, nf, None), coef=kappa)
    assert rel(ref.lam, ref_u.lam) < 1e-11
    op.close()
