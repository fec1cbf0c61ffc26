"""GPU parity: libhdg (gfx950) against the CPU oracle on the same seeded inputs.

Tolerances: the FP64 path must match the oracle to relative 1e-12 (max-norm
error over max-norm of the oracle value); integer outputs (dof maps, CSC
pattern) must match bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.max(np.abs(a - b)) if a.size else 0.0
    s = np.max(np.abs(b)) if b.size else 1.0
    return d / (s if s > 0 else 1.0)


def random_block_system(rng, sizes, nc, untouched=True):
    """Cell systems with the block pattern of test/StaticCondensationMapTests.jl:6-23
    (interior-interior off-diagonal blocks untouched = zero), made well conditioned."""
    n = sum(sizes)
    A = rng.standard_normal((nc, n, n))
    if untouched and len(sizes) == 3:
        o = np.cumsum([0] + list(sizes))
        A[:, o[1]:o[2], o[1]:o[3]] = 0.0
        A[:, o[2]:o[3], o[1]:o[3]] = 0.0
        A[:, o[1]:o[2], o[1]:o[2]] = rng.standard_normal((nc, sizes[1], sizes[1]))
    A += n * np.eye(n)[None]
    b = rng.standard_normal((nc, n))
    return A, b


@pytest.mark.parametrize("sizes", [(3, 3, 3), (8, 4, 8), (18, 9, 12), (24, 8, 24), (81, 27, 54)])
def test_static_condensation_map(ctx, sizes):
    import gridaphybrid as gh
    rng = np.random.default_rng(1234 + sum(sizes))
    nc = 37
    A, b = random_block_system(rng, sizes, nc)
    ni = sizes[0] + sizes[1]
    nb = sizes[2]
    print(f"static condensation sizes={sizes} ncells={nc}", flush=True)
    S, g = gh.StaticCondensationMap([1, 2], [3], ctx).evaluate(A, b, list(sizes))
    So, go, info = O.condense(A, b, ni, nb)
    assert not info.any()
    assert rel(S.cpu().numpy(), So) < RTOL
    assert rel(g.cpu().numpy(), go) < RTOL
    # backward map on the same cells
    x = rng.standard_normal((nc, nb))
    q, u, xx = gh.BackwardStaticCondensationMap([1, 2], [3], ctx).evaluate(A, b, x, list(sizes))
    xio, info = O.recover(A, b, x, ni, nb)
    xi = np.concatenate([q.cpu().numpy(), u.cpu().numpy()], axis=1)
    assert rel(xi, xio) < RTOL
    assert np.array_equal(xx.cpu().numpy(), x)


def test_static_condensation_field_order(ctx):
    """interior/boundary fields need not be contiguous: condense field 2 out."""
    import gridaphybrid as gh
    rng = np.random.default_rng(7)
    sizes = [5, 4, 6]
    A, b = random_block_system(rng, sizes, 9, untouched=False)
    S, g = gh.StaticCondensationMap([2], [1, 3], ctx).evaluate(A, b, sizes)
    perm = np.r_[5:9, 0:5, 9:15]
    Ap = A[:, perm][:, :, perm]
    So, go, info = O.condense(Ap, b[:, perm], 4, 11)
    assert rel(S.cpu().numpy(), So) < RTOL and rel(g.cpu().numpy(), go) < RTOL


def test_singular_cell_reports_info(ctx):
    """A cell whose interior block is all zero fails alone (getrf info = 1)."""
    import gridaphybrid as gh
    rng = np.random.default_rng(3)
    A, b = random_block_system(rng, (3, 3, 3), 4)
    A[2, :6, :6] = 0.0
    S, g, info = gh.StaticCondensationMap([1, 2], [3], ctx).evaluate(A, b, [3, 3, 3],
                                                                    return_info=True)
    info = info.cpu().numpy()
    assert info[2] == 1 and info[[0, 1, 3]].sum() == 0
    So, go, _ = O.condense(A[[0, 1, 3]], b[[0, 1, 3]], 6, 3)
    assert rel(S.cpu().numpy()[[0, 1, 3]], So) < RTOL
    with pytest.raises(gh.SingularError):
        gh.StaticCondensationMap([1, 2], [3], ctx).evaluate(A, b, [3, 3, 3])


def _device_cell_systems(op, first, count):
    import torch
    n = op.info.ni + op.info.nb
    A = torch.empty((count, n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((count, n), dtype=torch.float64, device="cuda")
    from gridaphybrid import _lib
    _lib.check(_lib.lib().hdg_cell_systems(op.mesh, first, count, A.data_ptr(), b.data_ptr()),
               op.ctx.h, "cell_systems")
    op.ctx.synchronize()
    return A.transpose(1, 2).cpu().numpy(), b.cpu().numpy()


def _operator(ctx, D, k, n, kappa=None, source=None, spaces=None, mms=O.MMS_LINEAR, tau=1.0):
    import gridaphybrid as gh
    model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
    X = spaces(model) if spaces else gh.HDGSpaces.darcy_baseline(model, k)
    wf = gh.DarcyHDG(tau=tau, kappa=1.0 if kappa is None else kappa,
                     source=mms if source is None else source)
    return gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx)


def check_local_factors(op, A, b, U, Y, ncheck):
    """U and Y = L^-1 P [B c] are partial-pivoting LU factors of the system the
    recovery back-substitutes: A_ii itself, or, after the flux-block
    elimination, R_uu of the reduced augmented system
    [R_uu R_ub r_u] = [A_(u,b)(u,b) b_(u,b)] - A_(u,b)q A_qq^-1 [A_q(u,b) b_q]
    formed here in numpy from the cell systems.  (The pivot order itself may
    differ from another LU on exact ties.)"""
    import scipy.linalg as sl
    ni = op.info.ni
    rows, fam = op.factor_rows()
    assert U is not None
    na = ni - rows
    for c in range(ncheck):
        Ab = np.concatenate([A[c], b[c][:, None]], axis=1)
        if na:
            assert fam in (8, 10)
            Z = np.linalg.solve(Ab[:na, :na], Ab[:na, na:])
            M = Ab[na:, na:] - Ab[na:, :na] @ Z
        else:
            M = Ab
        Aii = M[:rows, :rows]
        B = M[:rows, rows:]   # nb boundary columns and the right-hand side
        assert rel(sl.solve_triangular(U[c], Y[c]), np.linalg.solve(Aii, B)) < 1e-10
        # |L| <= 1 up to the rounding of inv(U) (cond(U) times eps; a factor
        # without partial pivoting has multipliers far above 1)
        PtL = Aii @ np.linalg.inv(U[c])
        assert np.max(np.abs(PtL)) < 1.0 + 1e-8
        assert np.allclose(np.sort(np.max(np.abs(PtL), axis=1)), 1.0, atol=0) or \
            np.count_nonzero(np.isclose(np.abs(PtL), 1.0, rtol=0, atol=1e-8)) >= rows


CASES = [
    # (D, k, n, mms, lognormal kappa)
    (2, 1, 8, O.MMS_LINEAR, False),     # cfg1
    (2, 2, 16, O.MMS_SINPROD, True),    # cfg2 shape, smaller mesh
    (2, 3, 6, O.MMS_SINPROD, True),
    (3, 1, 3, O.MMS_LINEAR, True),
    (3, 2, 3, O.MMS_SINPROD, True),     # cfg3 shape, smaller mesh
]


@pytest.mark.parametrize("D,k,n,mms,logn", CASES)
def test_operator_parity(ctx, D, k, n, mms, logn):
    print(f"operator parity D={D} k={k} n={n} mms={mms} lognormal={logn}", flush=True)
    import gridaphybrid as gh
    nc = n ** D
    kappa = O.lognormal_kappa(nc, seed=20261202) if logn else np.ones(nc)
    prob = O.Problem.darcy(D, k, mms=mms)
    ref = O.hdg_solve(prob, tuple([n] * D), coef=kappa)
    op = _operator(ctx, D, k, n, kappa=kappa, mms=mms)
    assert op.info.nfree == ref.nfree and op.info.nnz == len(ref.rowval)
    # element systems (fused generator) vs per-cell quadrature
    A, b = _device_cell_systems(op, 0, nc)
    assert rel(A, ref.A) < RTOL and rel(b, ref.b) < RTOL
    # condensed blocks and dof ids
    S, g, cd = op.cell_blocks()
    assert np.array_equal(cd, ref.cell_dofs)
    assert rel(S, ref.S) < RTOL and rel(g, ref.g) < RTOL
    # skeleton CSC: pattern bit-exact, values to tolerance
    K, rhs = op.skeleton_system()
    assert np.array_equal(K.indptr, ref.colptr) and np.array_equal(K.indices, ref.rowval)
    assert rel(K.data, ref.nzval) < RTOL and rel(rhs, ref.rhs) < RTOL
    # local factors: U and Y = L^-1 P [A_ib b_i] solve the local problems, and
    # the pivots are partial-pivoting ones: A_ii U^-1 = P^T L with |L| <= 1
    # (the pivot order itself may differ from another LU on exact ties)
    import scipy.linalg as sl
    U, Y = op.local_factors()
    check_local_factors(op, A, b, U, Y, min(nc, 16))
    # recovery from the oracle's skeleton solution
    fv = op.recover(ref.lam).cpu().numpy()
    assert rel(fv, ref.free_values) < RTOL
    # full solve through the public API
    sol = gh.solve(op)
    assert rel(sol.free_values, ref.free_values) < 1e-10
    op.close()


def test_reference_kat_on_gpu(ctx):
    """test/DarcyHDGTests.jl:170-174 on the GPU path: P1/P0/P1 on 2x2 quads,
    ||u - u_h|| < 1e-12 (DarcyHDGTests.jl:142)."""
    import gridaphybrid as gh
    model = gh.CartesianDiscreteModel((0.0, 1.0, 0.0, 1.0), (2, 2))
    X = gh.HDGSpaces.darcy_reference_test(model)
    op = gh.HybridAffineFEOperator(gh.DarcyHDG(tau=1.0, source=O.MMS_KAT), X, X, [1, 2], [3],
                                   ctx=ctx)
    sol = gh.solve(op)
    prob = O.Problem.darcy_reference_kat(2)
    mesh = O.cartesian(2, (2, 2))
    xi = np.concatenate([sol.flux, sol.potential], axis=1)
    err = O.l2_errors(prob, mesh, np.ones(4), xi)
    assert err[0] < 1e-12 and err[1] < 1e-12


def test_condense_deterministic(ctx):
    """Two assemblies of the same operator are bitwise identical (no atomics)."""
    kappa = O.lognormal_kappa(24 * 24, seed=5)
    op = _operator(ctx, 2, 2, 24, kappa=kappa, mms=O.MMS_SINPROD)
    K1, r1 = op.skeleton_system()
    op.assemble()
    K2, r2 = op.skeleton_system()
    assert np.array_equal(K1.data, K2.data) and np.array_equal(r1, r2)
    op.close()


@pytest.mark.parametrize("D,k,n", [(2, 2, 512), (3, 2, 32)])
def test_full_size_consistency(ctx, D, k, n):
    """At benchmark sizes: with a manufactured solution that lies in the
    discrete spaces (linear potential), recovery from the exact trace must
    reproduce the exact flux and potential (consistency of every cell's
    condense + recover), whatever the per-cell kappa."""
    print(f"full-size consistency D={D} k={k} n={n}", flush=True)
    import gridaphybrid as gh
    nc = n ** D
    kappa = O.lognormal_kappa(nc, seed=11)
    op = _operator(ctx, D, k, n, kappa=kappa, mms=O.MMS_LINEAR)
    # exact trace: the linear potential at the free trace nodes, in free-dof
    # order (free facets consecutive in facet order)
    prob = O.Problem.darcy(D, k, mms=O.MMS_LINEAR)
    mesh = O.cartesian(D, tuple([n] * D))
    cd, nfree, ndir, fd = O.trace_dofs(mesh, prob.sizes()["n_tf"])
    inv = 1 - fd
    lam_all = O.dirichlet_values(prob, mesh, inv, int(inv.sum()) * prob.sizes()["n_tf"])
    assert len(lam_all) == op.info.nfree
    fv = op.recover(lam_all).cpu().numpy()
    nf = op.info.n_flux
    xi = np.concatenate([fv[:nc * nf].reshape(nc, nf),
                         fv[nc * nf:nc * op.info.ni].reshape(nc, -1)], axis=1)
    err = O.l2_errors(prob, mesh, kappa, xi)
    assert err[0] < 1e-9 * np.sqrt(np.mean(kappa ** 2)) + 1e-9 and err[1] < 1e-10
    op.close()


@pytest.mark.parametrize("sizes", [(150, 50, 100), (120, 108, 180)])
def test_static_condensation_map_large(ctx, sizes):
    """Local systems beyond the register-resident kernels (blocked kernel)."""
    import gridaphybrid as gh
    rng = np.random.default_rng(99 + sum(sizes))
    nc = 5
    A, b = random_block_system(rng, sizes, nc, untouched=False)
    ni, nb = sizes[0] + sizes[1], sizes[2]
    print(f"large static condensation sizes={sizes}", flush=True)
    S, g = gh.StaticCondensationMap([1, 2], [3], ctx).evaluate(A, b, list(sizes))
    So, go, info = O.condense(A, b, ni, nb)
    assert rel(S.cpu().numpy(), So) < RTOL and rel(g.cpu().numpy(), go) < RTOL
    x = rng.standard_normal((nc, nb))
    q, u, _ = gh.BackwardStaticCondensationMap([1, 2], [3], ctx).evaluate(A, b, x, list(sizes))
    xio, _ = O.recover(A, b, x, ni, nb)
    assert rel(np.concatenate([q.cpu().numpy(), u.cpu().numpy()], axis=1), xio) < RTOL


def _elasticity_operator(ctx, D, k, n, nu, source):
    import gridaphybrid as gh
    model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
    X = gh.HDGSpaces.elasticity(model, k)
    wf = gh.ElasticityHDG(tau=float(n), nu=nu, E=1.0, source=source)
    return gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx)


@pytest.mark.parametrize("D,k,n", [(2, 1, 4), (3, 1, 2), (3, 2, 2), (3, 3, 1)])
def test_elasticity_operator_parity(ctx, D, k, n):
    """3D linear-elasticity HDG (test/LinearElasticityHDGTests.jl:140-150),
    nu varying per cell in [0.2, 0.4]."""
    print(f"elasticity parity D={D} k={k} n={n}", flush=True)
    import gridaphybrid as gh
    nc = n ** D
    nu = np.random.default_rng(5).uniform(0.2, 0.4, nc)
    prob = O.Problem.elasticity(D, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN)
    ref = O.hdg_solve(prob, (n,) * D, coef=nu)
    op = _elasticity_operator(ctx, D, k, n, nu, 12)
    A, b = _device_cell_systems(op, 0, nc)
    assert rel(A, ref.A) < RTOL and rel(b, ref.b) < RTOL
    # k=3 (ni=225, nb=180) has cond(A_ii) ~ 3e8: the oracle's own S and g move
    # at the 1e-12 .. 1e-11 level when only the order of the interior unknowns
    # changes; the tolerance is ten times that measured self-variation
    ni, nbb = op.info.ni, op.info.nb
    perm = np.r_[np.random.default_rng(1).permutation(ni), ni:ni + nbb]
    S2, g2, _ = O.condense(ref.A[:, perm][:, :, perm], ref.b[:, perm], ni, nbb)
    tol_S = max(RTOL, 10 * rel(S2, ref.S))
    tol_g = max(RTOL, 10 * rel(g2, ref.g))
    S, g, cd = op.cell_blocks()
    assert np.array_equal(cd, ref.cell_dofs)
    assert rel(S, ref.S) < tol_S and rel(g, ref.g) < tol_g
    K, rhs = op.skeleton_system()
    assert np.array_equal(K.indptr, ref.colptr) and np.array_equal(K.indices, ref.rowval)
    assert rel(K.data, ref.nzval) < tol_S and rel(rhs, ref.rhs) < tol_g
    fv = op.recover(ref.lam).cpu().numpy()
    assert rel(fv, ref.free_values) < (1e-9 if k == 3 else 1e-11)
    op.close()


def test_darcy_3d_k3_large(ctx):
    """3D Darcy HDG k=3: ni=256, nb=96, n_flux=192 (the flux elimination from
    reference products with a 160-row reduced system, the recovery reading
    G_e from the table; the blocked kernel is tested in
    test_blocked_launch_shapes)."""
    print("darcy 3D k=3", flush=True)
    nc = 8
    kappa = O.lognormal_kappa(nc, seed=8)
    prob = O.Problem.darcy(3, 3, mms=O.MMS_SINPROD)
    ref = O.hdg_solve(prob, (2, 2, 2), coef=kappa)
    op = _operator(ctx, 3, 3, 2, kappa=kappa, mms=O.MMS_SINPROD)
    assert op.factor_rows() == (64, 10)
    S, g, cd = op.cell_blocks()
    assert rel(S, ref.S) < RTOL and rel(g, ref.g) < RTOL
    fv = op.recover(ref.lam).cpu().numpy()
    assert rel(fv, ref.free_values) < 1e-11
    op.close()


@pytest.mark.parametrize("shape", [4, 8, 16, 108, 116])
def test_blocked_launch_shapes(ctx, shape):
    """Every launch shape of the blocked elimination (16-column panels with 4,
    8 or 16 wavefronts per workgroup; 32-column panels, shape 100 + w) on the
    3D Darcy k=3 class (ni=256, nb=96) against the oracle."""
    print(f"blocked elimination, launch shape {shape}", flush=True)
    nc = 8
    kappa = O.lognormal_kappa(nc, seed=9)
    prob = O.Problem.darcy(3, 3, mms=O.MMS_SINPROD)
    ref = O.hdg_solve(prob, (2, 2, 2), coef=kappa)
    op = _operator(ctx, 3, 3, 2, kappa=kappa, mms=O.MMS_SINPROD)
    op.set_family(4, shape)
    op.assemble()
    assert op.factor_rows()[1] == 4
    S, g, cd = op.cell_blocks()
    assert rel(S, ref.S) < RTOL and rel(g, ref.g) < RTOL
    K, rhs = op.skeleton_system()
    assert rel(K.data, ref.nzval) < RTOL and rel(rhs, ref.rhs) < RTOL
    fv = op.recover(ref.lam).cpu().numpy()
    assert rel(fv, ref.free_values) < 1e-11
    op.close()


@pytest.mark.parametrize("D,k,n", [(2, 1, 4), (3, 1, 2), (3, 2, 2)])
def test_elasticity_projection_parity(ctx, D, k, n):
    """Elasticity HDG with the bulk-to-skeleton projection P_m
    (test/P_m.jl:4-23, bulk_to_skeleton_projection=true of
    test/LinearElasticityHDGTests.jl:124-150) against the oracle."""
    import gridaphybrid as gh
    print(f"elasticity with P_m D={D} k={k} n={n}", flush=True)
    nc = n ** D
    nu = np.random.default_rng(6).uniform(0.2, 0.4, nc)
    prob = O.Problem.elasticity(D, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN, projection=1)
    ref = O.hdg_solve(prob, (n,) * D, coef=nu)
    model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
    X = gh.HDGSpaces.elasticity(model, k)
    wf = gh.ElasticityHDG(tau=float(n), nu=nu, E=1.0, source=12, bulk_to_skeleton_projection=True)
    op = gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx)
    A, b = _device_cell_systems(op, 0, nc)
    assert rel(A, ref.A) < RTOL and rel(b, ref.b) < RTOL
    S, g, cd = op.cell_blocks()
    assert rel(S, ref.S) < 1e-11 and rel(g, ref.g) < 1e-11
    K, rhs = op.skeleton_system()
    assert rel(K.data, ref.nzval) < 1e-11 and rel(rhs, ref.rhs) < 1e-11
    sol = gh.solve(op)
    assert rel(sol.free_values, ref.free_values) < 1e-10
    op.close()
