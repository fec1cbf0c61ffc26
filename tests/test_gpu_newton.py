"""The hybrid Newton path (src/HybridFEOperators.jl, src/HybridLinearSolvers.jl:15-59)
on the GPU: Newton's method whose linear solves condense, solve the skeleton
system on the host and recover, as test/DarcyHDGTests.jl:157-165 runs it."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_newton_darcy_reference_kat(ctx):
    """test/DarcyHDGTests.jl:157-165: HybridFEOperator(residual, X, Y, [1,2], [3]),
    NLSolver(method=:newton), solve! from xh0 = 0 on 2x2 quads with P1/P0/P1:
    ||u - u_h|| < 1e-12 (:165), and the result equals the oracle's affine
    hybrid solve."""
    import gridaphybrid as gh
    model = gh.CartesianDiscreteModel((0.0, 1.0, 0.0, 1.0), (2, 2))
    X = gh.HDGSpaces.darcy_reference_test(model)
    residual = gh.DarcyHDG(tau=1.0, source=O.MMS_KAT)
    op = gh.HybridFEOperator(residual, X, X, [1, 2], [3], ctx=ctx)
    xh0 = gh.FEFunction(np.zeros(op.num_free_dofs()))
    xh, tr = gh.solve_newton(xh0, gh.FESolver(gh.NLSolver(method="newton")), op)
    print("newton trace", tr, flush=True)
    assert tr.converged and tr.iterations <= 2
    prob = O.Problem.darcy_reference_kat(2)
    ref = O.hdg_solve(prob, (2, 2))
    assert O.rel_err(xh.free_values, ref.free_values) < 1e-12
    nc, ni, na = op.info.ncells, op.info.ni, op.info.n_flux
    fv = xh.free_values
    xi = np.concatenate([fv[:nc * na].reshape(nc, na), fv[nc * na:nc * ni].reshape(nc, -1)],
                        axis=1)
    err = O.l2_errors(prob, O.cartesian(2, (2, 2)), np.ones(nc), xi)
    assert err[0] < 1e-12 and err[1] < 1e-12
    op.close()


@pytest.mark.parametrize("phys,D,k,n", [("darcy", 3, 2, 4), ("darcy", 2, 3, 6),
                                        ("elasticity", 3, 1, 2)])
def test_newton_from_random_start(ctx, phys, D, k, n):
    """From a random initial guess (Dirichlet trace values still those of the
    trial space) Newton converges in one step to the affine solution."""
    import gridaphybrid as gh
    print(f"newton {phys} D={D} k={k} n={n}", flush=True)
    nc = n ** D
    model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
    if phys == "darcy":
        kappa = O.lognormal_kappa(nc, seed=3)
        X = gh.HDGSpaces.darcy_baseline(model, k)
        residual = gh.DarcyHDG(tau=1.0, kappa=kappa, source=O.MMS_SINPROD)
        ref = O.hdg_solve(O.Problem.darcy(D, k, mms=O.MMS_SINPROD), (n,) * D, coef=kappa)
    else:
        nu = np.random.default_rng(4).uniform(0.2, 0.4, nc)
        X = gh.HDGSpaces.elasticity(model, k)
        residual = gh.ElasticityHDG(tau=float(n), nu=nu, E=1.0, source=gh._lib.SOURCE_ELAS_SIN)
        ref = O.hdg_solve(O.Problem.elasticity(D, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN),
                          (n,) * D, coef=nu)
    op = gh.HybridFEOperator(residual, X, X, [1, 2], [3], ctx=ctx)
    x0 = np.random.default_rng(9).standard_normal(op.num_free_dofs())
    xh, tr = gh.solve_newton(gh.FEFunction(x0), gh.FESolver(gh.NLSolver(ftol=1e-11)), op)
    print("newton trace", tr, flush=True)
    assert tr.converged and tr.iterations == 1
    assert O.rel_err(xh.free_values, ref.free_values) < 1e-10
    op.close()
