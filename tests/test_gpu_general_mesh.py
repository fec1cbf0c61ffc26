"""hdg_mesh_create on meshes that are not CartesianDiscreteModels: sheared
parallelogram / parallelepiped cells, a seeded subset of them re-parametrised
by rotations of the reference cell, with the orientation code (the reference's
pindex, src/SkeletonArrays.jl:28-37, src/Skeleton.jl:587-612) of every
(cell, local facet).  The GPU path is compared with the oracle on the same
mesh description; the oracle itself is pinned on such meshes by
tests/test_oracle.py::test_oriented_mesh_reproduces_unrotated_solution."""
import numpy as np
import pytest

from oracle import oracle as O

import meshgen

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _model(gh, x0, J, cf, nf, perm):
    return gh.AffineDiscreteModel(x0=x0, J=J, cell_facets=cf, nfacets=nf, facet_perm=perm)


@pytest.mark.parametrize("phys,D,k,n,frac", [
    ("darcy", 2, 1, 5, 0.6), ("darcy", 2, 3, 4, 0.6), ("darcy", 3, 1, 3, 0.6),
    ("darcy", 3, 2, 3, 0.6), ("darcy", 3, 2, 3, 0.0), ("elasticity", 2, 1, 4, 0.6)])
def test_general_mesh_parity(ctx, phys, D, k, n, frac):
    import gridaphybrid as gh
    print(f"general mesh {phys} D={D} k={k} n={n} rotated fraction {frac}", flush=True)
    nc = n ** D
    x0, J, cf, nf, perm, rot = meshgen.sheared_rotated_mesh(D, n, seed=7 + D + k, frac=frac)
    fperm = perm if perm.any() else None
    model = _model(gh, x0, J, cf, nf, fperm)
    if phys == "darcy":
        coef = O.lognormal_kappa(nc, seed=17)
        X = gh.HDGSpaces.darcy_baseline(model, k)
        wf = gh.DarcyHDG(tau=1.0, kappa=coef, source=O.MMS_SINPROD)
        prob = O.Problem.darcy(D, k, mms=O.MMS_SINPROD)
    else:
        coef = np.random.default_rng(3).uniform(0.2, 0.4, nc)
        X = gh.HDGSpaces.elasticity(model, k)
        wf = gh.ElasticityHDG(tau=float(n), nu=coef, E=1.0, source=gh._lib.SOURCE_ELAS_SIN)
        prob = O.Problem.elasticity(D, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN)
    ref = O.hdg_solve_mesh(prob, meshgen.oracle_mesh(x0, J, cf, nf, fperm), coef=coef)
    op = gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx)
    try:
        assert op.info.nfree == ref.nfree and op.info.nnz == len(ref.rowval)
        S, g, cd = op.cell_blocks()
        assert np.array_equal(cd, ref.cell_dofs)
        assert O.rel_err(S, ref.S) < RTOL and O.rel_err(g, ref.g) < RTOL
        K, rhs = op.skeleton_system()
        assert np.array_equal(K.indptr, ref.colptr) and np.array_equal(K.indices, ref.rowval)
        assert O.rel_err(K.data, ref.nzval) < RTOL and O.rel_err(rhs, ref.rhs) < RTOL
        # the interpolated trace in each facet's own node order
        lam_all = op.interpolate_trace()
        mesh = meshgen.oracle_mesh(x0, J, cf, nf, fperm)
        inv = (1 - ref.mesh.facet_boundary).astype(np.int8)
        if fperm is None:
            cd_i, _, nd_i, fd_i = O.trace_dofs(mesh, prob.sizes()["n_tf"], inv)
            lam_o = O.dirichlet_values(prob, mesh, fd_i, nd_i)
        else:
            cd_i, _, nd_i, fd_i = O.trace_dofs_oriented(prob, mesh, inv)
            lam_o = O.dirichlet_values_oriented(prob, mesh, fd_i, nd_i)
        assert O.rel_err(lam_all, lam_o) < 1e-15
        fv = op.recover(ref.lam).cpu().numpy()
        assert O.rel_err(fv, ref.free_values) < 1e-11
        sol = gh.solve(op)
        assert O.rel_err(sol.free_values, ref.free_values) < 1e-10
    finally:
        op.close()


def test_unsupported_orientation_is_refused(ctx):
    """P_1 trace nodes on a hexahedron's facet ((0,0), (1,0), (0,1)) are not
    mapped onto themselves by a flip of the facet frame: the library refuses
    such a mesh instead of numbering the trace wrongly."""
    import gridaphybrid as gh
    x0, J, cf, nf, perm, rot = meshgen.sheared_rotated_mesh(3, 2, seed=5, frac=1.0)
    assert (perm & 6).any()
    model = _model(gh, x0, J, cf, nf, perm)
    X = gh.HDGSpaces.elasticity(model, 1)
    wf = gh.ElasticityHDG(tau=2.0, nu=0.3, E=1.0, source=gh._lib.SOURCE_ELAS_SIN)
    with pytest.raises(gh.HDGError) as e:
        gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx)
    assert e.value.code == gh._lib.HDG_ERR_UNSUPPORTED
