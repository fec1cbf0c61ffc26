"""Multi-field skeleton (src/HybridAffineFEOperators.jl:52-63, :343-356,
src/Scalar2ArrayBlockMap.jl:42-69): the elasticity HDG displacement trace as
D scalar trace fields, HybridAffineFEOperator(wf, X, Y, [1,2], [3,...,2+D]).
The skeleton system and the trace are numbered field by field (all dofs of
field 3, then field 4, ...); the oracle's single-field result, renumbered
here from the definition of that numbering, must match it."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def field_major(nfree_facets, D, nm):
    """facet-major index G (D nm) + c nm + j  ->  field-major c (F nm) + G nm + j"""
    G, c, j = np.meshgrid(np.arange(nfree_facets), np.arange(D), np.arange(nm), indexing="ij")
    o = (G * D * nm + c * nm + j).ravel()
    n = (c * nfree_facets * nm + G * nm + j).ravel()
    o2n = np.empty(len(o), np.int64)
    o2n[o] = n
    return o2n


@pytest.mark.parametrize("D,k,n", [(2, 1, 3), (2, 2, 2), (3, 1, 2)])
def test_multifield_trace(ctx, D, k, n):
    import gridaphybrid as gh
    print(f"multi-field trace D={D} k={k} n={n}", flush=True)
    nc = n ** D
    nu = np.random.default_rng(8).uniform(0.2, 0.4, nc)
    prob = O.Problem.elasticity(D, k, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN)
    ref = O.hdg_solve(prob, (n,) * D, coef=nu)
    model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
    X = gh.HDGSpaces.elasticity(model, k)
    wf = gh.ElasticityHDG(tau=float(n), nu=nu, E=1.0, source=gh._lib.SOURCE_ELAS_SIN)
    op = gh.HybridAffineFEOperator(wf, X, X, [1, 2], list(range(3, 3 + D)), ctx=ctx)
    try:
        ntf = op.info.n_trace_facet
        nm = ntf // D
        o2n = field_major(ref.nfree // ntf, D, nm)
        od2n = field_major(ref.ndir // ntf, D, nm)
        # skeleton system: the oracle's, renumbered field by field
        Kr = sp.csc_matrix((ref.nzval, ref.rowval, ref.colptr), shape=(ref.nfree, ref.nfree))
        P = sp.csc_matrix((np.ones(ref.nfree), (o2n, np.arange(ref.nfree))),
                          shape=(ref.nfree, ref.nfree))
        Kf = (P @ Kr @ P.T).tocsc()
        Kf.sort_indices()
        rhs_f = np.empty(ref.nfree)
        rhs_f[o2n] = ref.rhs
        K, rhs = op.skeleton_system()
        assert np.array_equal(K.indptr, Kf.indptr) and np.array_equal(K.indices, Kf.indices)
        assert O.rel_err(K.data, Kf.data) < 1e-12 and O.rel_err(rhs, rhs_f) < 1e-12
        # cell dofs in the field-major numbering (Dirichlet ids renumbered too)
        S, g, cd = op.cell_blocks()
        cdr = ref.cell_dofs
        exp = np.where(cdr >= 0, o2n[np.maximum(cdr, 0)], -od2n[np.maximum(-cdr - 1, 0)] - 1)
        assert np.array_equal(cd, exp)
        # condensed blocks re-blocked per field (Scalar2ArrayBlockMap)
        idx = op.skeleton_blocks()
        assert len(idx) == D and sorted(np.concatenate(idx)) == list(range(op.info.nb))
        perm = np.concatenate(idx)
        Sb, gb = gh.Scalar2ArrayBlockMap().evaluate((S[:, perm][:, :, perm], g[:, perm]),
                                                    [len(i) for i in idx])
        for a in range(D):
            assert O.rel_err(gb[a], ref.g[:, idx[a]]) < 1e-12
            for b in range(D):
                assert O.rel_err(Sb[a][b], ref.S[:, idx[a]][:, :, idx[b]]) < 1e-12
        # solve: interior fields as before, the trace field by field
        sol = gh.solve(op)
        ni = op.info.ni
        fv = sol.free_values
        assert O.rel_err(fv[:nc * ni], ref.free_values[:nc * ni]) < 1e-10
        lam_f = np.empty(ref.nfree)
        lam_f[o2n] = ref.lam
        assert O.rel_err(fv[nc * ni:], lam_f) < 1e-10
    finally:
        op.close()
