"""Pins of the CPU oracle (oracle/hdg_oracle.c) -- CPU only.

Each function of the oracle is checked against something it does not share
code with:
  * cell systems  -> exact rational integration (tests/golden/make_golden.py);
  * condensation / recovery -> LAPACK dgetrf/dgetrs golden vectors (scipy);
  * the end-to-end path -> the reference's own known answer
    (test/DarcyHDGTests.jl:142: ||u - u_h|| < 1e-12 on 2x2 quads), a
    monolithic solve of the uncondensed system, and convergence rates;
  * dof maps / CSC pattern -> brute-force counts and the survey's counts.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


def one_box_cell(D, x0, h, coef, signs):
    """A one-cell 'mesh' holding a box cell with the given geometry and signs."""
    mesh = O.cartesian(D, (1,) * D)
    mesh.x0[:] = x0
    J = np.zeros((D, D))
    np.fill_diagonal(J, h)
    mesh.J[:] = J.T.ravel()
    mesh.cell_sign[:] = signs
    return mesh


@pytest.mark.parametrize("name", ["darcy_q1_2d", "darcy_q2_2d", "darcy_p1p0p1_2d", "darcy_q1_3d",
                                  "darcy_q2_3d", "darcy_q3_2d"])
def test_darcy_cell_system_exact(name):
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    D = int(g["D"])
    kf, kp, kt = (int(v) for v in g["k"])
    prob = O.Problem(O.PHYS_DARCY, D, int(g["space"]), kf, kp, kt, 2 * (kf + 1), float(g["tau"]),
                     1.0, 0.3, O.MMS_KAT)
    mesh = one_box_cell(D, g["x0"], g["h"], g["kappa"], g["signs"])
    A, b = O.cell_systems(prob, mesh, np.array([float(g["kappa"])]))
    assert A.shape[1:] == g["A"].shape
    assert rel(A[0], g["A"]) < 1e-13
    assert rel(b[0], g["b"]) < 1e-13


@pytest.mark.parametrize("name", ["elas_p1_2d", "elas_p1_3d", "elas_p2_3d", "elas_p1_2d_pm",
                                  "elas_p2_3d_pm"])
def test_elasticity_cell_system_exact(name):
    """Elasticity HDG cell systems against exact rational integration, without
    and with the bulk-to-skeleton projection P_m (test/P_m.jl:4-23)."""
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    D, k = int(g["D"]), int(g["k"])
    pm = int(g["projection"]) if "projection" in g.files else 0
    prob = O.Problem.elasticity(D, k, tau=float(g["tau"]), nu=float(g["nu"]), mms=0,
                                projection=pm)
    mesh = one_box_cell(D, np.zeros(D), g["h"], float(g["nu"]), np.ones(2 * D))
    A, b = O.cell_systems(prob, mesh, np.array([float(g["nu"])]))
    assert rel(A[0], g["A"]) < 1e-13
    assert np.max(np.abs(b)) == 0.0


def test_condense_recover_vs_lapack():
    """StaticCondensationMap / BackwardStaticCondensationMap op sequence
    (src/StaticCondensationMap.jl:179-192, src/BackwardStaticCondensationMap.jl:91-99)
    against dgetrf/dgetrs, on the untouched-block pattern of
    test/StaticCondensationMapTests.jl."""
    g = np.load(os.path.join(GOLD, "condense_lapack.npz"))
    S, gg, info = O.condense(g["A"], g["b"], 6, 3)
    assert not info.any()
    assert rel(S, g["S"]) < 1e-13 and rel(gg, g["g"]) < 1e-13
    xi, info = O.recover(g["A"], g["b"], g["x"], 6, 3)
    assert rel(xi, g["xi"]) < 1e-13


def test_condense_singular_info():
    """A zero interior block gives LAPACK info = 1 (the reference's @check
    info==0 would throw, src/StaticCondensationMap.jl:180)."""
    g = np.load(os.path.join(GOLD, "condense_lapack.npz"))
    A = g["A"].copy()
    A[1, :6, :6] = 0.0
    _, _, info = O.condense(A, g["b"], 6, 3)
    assert info[1] == 1 and info[[0, 2, 3, 4, 5]].sum() == 0


def test_reference_known_answer_2x2_quads():
    """test/DarcyHDGTests.jl:170-174 with :142 -- P1 flux, P0 potential, P1
    trace on 2x2 quads, tau=1: ||u - u_h||_L2 < 1e-12."""
    p = O.Problem.darcy_reference_kat(2)
    sol = O.hdg_solve(p, (2, 2))
    err = O.l2_errors(p, sol.mesh, sol.coef, sol.xi)
    assert err[0] < 1e-12 and err[1] < 1e-12


@pytest.mark.parametrize("D,k,n", [(2, 1, 4), (2, 2, 3), (2, 3, 2), (3, 1, 2), (3, 2, 2)])
def test_known_answer_baseline_spaces(D, k, n):
    """The same known answer in the Q_k spaces (the exact flux/potential are
    in every Q_k, k >= 1)."""
    p = O.Problem.darcy(D, k, mms=O.MMS_KAT)
    sol = O.hdg_solve(p, (n,) * D)
    err = O.l2_errors(p, sol.mesh, sol.coef, sol.xi)
    assert err[0] < 1e-11 and err[1] < 1e-11


@pytest.mark.parametrize("D,k,n,logn", [(2, 1, 6, False), (2, 2, 5, True), (3, 1, 2, True)])
def test_hybrid_equals_monolithic(D, k, n, logn):
    """Condensation + skeleton solve + recovery equals a direct solve of the
    uncondensed hybridizable system (the reference's DarcyHDGTests.jl:144-155
    re-solve, without Newton)."""
    nc = n ** D
    kappa = O.lognormal_kappa(nc, 3) if logn else np.ones(nc)
    p = O.Problem.darcy(D, k, mms=O.MMS_SINPROD)
    sol = O.hdg_solve(p, (n,) * D, coef=kappa)
    mono = O.monolithic_solve(sol)
    assert rel(sol.free_values, mono) < 1e-10


def test_convergence_rates():
    """With the reference's ownership-signed stabilisation tau (n.n_o) the
    L2 errors converge at rate >= k; with tau > 0 on every facet (standard
    HDG) at rate k+1.  A sign or facet-measure mistake breaks one of them."""
    def rates(k, allplus):
        errs = []
        for n in (4, 8, 16):
            p = O.Problem.darcy(2, k, mms=O.MMS_SINPROD)
            if not allplus:
                sol = O.hdg_solve(p, (n, n))
                errs.append(O.l2_errors(p, sol.mesh, sol.coef, sol.xi))
                continue
            import scipy.sparse as sp
            import scipy.sparse.linalg as spla
            mesh = O.cartesian(2, (n, n))
            mesh.cell_sign[:] = 1
            s = p.sizes()
            coef = np.ones(mesh.ncells)
            A, b = O.cell_systems(p, mesh, coef)
            S, g, _ = O.condense(A, b, s["ni"], s["nb"])
            cd, nfree, ndir, fd = O.trace_dofs(mesh, s["n_tf"])
            ld = O.dirichlet_values(p, mesh, fd, ndir)
            cp, rv, nz, rhs = O.assemble(S, g, cd, nfree, ld)
            lam = spla.spsolve(sp.csc_matrix((nz, rv, cp), shape=(nfree, nfree)), rhs)
            xi, _ = O.recover(A, b, O.gather_lambda(cd, lam, ld), s["ni"], s["nb"])
            errs.append(O.l2_errors(p, mesh, coef, xi))
        errs = np.array(errs)
        return np.log2(errs[-2] / errs[-1])
    for k in (1, 2):
        assert np.all(rates(k, False) > k - 0.2)
        assert np.all(rates(k, True) > k + 1 - 0.25)


def test_owner_signs_and_facet_numbering():
    """Owner = first (lowest id) cell around a facet (src/Skeleton.jl:135-140);
    facets numbered at first encounter; boundary facets owned by their cell."""
    for D, n in ((2, (3, 2)), (3, (2, 2, 2))):
        m = O.cartesian(D, n)
        nF = 2 * D
        seen = {}
        for c in range(m.ncells):
            for lf in range(nF):
                f = m.cell_facets[c * nF + lf]
                if f not in seen:
                    assert f == len(seen)
                    seen[f] = c
                    assert m.cell_sign[c * nF + lf] == 1
                else:
                    assert seen[f] < c and m.cell_sign[c * nF + lf] == -1
        assert len(seen) == m.nfacets


def test_skeleton_counts():
    """cfg1 (8x8, k=1) has 224 free trace dofs and 2,784 nonzeros; cfg2
    (512^2, k=2) 1,569,792 free dofs and 32,910,408 nonzeros.  Brute force:
    nnz = sum over free facets G of ntf^2 x (free facets sharing a cell with G)."""
    for n, k, nfree_ref, nnz_ref in ((8, 1, 224, 2784), (512, 2, 1569792, 32910408)):
        p = O.Problem.darcy(2, k)
        m = O.cartesian(2, (n, n))
        ntf = p.sizes()["n_tf"]
        cd, nfree, ndir, fd = O.trace_dofs(m, ntf)
        assert nfree == nfree_ref
        nnz = O.lib().orc_assemble_nnz(4 * ntf, O.C.c_int64(m.ncells), O.C.c_int64(nfree),
                                       cd.ctypes.data_as(O.C.c_void_p)) if n <= 64 else None
        # facet adjacency count
        cf = m.cell_facets.reshape(-1, 4)
        adj = [set() for _ in range(m.nfacets)]
        if n <= 64:
            for c in range(m.ncells):
                fr = [f for f in cf[c] if not fd[f]]
                for f in fr:
                    adj[f].update(fr)
            brute = sum(len(a) for f, a in enumerate(adj) if not fd[f]) * ntf * ntf
            assert nnz == brute == nnz_ref
        else:
            # interior facets: 7 couplings minus boundary ones, counted per facet
            fb = fd.astype(bool)
            cnt = np.zeros(m.nfacets, np.int64)
            for lf in range(4):
                fcol = cf[:, lf]
                nfree_cell = (~fb[cf]).sum(axis=1)
                np.add.at(cnt, fcol, nfree_cell)
            # each free facet is counted once per adjacent cell; itself twice for
            # interior facets
            interior = np.zeros(m.nfacets, bool)
            fc = m.facet_cells.reshape(-1, 2)
            interior[fc[:, 1] >= 0] = True
            cnt = cnt - interior.astype(np.int64)
            assert int(cnt[~fb].sum()) * ntf * ntf == nnz_ref


def test_sum_facets_closed_form():
    """test/SumFacetMapTests.jl:29,109 analogue: the trace-trace block is
    block-diagonal by local facet (facet-major lambda layout) and, on a square
    cell with all owner signs +1, its four facet blocks are equal (hcat(a,a,a,a))."""
    p = O.Problem.darcy(2, 2, mms=O.MMS_ZERO)
    mesh = one_box_cell(2, np.zeros(2), np.array([0.5, 0.5]), 1.0, np.ones(4))
    A, _ = O.cell_systems(p, mesh, np.ones(1))
    s = p.sizes()
    ni, nm = s["ni"], s["n_tf"]
    L = A[0, ni:, ni:]
    blocks = [L[f * nm:(f + 1) * nm, f * nm:(f + 1) * nm] for f in range(4)]
    for f in range(4):
        for g in range(4):
            if f != g:
                assert np.all(L[f * nm:(f + 1) * nm, g * nm:(g + 1) * nm] == 0.0)
        assert np.allclose(blocks[f], blocks[0], rtol=1e-15, atol=0)
    # the facet sum of tau w u over the four facets equals tau x the boundary mass
    Auu = A[0, s["n_flux"]:ni, s["n_flux"]:ni]
    ones = np.ones(Auu.shape[0])
    perimeter = 4 * 0.5
    assert abs(ones @ Auu @ ones - 1.0 * perimeter) < 1e-14


def test_elasticity_linear_displacement_exact():
    """Elasticity HDG: a linear displacement (constant stress) lies in the
    spaces, so the hybrid solution reproduces it to round-off; compliance
    identity A(invA(t)) = t (test/LinearElasticityHDGTests.jl:70-71)."""
    for D, n in ((2, 3), (3, 2)):
        p = O.Problem.elasticity(D, 1, tau=float(n), nu=0.3, mms=O.ELAS_MMS_LINEAR)
        sol = O.hdg_solve(p, (n,) * D, coef=np.full(n ** D, 0.3))
        err = O.l2_errors(p, sol.mesh, sol.coef, sol.xi)
        assert err[0] < 1e-11 and err[1] < 1e-11
        assert rel(sol.free_values, O.monolithic_solve(sol)) < 1e-10


def test_elasticity_convergence():
    errs = []
    for n in (4, 8):
        p = O.Problem.elasticity(2, 1, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN)
        sol = O.hdg_solve(p, (n, n), coef=np.full(n * n, 0.3))
        errs.append(O.l2_errors(p, sol.mesh, sol.coef, sol.xi))
    r = np.log2(np.array(errs[0]) / np.array(errs[1]))
    assert np.all(r > 0.8)


@pytest.mark.parametrize("name", ["darcy_q1_2d", "darcy_q2_2d", "darcy_q1_3d", "darcy_q2_3d",
                                  "darcy_q3_2d"])
def test_darcy_uq_block_integrates_by_parts(name):
    """On the exact cell systems the potential equation's flux block
    -(grad w, q) + <w, q.n> (test/DarcyHDGTests.jl:125-129) equals the negated
    transpose of the flux equation's potential block -(div v, p): the identity
    behind the integrated-by-parts u-q block and the oracle's self-variation
    variant (oracle.darcy_ibp_systems)."""
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    D = int(g["D"])
    kf, kp, kt = (int(v) for v in g["k"])
    prob = O.Problem(O.PHYS_DARCY, D, int(g["space"]), kf, kp, kt, 2 * (kf + 1), float(g["tau"]),
                     1.0, 0.3, O.MMS_KAT)
    s = prob.sizes()
    na, ni = s["n_flux"], s["ni"]
    A = g["A"]
    assert rel(A[na:ni, :na], -A[:na, na:ni].T) < 1e-15
    A2 = O.darcy_ibp_systems(A[None], na, ni)[0]
    assert rel(A2, A) < 1e-15


def test_condense_recover_stored_vs_lapack():
    """The CPU-baseline variant that keeps W = A_ii^-1 A_ib and w = A_ii^-1 b_i
    (x_i = w - W x_b) against dgetrf/dgetrs."""
    g = np.load(os.path.join(GOLD, "condense_lapack.npz"))
    S, gg, xi, info = O.condense_recover_stored(g["A"], g["b"], g["x"], 6, 3)
    assert not info.any()
    assert rel(S, g["S"]) < 1e-13 and rel(gg, g["g"]) < 1e-13 and rel(xi, g["xi"]) < 1e-13


def test_self_variation_reorder_is_exact_permutation():
    """solve_reordered maps its results back: on a well-conditioned system it
    agrees with the unpermuted solve to rounding."""
    g = np.load(os.path.join(GOLD, "condense_lapack.npz"))
    perm = np.r_[3:6, 0:3, 6:9]
    S, gg, xi, info = O.solve_reordered(g["A"], g["b"], g["x"], 6, 3, perm)
    assert not info.any()
    assert rel(S, g["S"]) < 1e-13 and rel(gg, g["g"]) < 1e-13 and rel(xi, g["xi"]) < 1e-13


@pytest.mark.parametrize("D,k,n", [(2, 1, 4), (2, 2, 3), (2, 3, 3), (3, 1, 2), (3, 2, 2)])
def test_oriented_mesh_reproduces_unrotated_solution(D, k, n):
    """Facet orientation (pindex, src/SkeletonArrays.jl:28-37): re-parametrising
    cells by rotations of the reference cell permutes their local facets and
    flips their facet frames, but not the discrete problem.  The skeleton
    solution (numbered in each facet's own frame) must be the same, and every
    cell's flux and potential must be the same nodal values in permuted
    order (the flux components are physical)."""
    import meshgen
    kappa = O.lognormal_kappa(n ** D, seed=31)
    prob = O.Problem.darcy(D, k, mms=O.MMS_SINPROD)
    x0, J, cf, nf, perm, rot = meshgen.sheared_rotated_mesh(D, n, seed=D * 10 + k, frac=0.6)
    x0r, Jr, cfr, _, _, _ = meshgen.sheared_rotated_mesh(D, n, seed=D * 10 + k, frac=0.0)
    assert any(r is not None for r in rot) and perm.any()
    ref = O.hdg_solve_mesh(prob, meshgen.oracle_mesh(x0r, Jr, cfr, nf, None), coef=kappa)
    got = O.hdg_solve_mesh(prob, meshgen.oracle_mesh(x0, J, cf, nf, perm), coef=kappa)
    assert got.nfree == ref.nfree and len(got.rowval) == len(ref.rowval)
    assert rel(got.lam, ref.lam) < 1e-11
    s = prob.sizes()
    na, ni = s["n_flux"], s["ni"]
    npn = na // D
    nodes = O.space_nodes(D, k, 0)
    for c in range(n ** D):
        idx = np.arange(npn) if rot[c] is None else meshgen.node_map(nodes, *rot[c])
        qi = np.concatenate([a * npn + idx for a in range(D)])
        assert rel(got.xi[c, :na], ref.xi[c, qi]) < 1e-10
        assert rel(got.xi[c, na:ni], ref.xi[c, na:ni][idx]) < 1e-10


def test_orientation_codes_on_trace_nodes():
    """Every orientation code maps Q_k trace nodes onto themselves (a
    permutation); P_2 trace nodes on a hexahedron's facet are not invariant
    under a flip, only under the swap of the facet coordinates."""
    for D, k in ((2, 1), (2, 3), (3, 1), (3, 2), (3, 3)):
        prob = O.Problem.darcy(D, k)
        for p in range(2 if D == 2 else 8):
            perm = O.facet_node_perm(prob, p)
            assert perm is not None and sorted(perm) == list(range(len(perm)))
            if p == 0:
                assert list(perm) == list(range(len(perm)))
    pe = O.Problem.elasticity(3, 2, tau=1.0)
    assert O.facet_node_perm(pe, 1) is not None
    assert O.facet_node_perm(pe, 2) is None


@pytest.mark.parametrize("D,k", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_projection_reproduces_trace_space_functions(D, k):
    """P_m is the L2 projection onto the trace space (test/P_m.jl:4-23; the
    reference checks (mu, u - P_m u)_F = 0 at test/LinearElasticityHDGTests.jl:
    124-130): a displacement whose facet traces lie in P_k is left unchanged,
    so on the interpolant of a linear field the projected and unprojected cell
    systems agree; on a field of degree k+1 they differ."""
    mesh = O.cartesian(D, (1,) * D)
    mesh.J[:] = (np.diag([0.5, 0.25, 0.125][:D])).T.ravel()
    A0, _ = O.cell_systems(O.Problem.elasticity(D, k, tau=2.0, mms=0), mesh, np.array([0.3]))
    A1, _ = O.cell_systems(O.Problem.elasticity(D, k, tau=2.0, mms=0, projection=1), mesh,
                           np.array([0.3]))
    s = O.Problem.elasticity(D, k, tau=2.0).sizes()
    na, ni, nq = s["n_flux"], s["ni"], s["n_pot"] // D
    nodes = O.space_nodes(D, k + 1, 1) * np.array([0.5, 0.25, 0.125][:D])
    dA = A1[0] - A0[0]
    assert np.max(np.abs(dA[:na])) == 0.0 and np.max(np.abs(dA[ni:])) == 0.0
    lin = np.concatenate([nodes @ np.arange(1, D + 1) + c for c in range(D)])
    assert np.max(np.abs(dA[:, na:ni] @ lin)) < 1e-14 * np.max(np.abs(A0[0]))
    top = np.concatenate([nodes[:, 0] ** (k + 1)] * D)
    assert np.max(np.abs(dA[:, na:ni] @ top)) > 1e-6


def test_projection_linear_displacement_exact():
    """With P_m the linear displacement (exact in the spaces) is still
    reproduced to round-off (test/LinearElasticityHDGTests.jl:70-71)."""
    for D, n in ((2, 3), (3, 2)):
        p = O.Problem.elasticity(D, 1, tau=float(n), nu=0.3, mms=O.ELAS_MMS_LINEAR, projection=1)
        sol = O.hdg_solve(p, (n,) * D, coef=np.full(n ** D, 0.3))
        err = O.l2_errors(p, sol.mesh, sol.coef, sol.xi)
        assert err[0] < 1e-11 and err[1] < 1e-11
        assert rel(sol.free_values, O.monolithic_solve(sol)) < 1e-10


def test_projection_convergence():
    """With P_m, the stress and displacement errors of the smooth solution
    fall at least at rate k+1 and k+1 (test/LinearElasticityHDGTests.jl:212-221
    reports the slopes of this convergence study)."""
    errs = []
    for n in (4, 8, 16):
        p = O.Problem.elasticity(2, 1, tau=float(n), nu=0.3, mms=O.ELAS_MMS_SIN, projection=1)
        sol = O.hdg_solve(p, (n, n), coef=np.full(n * n, 0.3))
        errs.append(O.l2_errors(p, sol.mesh, sol.coef, sol.xi))
    r = np.log2(np.array(errs[1]) / np.array(errs[2]))
    print("P_m convergence rates (stress, displacement):", r)
    assert r[0] > 1.7 and r[1] > 1.7
