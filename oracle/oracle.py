"""ctypes front end of the CPU oracle (oracle/hdg_oracle.c).

TEST INFRASTRUCTURE ONLY.  This module is the checker: tests/, the smoke
check and bench.py's CPU-baseline leg import it; the product (libhdg and the
host package) never does.

It composes the oracle's C restatement into the reference's end-to-end flow:
HybridAffineFEOperator (src/HybridAffineFEOperators.jl:10-50: cell systems,
StaticCondensationMap, Dirichlet lifting, skeleton assembly), the host sparse
skeleton solve (:74) and solve!'s recovery (:72-150).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SPACE_Q, SPACE_P = 0, 1
MMS_ZERO, MMS_KAT, MMS_LINEAR, MMS_SINPROD = 0, 1, 2, 3
ELAS_MMS_LINEAR, ELAS_MMS_SIN = 1, 2
PHYS_DARCY, PHYS_ELASTICITY = 0, 1


class _Problem(C.Structure):
    _fields_ = [
        ("phys", C.c_int), ("D", C.c_int), ("space", C.c_int), ("k_flux", C.c_int),
        ("k_pot", C.c_int), ("k_trace", C.c_int), ("quad_degree", C.c_int),
        ("tau", C.c_double), ("E", C.c_double), ("nu", C.c_double), ("mms", C.c_int),
        ("projection", C.c_int),
    ]


def build():
    """Compile the oracle library (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libhdg_oracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _LIB.orc_assemble_nnz.restype = C.c_int64
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


@dataclass
class Problem:
    phys: int = PHYS_DARCY
    D: int = 2
    space: int = SPACE_Q
    k_flux: int = 1
    k_pot: int = 1
    k_trace: int = 1
    quad_degree: int = 4
    tau: float = 1.0
    E: float = 1.0
    nu: float = 0.3
    mms: int = MMS_LINEAR
    projection: int = 0   # elasticity: bulk-to-skeleton L2 projection P_m

    @staticmethod
    def darcy(D, k, space=SPACE_Q, mms=MMS_LINEAR, tau=1.0, k_pot=None, k_trace=None,
              quad_degree=None):
        """BASELINE spaces: Q_k flux, Q_k potential, Q_k trace, degree 2(k+1)."""
        return Problem(PHYS_DARCY, D, space, k, k if k_pot is None else k_pot,
                       k if k_trace is None else k_trace,
                       2 * (k + 1) if quad_degree is None else quad_degree, tau, 1.0, 0.3, mms)

    @staticmethod
    def darcy_reference_kat(D=2):
        """test/DarcyHDGTests.jl:34-37: P1 flux, P0 potential, P1 trace, tau=1,
        degree 2(order+1) = 4."""
        return Problem(PHYS_DARCY, D, SPACE_P, 1, 0, 1, 4, 1.0, 1.0, 0.3, MMS_KAT)

    @staticmethod
    def elasticity(D, k, tau, nu=0.3, E=1.0, mms=ELAS_MMS_SIN, projection=0):
        """test/LinearElasticityHDGTests.jl:88-94: P_k sym stress, P_{k+1}
        displacement, P_k trace, degree 2(k+1); projection = the
        bulk_to_skeleton_projection flag (P_m, :124-150)."""
        return Problem(PHYS_ELASTICITY, D, SPACE_P, k, k + 1, k, 2 * (k + 1), tau, E, nu, mms,
                       int(projection))

    def c(self):
        return _Problem(self.phys, self.D, self.space, self.k_flux, self.k_pot, self.k_trace,
                        self.quad_degree, self.tau, self.E, self.nu, self.mms, self.projection)

    def sizes(self):
        v = [C.c_int() for _ in range(5)]
        lib().orc_sizes(C.byref(self.c()), *[C.byref(x) for x in v])
        nflux, npot, ntf, ni, nb = [x.value for x in v]
        return dict(n_flux=nflux, n_pot=npot, n_tf=ntf, ni=ni, nb=nb)

    def quad_sizes(self):
        a, b = C.c_int(), C.c_int()
        lib().orc_quad_sizes(C.byref(self.c()), C.byref(a), C.byref(b))
        return a.value, b.value


@dataclass
class Mesh:
    D: int
    n: tuple
    domain: tuple
    ncells: int
    nfacets: int
    cell_facets: np.ndarray
    cell_sign: np.ndarray
    facet_cells: np.ndarray
    facet_lpos: np.ndarray
    facet_boundary: np.ndarray
    x0: np.ndarray
    J: np.ndarray
    facet_perm: np.ndarray = None   # [ncells*2D] orientation codes (None = identity)


def general_mesh(D, x0, J, cell_facets, nfacets, facet_perm=None):
    """Topology of a conforming mesh of affine cells given by its cell->facet
    ids in local facet order: facet -> (cell, local position) in ascending cell
    order, owner = first (lowest id) cell around a facet
    (src/Skeleton.jl:135-140), owner signs, boundary facets."""
    nc = len(x0) // D
    nF = 2 * D
    cf = np.ascontiguousarray(cell_facets, np.int32)
    fc = -np.ones(nfacets * 2, np.int32)
    fl = -np.ones(nfacets * 2, np.int8)
    cs = np.ones(nc * nF, np.int8)
    for c in range(nc):
        for lf in range(nF):
            f = cf[c * nF + lf]
            if fc[2 * f] < 0:
                fc[2 * f], fl[2 * f] = c, lf
            else:
                assert fc[2 * f + 1] < 0, "facet shared by more than two cells"
                fc[2 * f + 1], fl[2 * f + 1] = c, lf
                cs[c * nF + lf] = -1
    fb = (fc[1::2] < 0).astype(np.int8)
    fp = None if facet_perm is None else np.ascontiguousarray(facet_perm, np.int8)
    return Mesh(D, (nc,), (), nc, nfacets, cf, cs, fc, fl, fb,
                np.ascontiguousarray(x0, np.float64), np.ascontiguousarray(J, np.float64), fp)


def facet_node_perm(prob: Problem, code):
    """perm[i] = the facet's own index of the cell-local trace node i under
    orientation code `code` (None if the node set is not invariant)."""
    nm = prob.sizes()["n_tf"] // (1 if prob.phys == PHYS_DARCY else prob.D)
    perm = np.zeros(nm, np.int32)
    rc = lib().orc_facet_node_perm(C.byref(prob.c()), int(code), _p(perm))
    return None if rc else perm


def space_nodes(d, k, pspace):
    """Reference nodes (n, d) of the scalar Lagrange space of order k."""
    buf = np.zeros(4096 * max(d, 1))
    n = lib().orc_space_nodes(int(d), int(k), int(pspace), _p(buf))
    assert n > 0
    return buf[:n * d].reshape(n, d)


def trace_dofs_oriented(prob: Problem, mesh: Mesh, facet_dirichlet=None):
    """trace_dofs honouring mesh.facet_perm (cell-local trace nodes mapped to
    each facet's own node order)."""
    if facet_dirichlet is None:
        facet_dirichlet = mesh.facet_boundary.copy()
    fd = np.ascontiguousarray(facet_dirichlet, np.int8)
    nb = prob.sizes()["nb"]
    cd = np.zeros(mesh.ncells * nb, np.int32)
    nf, nd = C.c_int64(), C.c_int64()
    rc = lib().orc_trace_dofs_oriented(C.byref(prob.c()), C.c_int64(mesh.ncells),
                                       C.c_int64(mesh.nfacets), _p(mesh.cell_facets),
                                       _p(mesh.facet_perm), _p(fd), _p(cd), C.byref(nf),
                                       C.byref(nd))
    assert rc == 0, "orientation not supported by the trace space"
    return cd.reshape(mesh.ncells, nb), nf.value, nd.value, fd


def dirichlet_values_oriented(prob: Problem, mesh: Mesh, facet_dirichlet, ndir):
    out = np.zeros(max(ndir, 1))
    rc = lib().orc_dirichlet_values_oriented(
        C.byref(prob.c()), C.c_int64(mesh.nfacets), _p(mesh.facet_cells), _p(mesh.facet_lpos),
        _p(mesh.facet_perm), _p(np.ascontiguousarray(facet_dirichlet, np.int8)), _p(mesh.x0),
        _p(mesh.J), _p(out))
    assert rc == 0
    return out[:ndir]


def cartesian(D, n, domain=None):
    """CartesianDiscreteModel(domain, n) topology and affine geometry."""
    n = tuple(int(v) for v in n)
    if domain is None:
        domain = tuple([0.0, 1.0] * D)
    L = lib()
    na = (C.c_int64 * 3)(*(list(n) + [1] * (3 - D)))
    nc, nf = C.c_int64(), C.c_int64()
    L.orc_cartesian_counts(D, na, C.byref(nc), C.byref(nf))
    nc, nf = nc.value, nf.value
    nF = 2 * D
    cf = np.zeros(nc * nF, np.int32)
    cs = np.zeros(nc * nF, np.int8)
    fc = np.zeros(nf * 2, np.int32)
    fl = np.zeros(nf * 2, np.int8)
    fb = np.zeros(nf, np.int8)
    rc = L.orc_cartesian_topology(D, na, _p(cf), _p(cs), _p(fc), _p(fl), _p(fb))
    assert rc == 0
    x0 = np.zeros(nc * D)
    J = np.zeros(nc * D * D)
    dom = np.asarray(domain, np.float64)
    L.orc_cartesian_geometry(D, na, _p(dom), _p(x0), _p(J))
    return Mesh(D, n, tuple(domain), nc, nf, cf, cs, fc, fl, fb, x0, J)


def cell_systems(prob: Problem, mesh: Mesh, coef, src=None, cells=None):
    """Dense cell systems (A, b), column-major per cell."""
    s = prob.sizes()
    n = s["ni"] + s["nb"]
    D = mesh.D
    if cells is None:
        x0, J, sg, cf = mesh.x0, mesh.J, mesh.cell_sign, np.ascontiguousarray(coef, np.float64)
        nc = mesh.ncells
    else:
        cells = np.asarray(cells)
        x0 = np.ascontiguousarray(mesh.x0.reshape(-1, D)[cells]).ravel()
        J = np.ascontiguousarray(mesh.J.reshape(-1, D * D)[cells]).ravel()
        sg = np.ascontiguousarray(mesh.cell_sign.reshape(-1, 2 * D)[cells]).ravel()
        cf = np.ascontiguousarray(np.asarray(coef, np.float64)[cells])
        nc = len(cells)
    A = np.zeros((nc, n, n))
    b = np.zeros((nc, n))
    srcp = None if src is None else np.ascontiguousarray(src, np.float64)
    rc = lib().orc_cell_systems(C.byref(prob.c()), C.c_int64(nc), _p(x0), _p(J), _p(cf), _p(sg),
                                _p(srcp), _p(A), _p(b))
    assert rc == 0
    # A[c] is stored column-major: element (i, j) at flat i + j*n -> transpose view
    return A.transpose(0, 2, 1), b


def to_colmajor(A):
    """(nc, n, m) logical -> contiguous buffer in [cell][col][row] order."""
    return np.ascontiguousarray(A.transpose(0, 2, 1))


def condense(A, b, ni, nb):
    nc = A.shape[0]
    Ac = to_colmajor(A)
    bc = np.ascontiguousarray(b)
    S = np.zeros((nc, nb, nb))
    g = np.zeros((nc, nb))
    info = np.zeros(nc, np.int32)
    lib().orc_condense(ni, nb, C.c_int64(nc), _p(Ac), _p(bc), _p(S), _p(g), _p(info))
    return S.transpose(0, 2, 1), g, info


def recover(A, b, xb, ni, nb):
    nc = A.shape[0]
    Ac = to_colmajor(A)
    xi = np.zeros((nc, ni))
    info = np.zeros(nc, np.int32)
    lib().orc_recover(ni, nb, C.c_int64(nc), _p(Ac), _p(np.ascontiguousarray(b)),
                      _p(np.ascontiguousarray(xb, np.float64)), _p(xi), _p(info))
    return xi, info


def trace_dofs(mesh: Mesh, n_tf, facet_dirichlet=None):
    if facet_dirichlet is None:
        facet_dirichlet = mesh.facet_boundary.copy()
    fd = np.ascontiguousarray(facet_dirichlet, np.int8)
    nb = 2 * mesh.D * n_tf
    cd = np.zeros(mesh.ncells * nb, np.int32)
    nf, nd = C.c_int64(), C.c_int64()
    lib().orc_trace_dofs(mesh.D, n_tf, C.c_int64(mesh.ncells), C.c_int64(mesh.nfacets),
                         _p(mesh.cell_facets), _p(fd), _p(cd), C.byref(nf), C.byref(nd))
    return cd.reshape(mesh.ncells, nb), nf.value, nd.value, fd


def dirichlet_values(prob: Problem, mesh: Mesh, facet_dirichlet, ndir):
    out = np.zeros(max(ndir, 1))
    lib().orc_dirichlet_values(C.byref(prob.c()), C.c_int64(mesh.nfacets), _p(mesh.facet_cells),
                               _p(mesh.facet_lpos), _p(np.ascontiguousarray(facet_dirichlet, np.int8)),
                               _p(mesh.x0), _p(mesh.J), _p(out))
    return out[:ndir]


def assemble(S, g, cell_dofs, nfree, lam_dir):
    """Returns (colptr, rowval, nzval, rhs) of the CSC skeleton system."""
    import scipy.sparse as sp  # noqa: F401  (callers build csc from these)
    nc, nb, _ = S.shape
    cd = np.ascontiguousarray(cell_dofs, np.int32)
    nnz = lib().orc_assemble_nnz(nb, C.c_int64(nc), C.c_int64(nfree), _p(cd))
    colptr = np.zeros(nfree + 1, np.int64)
    rowval = np.zeros(max(nnz, 1), np.int32)
    nzval = np.zeros(max(nnz, 1))
    rhs = np.zeros(max(nfree, 1))
    ld = np.ascontiguousarray(lam_dir if len(lam_dir) else np.zeros(1), np.float64)
    lib().orc_assemble(nb, C.c_int64(nc), C.c_int64(nfree), _p(cd), _p(to_colmajor(S)),
                       _p(np.ascontiguousarray(g)), _p(ld), _p(colptr), _p(rowval), _p(nzval),
                       _p(rhs))
    return colptr, rowval[:nnz], nzval[:nnz], rhs[:nfree]


def gather_lambda(cell_dofs, lam_free, lam_dir):
    cd = np.ascontiguousarray(cell_dofs, np.int32)
    nc, nb = cd.shape
    xb = np.zeros((nc, nb))
    lf = np.ascontiguousarray(lam_free if len(lam_free) else np.zeros(1), np.float64)
    ld = np.ascontiguousarray(lam_dir if len(lam_dir) else np.zeros(1), np.float64)
    lib().orc_gather_lambda(nb, C.c_int64(nc), _p(cd), _p(lf), _p(ld), _p(xb))
    return xb


def l2_errors(prob: Problem, mesh: Mesh, coef, xi):
    err = np.zeros(2)
    lib().orc_l2_errors(C.byref(prob.c()), C.c_int64(mesh.ncells), _p(mesh.x0), _p(mesh.J),
                        _p(np.ascontiguousarray(coef, np.float64)),
                        _p(np.ascontiguousarray(xi, np.float64)), _p(err))
    return err


def dense_solve(A, B):
    A = np.asfortranarray(A, np.float64).copy(order="F")
    B = np.asfortranarray(B, np.float64).copy(order="F")
    n = A.shape[0]
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    info = lib().orc_dense_solve(n, nrhs, _p(A), _p(B))
    return B, info


def lognormal_kappa(ncells, seed, sigma=1.0):
    """kappa_c = exp(N(0, sigma^2)) drawn in cell order from the seed."""
    rng = np.random.default_rng(seed)
    return np.exp(sigma * rng.standard_normal(ncells))


@dataclass
class Solution:
    prob: Problem
    mesh: Mesh
    coef: np.ndarray
    A: np.ndarray
    b: np.ndarray
    S: np.ndarray
    g: np.ndarray
    cell_dofs: np.ndarray
    nfree: int
    ndir: int
    lam_dir: np.ndarray
    colptr: np.ndarray
    rowval: np.ndarray
    nzval: np.ndarray
    rhs: np.ndarray
    lam: np.ndarray
    xb: np.ndarray
    xi: np.ndarray
    free_values: np.ndarray
    extra: dict = field(default_factory=dict)


def hdg_solve(prob: Problem, n, coef=None, domain=None, lam=None):
    """Full reference flow: condense -> assemble -> host sparse solve -> recover.
    If lam is given (free trace values), the skeleton solve is skipped."""
    return hdg_solve_mesh(prob, cartesian(prob.D, n, domain), coef, lam)


def hdg_solve_mesh(prob: Problem, mesh: Mesh, coef=None, lam=None):
    """hdg_solve on any mesh of affine cells (general_mesh or cartesian),
    honouring its facet orientation codes."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    s = prob.sizes()
    ni, nb = s["ni"], s["nb"]
    if coef is None:
        coef = np.full(mesh.ncells, 1.0 if prob.phys == PHYS_DARCY else prob.nu)
    A, b = cell_systems(prob, mesh, coef)
    S, g, info = condense(A, b, ni, nb)
    assert not info.any(), "singular A_ii"
    if mesh.facet_perm is None:
        cd, nfree, ndir, fd = trace_dofs(mesh, s["n_tf"])
        lam_dir = dirichlet_values(prob, mesh, fd, ndir)
    else:
        cd, nfree, ndir, fd = trace_dofs_oriented(prob, mesh)
        lam_dir = dirichlet_values_oriented(prob, mesh, fd, ndir)
    colptr, rowval, nzval, rhs = assemble(S, g, cd, nfree, lam_dir)
    if lam is None:
        K = sp.csc_matrix((nzval, rowval, colptr), shape=(nfree, nfree))
        lam = spla.spsolve(K, rhs) if nfree else np.zeros(0)
    xb = gather_lambda(cd, lam, lam_dir)
    xi, info = recover(A, b, xb, ni, nb)
    assert not info.any()
    nf = s["n_flux"]
    free = np.concatenate([xi[:, :nf].ravel(), xi[:, nf:].ravel(), lam])
    return Solution(prob, mesh, np.asarray(coef), A, b, S, g, cd, nfree, ndir, lam_dir, colptr,
                    rowval, nzval, rhs, lam, xb, xi, free)


def monolithic_solve(sol: Solution):
    """Assemble the full (uncondensed) hybridizable system with the global
    numbering [q of all cells; u of all cells; free lambda] and solve it with a
    sparse direct solver.  Independent of the condensation path."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    s = sol.prob.sizes()
    ni, nb, nf, npot = s["ni"], s["nb"], s["n_flux"], s["n_pot"]
    nc = sol.mesh.ncells
    N = nc * ni + sol.nfree
    gid = np.zeros((nc, ni + nb), np.int64)
    for c in range(nc):
        gid[c, :nf] = c * nf + np.arange(nf)
        gid[c, nf:ni] = nc * nf + c * npot + np.arange(npot)
        d = sol.cell_dofs[c]
        gid[c, ni:] = np.where(d >= 0, nc * ni + d, -1)
    rows, cols, vals = [], [], []
    rhs = np.zeros(N)
    for c in range(nc):
        Ac, bc, g = sol.A[c], sol.b[c].copy(), gid[c]
        d = sol.cell_dofs[c]
        xd = np.zeros(ni + nb)
        dm = d < 0
        xd[ni:][dm] = sol.lam_dir[-d[dm] - 1]
        bc = bc - Ac @ xd
        keep = g >= 0
        ii, jj = np.meshgrid(np.nonzero(keep)[0], np.nonzero(keep)[0], indexing="ij")
        rows.append(g[ii].ravel())
        cols.append(g[jj].ravel())
        vals.append(Ac[ii, jj].ravel())
        np.add.at(rhs, g[keep], bc[keep])
    K = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    return spla.spsolve(K, rhs)


# ----------------------------------------------------------------------------
# Self-variation: how far the oracle's own FP64 answer moves under changes that
# are exact in real arithmetic.  Two independent FP64 implementations of the
# same cell systems cannot agree more closely than this.
# ----------------------------------------------------------------------------
def darcy_ibp_systems(A, n_flux, ni):
    """The same Darcy cell systems with the u-q block in integrated-by-parts
    form.  test/DarcyHDGTests.jl:125-129 writes the potential equation's flux
    terms as -(grad w, q) + <w, q.n>; on an affine cell that equals (w, div q),
    the negated transpose of the q-u block -(div v, p) (both integrated exactly
    by the degree 2(k+1) rules).  The two differ only by rounding."""
    A2 = np.array(A, copy=True)
    A2[:, n_flux:ni, :n_flux] = -np.transpose(A[:, :n_flux, n_flux:ni], (0, 2, 1))
    return A2


def interior_reorder(ni, n_flux, nb):
    """Permutation of the cell unknowns that puts the potential block before
    the flux block (boundary unknowns unchanged): getrf then pivots in another
    order."""
    return np.concatenate([np.arange(n_flux, ni), np.arange(n_flux), np.arange(ni, ni + nb)])


def solve_reordered(A, b, xb, ni, nb, perm):
    """condense + recover with the unknowns permuted by perm, results mapped
    back to the natural order."""
    Ap = A[:, perm][:, :, perm]
    bp = b[:, perm]
    S, g, info = condense(Ap, bp, ni, nb)
    xp, info2 = recover(Ap, bp, xb, ni, nb)
    xi = np.empty_like(xp)
    xi[:, perm[:ni]] = xp
    return S, g, xi, info | info2


def refined_interior(A, b, xb, ni):
    """x_i = A_ii^-1 (b_i - A_ib x_b) of the given FP64 systems, refined with
    extended-precision residuals until it is the correctly rounded solution up
    to a few ulps: the reference point that separates a solver's own rounding
    from the sensitivity of the cell systems."""
    L = np.longdouble
    Aii = np.ascontiguousarray(A[:, :ni, :ni])
    rhsL = b[:, :ni].astype(L) - np.einsum("cij,cj->ci", A[:, :ni, ni:].astype(L), xb.astype(L))
    x = np.linalg.solve(Aii, rhsL.astype(np.float64)[..., None])[..., 0]
    AL = Aii.astype(L)
    for _ in range(3):
        r = rhsL - np.einsum("cij,cj->ci", AL, x.astype(L))
        d = np.linalg.solve(Aii, r.astype(np.float64)[..., None])[..., 0]
        x = (x.astype(L) + d.astype(L)).astype(np.float64)
    return x


def rel_err(a, b):
    """max |a - b| / max |b| (0 for empty arrays)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def darcy_self_variation(A, b, xb, ni, nb, n_flux):
    """The oracle's answers on Darcy cell systems (S, g, q, u) and how far they
    move (relative max norm, per field) under (1) the integrated-by-parts u-q
    block and (2) the reordered interior unknowns.  Returns (base, variation)
    with base = dict(S, g, xi) and variation = dict(S, g, q, u)."""
    S0, g0, info = condense(A, b, ni, nb)
    x0, info2 = recover(A, b, xb, ni, nb)
    assert not (info.any() or info2.any()), "singular A_ii"
    base = dict(S=S0, g=g0, xi=x0)
    A2 = darcy_ibp_systems(A, n_flux, ni)
    S1, g1, _ = condense(A2, b, ni, nb)
    x1, _ = recover(A2, b, xb, ni, nb)
    S2, g2, x2, _ = solve_reordered(A, b, xb, ni, nb, interior_reorder(ni, n_flux, nb))
    var = {}
    for key, ref, alts in (("S", S0, (S1, S2)), ("g", g0, (g1, g2)),
                           ("q", x0[:, :n_flux], (x1[:, :n_flux], x2[:, :n_flux])),
                           ("u", x0[:, n_flux:], (x1[:, n_flux:], x2[:, n_flux:]))):
        var[key] = max(rel_err(a, ref) for a in alts)
    return base, var


def set_threads(n):
    """OpenMP threads of the oracle's following calls; returns the previous count."""
    return lib().orc_set_threads(int(n))


def condense_recover_stored(A, b, xb, ni, nb):
    """Condensation keeping W = A_ii^-1 A_ib and w = A_ii^-1 b_i, then x_i = w - W x_b
    (no re-factorisation in the recovery)."""
    nc = A.shape[0]
    S = np.zeros((nc, nb, nb))
    g = np.zeros((nc, nb))
    xi = np.zeros((nc, ni))
    info = np.zeros(nc, np.int32)
    lib().orc_condense_recover_stored(ni, nb, C.c_int64(nc), _p(to_colmajor(A)),
                                      _p(np.ascontiguousarray(b)),
                                      _p(np.ascontiguousarray(xb, np.float64)), _p(S), _p(g),
                                      _p(xi), _p(info))
    return S.transpose(0, 2, 1), g, xi, info
