"""Writes the golden fixtures under tests/golden/ from independent exact or
library computations (never from the oracle):

* exact rational integration (fractions.Fraction polynomials) of the Darcy HDG
  weak form of test/DarcyHDGTests.jl:125-135 and of the elasticity HDG weak
  form of test/LinearElasticityHDGTests.jl:140-150 on single box cells, with
  nodal Lagrange bases built by exact Vandermonde inversion;
* LAPACK (scipy.linalg.lu_factor / lu_solve, i.e. dgetrf / dgetrs) static
  condensation and recovery of seeded block systems with the block pattern of
  test/StaticCondensationMapTests.jl:6-23.

Conventions restated here (not derived from the oracle): Gridap node order on
n-cubes (vertices, edges, faces, interior; by free-direction set, fixed
coordinates first-direction fastest), local facets fixing direction D-1-f//2
at f%2, facet coordinates = the free directions ascending, field order
(flux/stress, potential/displacement, trace), component-major vector dofs.

Run:  python tests/golden/make_golden.py
"""
import itertools
import os
from fractions import Fraction as Fr

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# --------------------------------------------------------------------------
# exact polynomials in up to 3 variables: dict {(e0,e1,e2): Fraction}
# --------------------------------------------------------------------------
def padd(p, q, s=1):
    r = dict(p)
    for k, v in q.items():
        r[k] = r.get(k, 0) + s * v
        if r[k] == 0:
            del r[k]
    return r


def pmul(p, q):
    r = {}
    for k1, v1 in p.items():
        for k2, v2 in q.items():
            k = (k1[0] + k2[0], k1[1] + k2[1], k1[2] + k2[2])
            r[k] = r.get(k, 0) + v1 * v2
    return {k: v for k, v in r.items() if v != 0}


def pscale(p, s):
    return {k: v * s for k, v in p.items() if v * s != 0}


def pderiv(p, a):
    r = {}
    for k, v in p.items():
        if k[a] > 0:
            kk = list(k)
            kk[a] -= 1
            r[tuple(kk)] = r.get(tuple(kk), 0) + v * k[a]
    return r


def peval(p, x):
    s = Fr(0)
    for k, v in p.items():
        t = v
        for a in range(3):
            if k[a]:
                t *= Fr(x[a]) ** k[a]
        s += t
    return s


def pint_cube(p, D):
    """integral over [0,1]^D"""
    s = Fr(0)
    for k, v in p.items():
        t = v
        for a in range(D):
            t /= (k[a] + 1)
        s += t
    return s


def psubst(p, a, val):
    """fix variable a at val (0 or 1)"""
    r = {}
    for k, v in p.items():
        kk = list(k)
        if kk[a] and val == 0:
            continue
        kk[a] = 0
        r[tuple(kk)] = r.get(tuple(kk), 0) + v
    return {k: v for k, v in r.items() if v != 0}


def prelabel(p, mapping):
    """rename variables: new index mapping[old]"""
    r = {}
    for k, v in p.items():
        kk = [0, 0, 0]
        for a in range(3):
            if k[a]:
                kk[mapping[a]] += k[a]
        r[tuple(kk)] = r.get(tuple(kk), 0) + v
    return r


# --------------------------------------------------------------------------
# nodal Lagrange bases on the reference d-cube
# --------------------------------------------------------------------------
def cube_lattice(d, k, pspace):
    if d == 0 or k == 0:
        return [tuple([Fr(1, 2)] * d)] if k == 0 and d > 0 else [()]
    out = []
    for fd in range(d + 1):
        for free in itertools.combinations(range(d), fd):
            fixed = [a for a in range(d) if a not in free]
            for fv in range(1 << len(fixed)):
                nown = (k - 1) ** fd
                for o in range(nown):
                    a = [0] * d
                    t = o
                    for f in free:
                        a[f] = 1 + t % (k - 1)
                        t //= (k - 1)
                    for i, f in enumerate(fixed):
                        a[f] = k if (fv >> i) & 1 else 0
                    if pspace and sum(a) > k:
                        continue
                    out.append(tuple(Fr(v, k) for v in a))
    return out


def monomials(d, k, pspace):
    ms = []
    for e in itertools.product(range(k + 1), repeat=d):
        if (sum(e) if pspace else max(e, default=0)) <= k:
            ms.append(tuple(list(e) + [0] * (3 - d)))
    return ms


def nodal_basis(d, k, pspace):
    nodes = cube_lattice(d, k, pspace)
    if k == 0:
        return [{(0, 0, 0): Fr(1)}], nodes
    ms = monomials(d, k, pspace)
    n = len(nodes)
    assert n == len(ms)
    V = [[Fr(1) for _ in range(n)] for _ in range(n)]
    for i, x in enumerate(nodes):
        for m, e in enumerate(ms):
            t = Fr(1)
            for a in range(d):
                t *= x[a] ** e[a]
            V[i][m] = t
    # invert V exactly (Gauss-Jordan)
    A = [row[:] + [Fr(int(i == j)) for j in range(n)] for i, row in enumerate(V)]
    for j in range(n):
        p = next(i for i in range(j, n) if A[i][j] != 0)
        A[j], A[p] = A[p], A[j]
        piv = A[j][j]
        A[j] = [v / piv for v in A[j]]
        for i in range(n):
            if i != j and A[i][j] != 0:
                f = A[i][j]
                A[i] = [vi - f * vj for vi, vj in zip(A[i], A[j])]
    Vinv = [row[n:] for row in A]
    basis = []
    for i in range(n):
        basis.append({ms[m]: Vinv[m][i] for m in range(n) if Vinv[m][i] != 0})
    return basis, nodes


# --------------------------------------------------------------------------
# cell systems on a box cell [x0, x0+h]: integrate on the reference cell
# --------------------------------------------------------------------------
def facet_info(D, f):
    fix = D - 1 - f // 2
    side = f % 2
    free = [a for a in range(D) if a != fix]
    return fix, side, free


def restrict_to_facet(p, D, f):
    """cell polynomial -> facet polynomial in the facet coordinates"""
    fix, side, free = facet_info(D, f)
    q = psubst(p, fix, side)
    mapping = {a: 0 for a in range(3)}
    for j, a in enumerate(free):
        mapping[a] = j
    return prelabel(q, mapping)


def darcy_cell_exact(D, kf, kp, kt, pspace, h, kappa, tau, signs, src):
    """Exact A (n x n) and b of the Darcy HDG weak form on a box cell.
    src(x) polynomial sources are given as (f list of polys in physical x, g poly)
    over the reference coordinates already (we pass them reference-mapped)."""
    phi, _ = nodal_basis(D, kf, pspace)
    psi, _ = nodal_basis(D, kp, pspace)
    mu, _ = nodal_basis(D - 1, kt, pspace)
    npf, nps, nm = len(phi), len(psi), len(mu)
    na = D * npf
    ni = na + nps
    nF = 2 * D
    n = ni + nF * nm
    detJ = Fr(1)
    for a in range(D):
        detJ *= h[a]
    A = [[Fr(0)] * n for _ in range(n)]
    b = [Fr(0)] * n
    # bulk
    for a in range(D):
        for i in range(npf):
            for j in range(npf):
                A[a * npf + i][a * npf + j] += pint_cube(pmul(phi[i], phi[j]), D) * detJ / kappa
    for a in range(D):
        for i in range(npf):
            dphi = pscale(pderiv(phi[i], a), 1 / h[a])
            for j in range(nps):
                A[a * npf + i][na + j] += -pint_cube(pmul(dphi, psi[j]), D) * detJ
    for a in range(D):
        for i in range(nps):
            dpsi = pscale(pderiv(psi[i], a), 1 / h[a])
            for j in range(npf):
                A[na + i][a * npf + j] += -pint_cube(pmul(dpsi, phi[j]), D) * detJ
    f_polys, g_poly = src
    for a in range(D):
        for i in range(npf):
            b[a * npf + i] += pint_cube(pmul(phi[i], f_polys[a]), D) * detJ
    for i in range(nps):
        b[na + i] += pint_cube(pmul(psi[i], g_poly), D) * detJ
    # facets
    for f in range(nF):
        fix, side, free = facet_info(D, f)
        area = detJ / h[fix]
        nrm = [Fr(0)] * D
        nrm[fix] = Fr(1) if side else Fr(-1)
        s = signs[f]
        phF = [restrict_to_facet(p, D, f) for p in phi]
        psF = [restrict_to_facet(p, D, f) for p in psi]
        lo = ni + f * nm
        I = lambda p: pint_cube(p, D - 1) * area  # noqa: E731
        for m in range(nm):
            for a in range(D):
                for i in range(npf):
                    A[a * npf + i][lo + m] += I(pmul(phF[i], mu[m])) * nrm[a]
                    A[lo + m][a * npf + i] += I(pmul(mu[m], phF[i])) * nrm[a]
            for i in range(nps):
                A[na + i][lo + m] += -tau * s * I(pmul(psF[i], mu[m]))
                A[lo + m][na + i] += tau * s * I(pmul(mu[m], psF[i]))
            for m2 in range(nm):
                A[lo + m][lo + m2] += -tau * s * I(pmul(mu[m], mu[m2]))
        for a in range(D):
            for i in range(nps):
                for j in range(npf):
                    A[na + i][a * npf + j] += I(pmul(psF[i], phF[j])) * nrm[a]
        for i in range(nps):
            for j in range(nps):
                A[na + i][na + j] += tau * s * I(pmul(psF[i], psF[j]))
    return (np.array([[float(v) for v in row] for row in A]), np.array([float(v) for v in b]))


def sym_ij(D, c):
    t = 0
    for a in range(D):
        for bb in range(a, D):
            if t == c:
                return a, bb
            t += 1


def frac_solve(A, B):
    """Exact A^-1 B (Fraction Gauss-Jordan); A n x n, B n x m (lists of rows)."""
    n, m = len(A), len(B[0])
    M = [list(A[i]) + list(B[i]) for i in range(n)]
    for j in range(n):
        p = next(i for i in range(j, n) if M[i][j] != 0)
        M[j], M[p] = M[p], M[j]
        piv = M[j][j]
        M[j] = [v / piv for v in M[j]]
        for i in range(n):
            if i != j and M[i][j] != 0:
                f = M[i][j]
                M[i] = [vi - f * vj for vi, vj in zip(M[i], M[j])]
    return [row[n:n + m] for row in M]


def elasticity_cell_exact(D, k, h, E, nu, tau, projection=False):
    """Exact A of the elasticity HDG form on a box cell, zero load; with
    projection, the tau (omega, u) facet block uses the bulk-to-skeleton L2
    projection P_m of test/P_m.jl:4-23: tau (omega, P_m u)_F = tau B^T A^-1 B
    with A = (mu, mu')_F, B = (mu, u)_F (exact rational inverse)."""
    phi, _ = nodal_basis(D, k, True)       # stress scalar P_k
    psi, _ = nodal_basis(D, k + 1, True)   # displacement scalar P_{k+1}
    mu, _ = nodal_basis(D - 1, k, True)    # trace scalar P_k
    npf, nps, nm = len(phi), len(psi), len(mu)
    nsym = D * (D + 1) // 2
    na, nu_ = nsym * npf, D * nps
    ni = na + nu_
    nF = 2 * D
    ntf = D * nm
    n = ni + nF * ntf
    detJ = Fr(1)
    for a in range(D):
        detJ *= h[a]
    c1, c2 = (1 + nu) / E, nu / E
    Et = []
    for c in range(nsym):
        i, j = sym_ij(D, c)
        M = [[Fr(0)] * D for _ in range(D)]
        M[i][j] = Fr(1)
        M[j][i] = Fr(1)
        Et.append(M)
    A = [[Fr(0)] * n for _ in range(n)]
    for ca in range(nsym):
        for cb in range(nsym):
            dd = sum(Et[ca][x][y] * Et[cb][x][y] for x in range(D) for y in range(D))
            tra = sum(Et[ca][x][x] for x in range(D))
            trb = sum(Et[cb][x][x] for x in range(D))
            kab = c1 * dd - c2 * tra * trb
            if kab == 0:
                continue
            for i in range(npf):
                for j in range(npf):
                    A[ca * npf + i][cb * npf + j] += kab * pint_cube(pmul(phi[i], phi[j]), D) * detJ
    for ca in range(nsym):
        for i in range(npf):
            # (div v)_comp = sum_x E(x, comp) d_x phi
            for comp in range(D):
                dv = {}
                for x in range(D):
                    if Et[ca][x][comp]:
                        dv = padd(dv, pscale(pderiv(phi[i], x), Et[ca][x][comp] / h[x]))
                if not dv:
                    continue
                for j in range(nps):
                    v = pint_cube(pmul(dv, psi[j]), D) * detJ
                    A[ca * npf + i][na + comp * nps + j] += v
                    A[na + comp * nps + j][ca * npf + i] += -v
    for f in range(nF):
        fix, side, free = facet_info(D, f)
        area = detJ / h[fix]
        nrm = [Fr(0)] * D
        nrm[fix] = Fr(1) if side else Fr(-1)
        phF = [restrict_to_facet(p, D, f) for p in phi]
        psF = [restrict_to_facet(p, D, f) for p in psi]
        lo = ni + f * ntf
        I = lambda p: pint_cube(p, D - 1) * area  # noqa: E731
        for ca in range(nsym):
            vn = [sum(Et[ca][x][y] * nrm[y] for y in range(D)) for x in range(D)]
            for comp in range(D):
                if vn[comp] == 0:
                    continue
                for m in range(nm):
                    for i in range(npf):
                        v = I(pmul(phF[i], mu[m])) * vn[comp]
                        A[ca * npf + i][lo + comp * nm + m] += -v
                        A[lo + comp * nm + m][ca * npf + i] += v
        if projection:
            Am = [[I(pmul(mu[m], mu[m2])) for m2 in range(nm)] for m in range(nm)]
            Bm = [[I(pmul(mu[m], psF[j])) for j in range(nps)] for m in range(nm)]
            X = frac_solve(Am, Bm)
            PP = [[sum(Bm[m][i] * X[m][j] for m in range(nm)) for j in range(nps)]
                  for i in range(nps)]
        for comp in range(D):
            for i in range(nps):
                for j in range(nps):
                    A[na + comp * nps + i][na + comp * nps + j] += tau * (
                        PP[i][j] if projection else I(pmul(psF[i], psF[j])))
                for m in range(nm):
                    v = tau * I(pmul(psF[i], mu[m]))
                    A[na + comp * nps + i][lo + comp * nm + m] += -v
                    A[lo + comp * nm + m][na + comp * nps + i] += -v
            for m in range(nm):
                for m2 in range(nm):
                    A[lo + comp * nm + m][lo + comp * nm + m2] += tau * I(pmul(mu[m], mu[m2]))
    return np.array([[float(v) for v in row] for row in A])


def kat_sources(D, x0, h):
    """test/DarcyHDGTests.jl:9-14: f = (1 + x_a), g = D, in reference coordinates
    x_a = x0_a + h_a xi_a."""
    f = []
    for a in range(D):
        e = [0, 0, 0]
        e[a] = 1
        f.append({(0, 0, 0): Fr(1) + x0[a], tuple(e): h[a]})
    return f, {(0, 0, 0): Fr(D)}


def main():
    # --- Darcy cells -----------------------------------------------------
    cases = {}
    x0_2, h_2 = (Fr(1, 4), Fr(1, 2)), (Fr(1, 2), Fr(1, 4))
    A, b = darcy_cell_exact(2, 1, 1, 1, False, h_2, Fr(2), Fr(3, 2), (1, -1, 1, -1),
                            kat_sources(2, x0_2, h_2))
    cases["darcy_q1_2d"] = dict(A=A, b=b, D=2, k=(1, 1, 1), space=0, x0=[0.25, 0.5],
                                h=[0.5, 0.25], kappa=2.0, tau=1.5, signs=[1, -1, 1, -1])
    A, b = darcy_cell_exact(2, 2, 2, 2, False, h_2, Fr(1, 3), Fr(1), (-1, 1, 1, -1),
                            kat_sources(2, x0_2, h_2))
    cases["darcy_q2_2d"] = dict(A=A, b=b, D=2, k=(2, 2, 2), space=0, x0=[0.25, 0.5],
                                h=[0.5, 0.25], kappa=1.0 / 3.0, tau=1.0, signs=[-1, 1, 1, -1])
    A, b = darcy_cell_exact(2, 1, 0, 1, True, h_2, Fr(1), Fr(1), (1, 1, -1, 1),
                            kat_sources(2, x0_2, h_2))
    cases["darcy_p1p0p1_2d"] = dict(A=A, b=b, D=2, k=(1, 0, 1), space=1, x0=[0.25, 0.5],
                                    h=[0.5, 0.25], kappa=1.0, tau=1.0, signs=[1, 1, -1, 1])
    x0_3, h_3 = (Fr(1, 2), Fr(0), Fr(1, 4)), (Fr(1, 4), Fr(1, 2), Fr(1, 8))
    A, b = darcy_cell_exact(3, 1, 1, 1, False, h_3, Fr(5, 4), Fr(1), (1, -1, 1, 1, -1, 1),
                            kat_sources(3, x0_3, h_3))
    cases["darcy_q1_3d"] = dict(A=A, b=b, D=3, k=(1, 1, 1), space=0, x0=[0.5, 0.0, 0.25],
                                h=[0.25, 0.5, 0.125], kappa=1.25, tau=1.0,
                                signs=[1, -1, 1, 1, -1, 1])
    # the 3D Darcy workload's own space (Q2/Q2/Q2, 162 x 162) and the cubic
    # 2D space, on anisotropic box cells with non-unit kappa and tau
    A, b = darcy_cell_exact(3, 2, 2, 2, False, h_3, Fr(3, 4), Fr(2), (-1, 1, 1, -1, 1, -1),
                            kat_sources(3, x0_3, h_3))
    cases["darcy_q2_3d"] = dict(A=A, b=b, D=3, k=(2, 2, 2), space=0, x0=[0.5, 0.0, 0.25],
                                h=[0.25, 0.5, 0.125], kappa=0.75, tau=2.0,
                                signs=[-1, 1, 1, -1, 1, -1])
    A, b = darcy_cell_exact(2, 3, 3, 3, False, h_2, Fr(5, 2), Fr(1, 2), (1, 1, -1, -1),
                            kat_sources(2, x0_2, h_2))
    cases["darcy_q3_2d"] = dict(A=A, b=b, D=2, k=(3, 3, 3), space=0, x0=[0.25, 0.5],
                                h=[0.5, 0.25], kappa=2.5, tau=0.5, signs=[1, 1, -1, -1])
    for name, c in cases.items():
        np.savez(os.path.join(HERE, f"{name}.npz"), A=c["A"], b=c["b"], D=c["D"],
                 k=np.array(c["k"]), space=c["space"], x0=np.array(c["x0"]),
                 h=np.array(c["h"]), kappa=c["kappa"], tau=c["tau"],
                 signs=np.array(c["signs"], np.int8))
        print("wrote", name, c["A"].shape)
    # --- elasticity cells (zero load) -----------------------------------
    for D, k, h, nu, tau, name in ((2, 1, (Fr(1, 2), Fr(1, 4)), Fr(3, 10), Fr(4), "elas_p1_2d"),
                                   (3, 1, (Fr(1, 4), Fr(1, 2), Fr(1, 8)), Fr(1, 4), Fr(2),
                                    "elas_p1_3d"),
                                   (3, 2, (Fr(1, 4), Fr(1, 2), Fr(1, 8)), Fr(1, 3), Fr(3),
                                    "elas_p2_3d")):
        A = elasticity_cell_exact(D, k, h, Fr(1), nu, tau)
        np.savez(os.path.join(HERE, f"{name}.npz"), A=A, D=D, k=k,
                 h=np.array([float(v) for v in h]), nu=float(nu), tau=float(tau))
        print("wrote", name, A.shape)
    # the same with the bulk-to-skeleton projection P_m
    for D, k, h, nu, tau, name in ((2, 1, (Fr(1, 2), Fr(1, 4)), Fr(3, 10), Fr(4), "elas_p1_2d_pm"),
                                   (3, 2, (Fr(1, 4), Fr(1, 2), Fr(1, 8)), Fr(1, 3), Fr(3),
                                    "elas_p2_3d_pm")):
        A = elasticity_cell_exact(D, k, h, Fr(1), nu, tau, projection=True)
        np.savez(os.path.join(HERE, f"{name}.npz"), A=A, D=D, k=k,
                 h=np.array([float(v) for v in h]), nu=float(nu), tau=float(tau), projection=1)
        print("wrote", name, A.shape)
    # --- LAPACK condensation --------------------------------------------
    import scipy.linalg as sl
    rng = np.random.default_rng(20261202)
    sizes = (3, 3, 3)
    nc = 6
    n = sum(sizes)
    A = rng.standard_normal((nc, n, n))
    # untouched interior off-diagonal blocks, test/StaticCondensationMapTests.jl:15
    A[:, 3:6, 3:9] = 0.0
    A[:, 6:9, 3:9] = 0.0
    A[:, 3:6, 3:6] = rng.standard_normal((nc, 3, 3))
    A += 3.0 * np.eye(n)[None]
    b = rng.standard_normal((nc, n))
    x = rng.standard_normal((nc, 3))
    ni = 6
    S = np.zeros((nc, 3, 3))
    g = np.zeros((nc, 3))
    xi = np.zeros((nc, ni))
    for c in range(nc):
        lu = sl.lu_factor(A[c, :ni, :ni])
        W = sl.lu_solve(lu, A[c, :ni, ni:])
        w = sl.lu_solve(lu, b[c, :ni])
        S[c] = A[c, ni:, ni:] - A[c, ni:, :ni] @ W
        g[c] = b[c, ni:] - A[c, ni:, :ni] @ w
        xi[c] = sl.lu_solve(lu, b[c, :ni] - A[c, :ni, ni:] @ x[c])
    np.savez(os.path.join(HERE, "condense_lapack.npz"), A=A, b=b, x=x, S=S, g=g, xi=xi,
             sizes=np.array(sizes))
    print("wrote condense_lapack")


if __name__ == "__main__":
    main()
