"""Test meshes of affine cells that are not CartesianDiscreteModels: a
Cartesian grid mapped by a global shear, with a seeded subset of cells
re-parametrised by a rotation of the reference cell.  A rotated cell's local
facets come in another order and its local facet frames no longer agree with
its neighbours', which the orientation codes (the reference's pindex,
src/SkeletonArrays.jl:28-37) describe.  Geometry only: no finite-element
quantity is computed here."""
import itertools

import numpy as np

from oracle import oracle as O


def facet_to_cell(D, lf, y):
    """Reference-cell point of local facet lf at facet coordinates y (Gridap
    order: lf fixes direction D-1-lf//2 at lf%2, the others ascending)."""
    fix = D - 1 - lf // 2
    xi = np.zeros(D)
    j = 0
    for a in range(D):
        if a == fix:
            xi[a] = lf % 2
        else:
            xi[a] = y[j]
            j += 1
    return xi


def facet_coords(D, lf, xi):
    fix = D - 1 - lf // 2
    return np.array([xi[a] for a in range(D) if a != fix])


def orient(D, p, y):
    """The orientation map of code p (hdg_mesh_desc)."""
    if D == 2:
        return np.array([1.0 - y[0] if p & 1 else y[0]])
    a, b = (y[1], y[0]) if p & 1 else (y[0], y[1])
    return np.array([1.0 - a if p & 2 else a, 1.0 - b if p & 4 else b])


def rotations(D):
    """Proper symmetries xi = R xi' + t of the reference cube (det R = +1)."""
    out = []
    for perm in itertools.permutations(range(D)):
        for signs in itertools.product((1, -1), repeat=D):
            R = np.zeros((D, D))
            for i, j in enumerate(perm):
                R[i, j] = signs[i]
            if np.linalg.det(R) > 0:
                out.append((R, np.array([1.0 if s < 0 else 0.0 for s in signs])))
    return out


def sheared_rotated_mesh(D, n, seed, frac=0.5, shear=None):
    """Returns (x0 (nc, D), J (nc, D, D), cell_facets (nc, 2D), nfacets,
    facet_perm (nc, 2D), rot) for the n^D Cartesian grid of the unit cube
    mapped by x -> B x, with about `frac` of the cells re-parametrised by a
    random proper rotation of the reference cell; rot[c] = (R, t) or None."""
    m = O.cartesian(D, (n,) * D)
    nc, nF = m.ncells, 2 * D
    if shear is None:
        shear = np.eye(D) + (np.array([[0, .3], [.2, 0]]) if D == 2 else
                             np.array([[0, .3, .1], [.2, 0, -.15], [.05, .1, 0]]))
    B = np.asarray(shear, float)
    x0 = m.x0.reshape(nc, D) @ B.T
    J = np.einsum("ab,cbe->cae", B, m.J.reshape(nc, D, D).transpose(0, 2, 1))
    cf = m.cell_facets.reshape(nc, nF).copy()
    perm = np.zeros((nc, nF), np.int8)
    rng = np.random.default_rng(seed)
    rots = rotations(D)
    rot = [None] * nc
    half = np.full(D - 1, 0.5)
    probes = [np.zeros(D - 1)] + [np.eye(D - 1)[i] for i in range(D - 1)]
    for c in range(nc):
        if rng.random() >= frac:
            continue
        R, t = rots[rng.integers(1, len(rots))]     # index 0 is the identity
        rot[c] = (R, t)
        x0[c] = x0[c] + J[c] @ t
        Jc = J[c] @ R
        newcf = np.zeros(nF, np.int32)
        for lfn in range(nF):
            # the old local facet this new local facet lies on
            ctr = R @ facet_to_cell(D, lfn, half) + t
            lfo = next(lf for lf in range(nF)
                       if np.allclose(facet_to_cell(D, lf, half), ctr))
            newcf[lfn] = cf[c, lfo]
            # facet-own coordinates (the Cartesian frame) of the new local
            # facet's probe points
            own = [facet_coords(D, lfo, R @ facet_to_cell(D, lfn, y) + t) for y in probes]
            perm[c, lfn] = next(p for p in range(2 if D == 2 else 8)
                                if all(np.allclose(orient(D, p, y), o)
                                       for y, o in zip(probes, own)))
        cf[c] = newcf
        J[c] = Jc
    return x0, J, cf, m.nfacets, perm, rot


def oracle_mesh(x0, J, cf, nfacets, perm):
    D = x0.shape[1]
    return O.general_mesh(D, x0.ravel(), J.transpose(0, 2, 1).ravel(), cf.ravel(), nfacets,
                          None if perm is None else perm.ravel())


def node_map(nodes, R, t):
    """For reference nodes (n, D): index in `nodes` of R xi' + t for every node
    xi' (the node of the original parametrisation that a re-parametrised
    cell's node i sits on)."""
    out = np.empty(len(nodes), int)
    for i, x in enumerate(nodes):
        y = R @ x + t
        out[i] = int(np.argmin(np.abs(nodes - y).sum(axis=1)))
        assert np.abs(nodes[out[i]] - y).sum() < 1e-12
    return out
