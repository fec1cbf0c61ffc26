"""GridapHybrid's hybridization API on MI355X.

Host-side mirror of the reference interface for the per-cell hybrid path:

* ``StaticCondensationMap(interior_fields, boundary_fields)`` and
  ``BackwardStaticCondensationMap(...)`` -- the reference's Gridap maps
  (src/StaticCondensationMap.jl, src/BackwardStaticCondensationMap.jl),
  evaluated for a whole batch of cells in one GPU launch;
* ``HybridAffineFEOperator(weakform, trial, test, bulk_fields, skeleton_fields)``
  and ``solve(op)`` (src/HybridAffineFEOperators.jl:10-100): fused cell
  integration + condensation + deterministic skeleton assembly on the GPU,
  the skeleton solve on the host (sparse direct), recovery on the GPU.

Field indices are 1-based, as in the reference's Julia calls
(``StaticCondensationMap([1,2],[3])``).  All numerics run in libhdg (HIP,
gfx950); nothing here computes on the CPU except the host sparse skeleton
solve, which the reference also keeps on the host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (DARCY, ELASTICITY, FP64, MIXED, SPACE_P, SPACE_Q, HDGError,  # noqa: F401
                   SingularError, check, lib)

__all__ = [
    "Context", "StaticCondensationMap", "BackwardStaticCondensationMap",
    "CartesianDiscreteModel", "AffineDiscreteModel", "HDGSpaces", "DarcyHDG", "ElasticityHDG",
    "HybridAffineFEOperator", "solve", "HDGError", "SingularError", "FP64", "MIXED",
    "HybridFEOperator", "NLSolver", "FESolver", "FEFunction", "solve_newton",
    "Scalar2ArrayBlockMap",
]

_torch = None


def _t():
    global _torch
    if _torch is None:
        import torch
        _torch = torch
    return _torch


# ----------------------------------------------------------------------------
# Context
# ----------------------------------------------------------------------------
class Context:
    """One libhdg context (device + stream).  Torch work that feeds or reads
    the library should run inside ``with ctx.stream():`` so it is ordered
    with the library's kernels."""

    _default = None

    def __init__(self, device=0):
        torch = _t()
        if not torch.cuda.is_available():
            raise HDGError(_lib.HDG_ERR_NO_DEVICE, "no HIP device visible")
        h = C.c_void_p()
        check(lib().hdg_ctx_create(int(device), C.byref(h)), None, "hdg_ctx_create")
        self.h = h
        self.device = int(device)
        self._stream = torch.cuda.ExternalStream(lib().hdg_ctx_stream(h),
                                                 device=torch.device("cuda", self.device))

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def stream(self):
        return _t().cuda.stream(self._stream)

    @property
    def torch_stream(self):
        return self._stream

    def synchronize(self):
        check(lib().hdg_ctx_synchronize(self.h), self.h, "synchronize")

    def profile_start(self):
        """Time every library kernel launch with HIP events until profile_stop."""
        check(lib().hdg_profile_start(self.h), self.h, "profile_start")

    def profile_stop(self):
        """{kernel family: (launches, total ms)} since profile_start."""
        n = C.c_int(0)
        arr = (_lib.KernelTime * 64)()
        check(lib().hdg_profile_stop(self.h, 64, arr, C.byref(n)), self.h, "profile_stop")
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms)
                for i in range(min(n.value, 64))}

    def close(self):
        if self.h:
            lib().hdg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dev(x, dtype):
    """Contiguous device tensor of the given dtype."""
    torch = _t()
    return torch.as_tensor(x, dtype=dtype, device="cuda").contiguous()


def _colmajor(A):
    """(ncells, n, m) logical -> device buffer laid out [cell][col][row]."""
    return A.transpose(1, 2).contiguous()


def _fields(field_sizes, interior_fields, boundary_fields):
    fs = (C.c_int32 * len(field_sizes))(*[int(v) for v in field_sizes])
    it = (C.c_int32 * max(len(interior_fields), 1))(*[int(v) - 1 for v in interior_fields])
    bd = (C.c_int32 * max(len(boundary_fields), 1))(*[int(v) - 1 for v in boundary_fields])
    f = _lib.Fields(len(field_sizes), fs, len(interior_fields), it, len(boundary_fields), bd)
    return f, (fs, it, bd)


def _check_preconditions(interior_fields, boundary_fields):
    """src/StaticCondensationMap.jl:16-34 (1-based fields)."""
    nf = len(interior_fields) + len(boundary_fields)
    seen = set()
    for f in list(interior_fields) + list(boundary_fields):
        if not (1 <= f <= nf) or f in seen:
            return False
        seen.add(f)
    return len(seen) == nf


# ----------------------------------------------------------------------------
# Map-level API
# ----------------------------------------------------------------------------
class StaticCondensationMap:
    """Batched StaticCondensationMap (src/StaticCondensationMap.jl:3-196).

    ``evaluate(A, b, field_sizes)``: A (ncells, n, n) and b (ncells, n) float64
    with the fields in their natural order; returns (S, g) with
    S = A22 - A21 A11^-1 A12 and g = b2 - A21 A11^-1 b1 per cell."""

    def __init__(self, interior_fields, boundary_fields, ctx=None):
        if not _check_preconditions(interior_fields, boundary_fields):
            raise HDGError(_lib.HDG_ERR_ARG, "interior/boundary fields must partition 1:nfields")
        self.interior_fields = list(interior_fields)
        self.boundary_fields = list(boundary_fields)
        self.ctx = ctx

    def _sizes(self, field_sizes):
        ni = sum(field_sizes[f - 1] for f in self.interior_fields)
        nb = sum(field_sizes[f - 1] for f in self.boundary_fields)
        return ni, nb

    def evaluate(self, A, b, field_sizes, return_info=False):
        torch = _t()
        ctx = self.ctx or Context.default()
        ni, nb = self._sizes(field_sizes)
        with ctx.stream():
            Ad = _colmajor(_dev(A, torch.float64))
            bd = _dev(b, torch.float64)
            nc = Ad.shape[0]
            S = torch.empty((nc, nb, nb), dtype=torch.float64, device="cuda")
            g = torch.empty((nc, nb), dtype=torch.float64, device="cuda")
            info = torch.empty((nc,), dtype=torch.int32, device="cuda")
        f, keep = _fields(field_sizes, self.interior_fields, self.boundary_fields)
        rc = lib().hdg_static_condensation(ctx.h, C.byref(f), nc, Ad.data_ptr(), bd.data_ptr(),
                                           S.data_ptr(), g.data_ptr(), info.data_ptr())
        ctx.synchronize()
        if rc != _lib.HDG_SINGULAR or not return_info:
            check(rc, ctx.h, "StaticCondensationMap")
        S = S.transpose(1, 2)  # stored column-major per cell
        return (S, g, info) if return_info else (S, g)


class BackwardStaticCondensationMap:
    """Batched BackwardStaticCondensationMap
    (src/BackwardStaticCondensationMap.jl:6-151).  ``evaluate(A, b, x,
    field_sizes)`` returns the per-field blocks [interior fields..., x] as the
    reference's ReblockInteriorDofsMap does."""

    def __init__(self, interior_fields, boundary_fields, ctx=None):
        self.sc = StaticCondensationMap(interior_fields, boundary_fields, ctx)

    def evaluate(self, A, b, x, field_sizes):
        torch = _t()
        ctx = self.sc.ctx or Context.default()
        ni, nb = self.sc._sizes(field_sizes)
        with ctx.stream():
            Ad = _colmajor(_dev(A, torch.float64))
            bd = _dev(b, torch.float64)
            xd = _dev(x, torch.float64)
            nc = Ad.shape[0]
            xi = torch.empty((nc, ni), dtype=torch.float64, device="cuda")
            info = torch.empty((nc,), dtype=torch.int32, device="cuda")
        f, keep = _fields(field_sizes, self.sc.interior_fields, self.sc.boundary_fields)
        rc = lib().hdg_backward_static_condensation(ctx.h, C.byref(f), nc, Ad.data_ptr(),
                                                    bd.data_ptr(), xd.data_ptr(), xi.data_ptr(),
                                                    info.data_ptr())
        ctx.synchronize()
        check(rc, ctx.h, "BackwardStaticCondensationMap")
        out, off = [], 0
        for fld in self.sc.interior_fields:
            out.append(xi[:, off:off + field_sizes[fld - 1]])
            off += field_sizes[fld - 1]
        out.append(xd)
        return out


# ----------------------------------------------------------------------------
# Operator-level API
# ----------------------------------------------------------------------------
class Scalar2ArrayBlockMap:
    """Scalar2ArrayBlockMap (src/Scalar2ArrayBlockMap.jl:1-69): splits a
    dense (A, b) into the (nblocks x nblocks) block matrix and block vector of
    block sizes bs.  Works on one cell (A: n x n) or a batch (A: ncells x n x
    n), on numpy arrays or torch tensors (the blocks are views)."""

    def evaluate(self, Ab, bs):
        A, b = Ab
        off = np.concatenate([[0], np.cumsum([int(v) for v in bs])])
        if off[-1] != A.shape[-1] or off[-1] != A.shape[-2] or off[-1] != b.shape[-1]:
            raise HDGError(_lib.HDG_ERR_ARG, "block sizes do not add up to the matrix size")
        Ab_ = [[A[..., off[i]:off[i + 1], off[j]:off[j + 1]] for j in range(len(bs))]
               for i in range(len(bs))]
        bb = [b[..., off[i]:off[i + 1]] for i in range(len(bs))]
        return Ab_, bb


@dataclass
class CartesianDiscreteModel:
    """CartesianDiscreteModel(domain, partition): cells lexicographic, x
    fastest; Dirichlet trace on every boundary facet."""
    domain: tuple
    partition: tuple

    @property
    def D(self):
        return len(self.partition)

    @property
    def num_cells(self):
        return int(np.prod(self.partition))


@dataclass
class AffineDiscreteModel:
    """A conforming mesh of affine cells (parallelograms / parallelepipeds)
    given cell by cell: origin x0 (ncells, D), column-major Jacobians J
    (ncells, D, D) as J[c][:, e] = the image of reference direction e, facet
    ids per cell in Gridap's local facet order (ncells, 2D), and the
    orientation code of each (cell, local facet) (the reference's pindex, see
    hdg_mesh_desc).  Boundary facets (one cell) carry the Dirichlet trace
    unless facet_dirichlet says otherwise."""
    x0: np.ndarray
    J: np.ndarray
    cell_facets: np.ndarray
    nfacets: int
    facet_perm: np.ndarray = None
    facet_dirichlet: np.ndarray = None

    @property
    def D(self):
        return int(np.asarray(self.x0).shape[1])

    @property
    def num_cells(self):
        return int(np.asarray(self.x0).shape[0])


@dataclass
class HDGSpaces:
    """The trial/test MultiFieldFESpace [flux, potential, trace] (Darcy) or
    [stress, displacement, trace] (elasticity) on a model."""
    model: CartesianDiscreteModel
    k_flux: int
    k_pot: int
    k_trace: int
    space: int = SPACE_Q

    @staticmethod
    def darcy_baseline(model, order):
        """Q_k flux, Q_k potential, Q_k trace."""
        return HDGSpaces(model, order, order, order, SPACE_Q)

    @staticmethod
    def darcy_reference_test(model):
        """test/DarcyHDGTests.jl:34-37: P1 flux, P0 potential, P1 trace."""
        return HDGSpaces(model, 1, 0, 1, SPACE_P)

    @staticmethod
    def elasticity(model, order):
        """test/LinearElasticityHDGTests.jl:88-94: P_k stress, P_{k+1} u, P_k trace."""
        return HDGSpaces(model, order, order + 1, order, SPACE_P)


@dataclass
class DarcyHDG:
    """Darcy HDG weak form of test/DarcyHDGTests.jl:125-135, with kappa^-1 on
    the mass term.  source: a built-in manufactured solution id
    (_lib.SOURCE_*) or per-cell source values at quadrature points."""
    tau: float = 1.0
    kappa: object = 1.0
    source: int = _lib.SOURCE_DARCY_LINEAR
    source_qp: object = None
    quad_degree: int = None
    physics: int = DARCY


@dataclass
class ElasticityHDG:
    """Elasticity HDG weak form of test/LinearElasticityHDGTests.jl:140-150.
    nu may vary per cell.  bulk_to_skeleton_projection: the displacement's
    facet traces in the tau terms are L2-projected onto the trace space (P_m,
    test/P_m.jl:4-23), as the reference's bulk_to_skeleton_projection=true."""
    tau: float = 1.0
    nu: object = 0.3
    E: float = 1.0
    source: int = _lib.SOURCE_ELAS_SIN
    source_qp: object = None
    quad_degree: int = None
    bulk_to_skeleton_projection: bool = False
    physics: int = ELASTICITY


class HybridAffineFEOperator:
    """HybridAffineFEOperator(weakform, trial, test, bulk_fields,
    skeleton_fields) -- src/HybridAffineFEOperators.jl:10-50.

    Construction runs the fused cell integration + StaticCondensationMap for
    every cell, the Dirichlet lifting and the deterministic skeleton assembly
    on the GPU; the condensed system and the resident local solves stay on the
    device until ``solve``."""

    def __init__(self, weakform, trial, test, bulk_fields, skeleton_fields, precision=FP64,
                 ctx=None, assemble=True, family=None, grid_per_cu=0):
        D = trial.model.D
        ntrace = 1
        if (list(bulk_fields) == [1, 2] and weakform.physics == ELASTICITY and
                list(skeleton_fields) == list(range(3, 3 + D))):
            ntrace = D   # the displacement trace as D scalar trace fields
        elif list(bulk_fields) != [1, 2] or list(skeleton_fields) != [3]:
            raise HDGError(_lib.HDG_ERR_UNSUPPORTED,
                           "hybrid operators here hybridize fields [1,2] onto trace field [3] "
                           "(elasticity: or onto its components [3, ..., 2+D])")
        if trial is not test and trial != test:
            raise HDGError(_lib.HDG_ERR_ARG, "trial and test spaces must match")
        self.weakform, self.trial, self.test = weakform, trial, test
        self.bulk_fields, self.skeleton_fields = list(bulk_fields), list(skeleton_fields)
        self.precision = precision
        self.ctx = ctx or Context.default()
        self.mesh = _create_mesh(self.ctx, weakform, trial)
        self.trace_fields = ntrace
        if ntrace > 1:
            check(lib().hdg_mesh_set_trace_fields(self.mesh, ntrace), self.ctx.h, "trace_fields")
        self.info = _lib.MeshInfo()
        check(lib().hdg_mesh_get_info(self.mesh, C.byref(self.info)), self.ctx.h, "info")
        self._src_keep = None
        if getattr(weakform, "source_qp", None) is not None:
            torch = _t()
            with self.ctx.stream():
                self._src_keep = _dev(weakform.source_qp, torch.float64)
            check(lib().hdg_mesh_set_source(self.mesh, self._src_keep.data_ptr()), self.ctx.h,
                  "set_source")
        check(lib().hdg_mesh_set_dirichlet_mms(self.mesh), self.ctx.h, "dirichlet")
        self._batch = None
        if family is not None:
            self.set_family(family, grid_per_cu)
        if assemble:
            self.assemble()

    def set_family(self, family, grid_per_cu=0):
        """Kernel family of the FP64 condensation (None or -1: the default for
        the class); see _lib.FAMILY_NAMES."""
        fam = -1 if family is None else int(family)
        check(lib().hdg_mesh_set_family(self.mesh, fam, int(grid_per_cu)), self.ctx.h,
              "set_family")

    def factor_layout(self):
        """hdg_mesh_factor_layout: rows of the resident factors, kernel family of
        the last condensation, per-cell strides of the U and Y slots."""
        lay = _lib.FactorLayout()
        check(lib().hdg_mesh_factor_layout(self.mesh, C.byref(lay)), self.ctx.h, "factor_layout")
        return lay

    def factor_rows(self):
        """(rows of the resident factors, kernel family of the last condensation)."""
        lay = self.factor_layout()
        return int(lay.rows), int(lay.family)

    def assemble(self):
        nbad = C.c_int64(0)
        rc = lib().hdg_condense_assemble(self.mesh, int(self.precision), C.byref(nbad))
        check(rc, self.ctx.h, "HybridAffineFEOperator")

    def skeleton_system(self):
        """Host scipy CSC matrix and rhs of the skeleton system."""
        import scipy.sparse as sp
        torch = _t()
        cp, rv, nz, rh = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().hdg_skeleton_system(self.mesh, C.byref(cp), C.byref(rv), C.byref(nz),
                                        C.byref(rh)), self.ctx.h, "skeleton_system")
        nfree, nnz = self.info.nfree, self.info.nnz
        colptr = np.empty(nfree + 1, np.int64)
        rowval = np.empty(nnz, np.int32)
        nzval = np.empty(nnz, np.float64)
        rhs = np.empty(nfree, np.float64)
        L = lib()
        for host, dev in ((colptr, cp), (rowval, rv), (nzval, nz), (rhs, rh)):
            check(L.hdg_memcpy_d2h(self.ctx.h, host.ctypes.data, dev, host.nbytes), self.ctx.h,
                  "d2h")
        _ = torch
        return sp.csc_matrix((nzval, rowval, colptr), shape=(nfree, nfree)), rhs

    def skeleton_blocks(self):
        """Per-cell local indices of each skeleton field in S and g (cell-local
        order: facet by facet, each facet's trace dofs component-major): one
        index array per trace field -- the block sizes the reference passes to
        Scalar2ArrayBlockMap (src/HybridAffineFEOperators.jl:343-356)."""
        nb, ntf = self.info.nb, self.info.n_trace_facet
        ncomp = self.trace_fields
        nm = ntf // max(ncomp, 1)
        idx = np.arange(nb).reshape(nb // ntf, ncomp, nm)
        return [idx[:, c, :].ravel() for c in range(ncomp)]

    def cell_blocks(self):
        """(S, g, cell_dofs) device tensors: S (ncells, nb, nb)."""
        torch = _t()
        S, g, d = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().hdg_cell_blocks(self.mesh, C.byref(S), C.byref(g), C.byref(d)), self.ctx.h,
              "cell_blocks")
        nc, nb = self.info.ncells, self.info.nb
        Sh = np.empty((nc, nb, nb)), np.empty((nc, nb)), np.empty((nc, nb), np.int32)
        for host, dev in zip(Sh, (S, g, d)):
            check(lib().hdg_memcpy_d2h(self.ctx.h, host.ctypes.data, dev, host.nbytes),
                  self.ctx.h, "d2h")
        _ = torch
        return Sh[0].transpose(0, 2, 1), Sh[1], Sh[2]

    def local_factors(self):
        """(U, Y) host copies: U (ncells, m, m) upper factor in LAPACK row order,
        Y (ncells, m, nb+1) = L^-1 P [B b] of the system the recovery
        back-substitutes: A_ii (m = ni) for the elimination families, the
        reduced system after the flux block (m = ni - n_flux) for flux_schur."""
        U, Y = C.c_void_p(), C.c_void_p()
        check(lib().hdg_local_factors(self.mesh, C.byref(U), C.byref(Y)), self.ctx.h, "factors")
        nc, nb = self.info.ncells, self.info.nb
        lay = self.factor_layout()
        ni = int(lay.rows)
        if not U.value:   # W = A_ii^-1 [A_ib b_i], natural order (mixed precision)
            Yh = np.empty((nc, nb + 1, ni))
            check(lib().hdg_memcpy_d2h(self.ctx.h, Yh.ctypes.data, Y, Yh.nbytes), self.ctx.h,
                  "d2h")
            return None, Yh.transpose(0, 2, 1)
        # per-cell slots are sized for the full interior; the factors of a
        # reduced system sit at the start of each slot
        Us = np.empty((nc, int(lay.u_stride)))
        Ys = np.empty((nc, int(lay.y_stride)))
        check(lib().hdg_memcpy_d2h(self.ctx.h, Us.ctypes.data, U, Us.nbytes), self.ctx.h, "d2h")
        check(lib().hdg_memcpy_d2h(self.ctx.h, Ys.ctypes.data, Y, Ys.nbytes), self.ctx.h, "d2h")
        Up = Us[:, :ni * (ni + 1) // 2]
        Yh = Ys[:, :(nb + 1) * ni].reshape(nc, nb + 1, ni)
        Uf = np.zeros((nc, ni, ni))
        for c in range(ni):
            Uf[:, :c + 1, c] = Up[:, c * (c + 1) // 2: c * (c + 1) // 2 + c + 1]
        return Uf, Yh.transpose(0, 2, 1)

    def recover(self, lam_free):
        """BackwardStaticCondensationMap over all cells + scatter
        (src/HybridAffineFEOperators.jl:102-150): free values [q; u; lambda]."""
        torch = _t()
        with self.ctx.stream():
            lam = _dev(lam_free, torch.float64)
            out = torch.empty((self.info.nfree_total,), dtype=torch.float64, device="cuda")
        check(lib().hdg_recover(self.mesh, lam.data_ptr() if lam.numel() else None,
                                out.data_ptr()), self.ctx.h, "recover")
        self.ctx.synchronize()
        return out

    def skeleton_solve(self, rtol=1e-12, maxit=0):
        """Device-side skeleton solve (hdg_skeleton_solve: MINRES with a facet
        block-Jacobi preconditioner, on the resident skeleton system) in place of the host
        ``A \\ b`` of src/HybridAffineFEOperators.jl:74.  Returns the free
        trace values (device tensor), the iterations and the final relative
        residual; raises when rtol was not met."""
        torch = _t()
        with self.ctx.stream():
            lam = torch.empty((self.info.nfree,), dtype=torch.float64, device="cuda")
        it, res = C.c_int(0), C.c_double(0.0)
        check(lib().hdg_skeleton_solve(self.mesh, float(rtol), int(maxit),
                                       lam.data_ptr() if lam.numel() else None, C.byref(it),
                                       C.byref(res)), self.ctx.h, "skeleton_solve")
        return lam, it.value, res.value

    def refinement_stats(self):
        """Refinement sweeps of the last condensation (mixed precision)."""
        st = _lib.RefinementStats()
        check(lib().hdg_mesh_refinement_stats(self.mesh, C.byref(st)), self.ctx.h, "stats")
        return {"mixed": bool(st.mixed), "fp64_fallback": bool(st.fp64_fallback),
                "max_sweeps": st.max_sweeps, "mean_sweeps": st.mean_sweeps,
                "unconverged": st.unconverged}

    def interpolate_trace(self):
        """Host interpolant of the built-in manufactured trace at the free
        trace dofs (Gridap's interpolate on the skeleton space)."""
        lam = np.empty(self.info.nfree)
        check(lib().hdg_mesh_interpolate_trace(self.mesh, lam.ctypes.data), self.ctx.h,
              "interpolate_trace")
        return lam

    def close(self):
        if getattr(self, "_batch", None) is not None:
            self._batch.close()
        # a context closed first (interpreter shutdown) took the device with it
        if getattr(self, "mesh", None) and self.ctx.h:
            lib().hdg_mesh_destroy(self.mesh)
        self.mesh = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HybridOperatorBatch:
    """Many independent HybridAffineFEOperators condensed and assembled
    together: one condensation launch per class of equal local systems, the
    whole sequence replayed from one captured HIP graph per ``launch``.  Build
    the operators with ``assemble=False``; afterwards their ``skeleton_system``,
    ``cell_blocks`` and ``recover`` work as for a lone operator."""

    def __init__(self, ops, autotune=True):
        if not ops:
            raise HDGError(_lib.HDG_ERR_ARG, "empty batch")
        self.ops = list(ops)
        self.ctx = self.ops[0].ctx
        arr = (C.c_void_p * len(self.ops))(*[op.mesh for op in self.ops])
        h = C.c_void_p()
        flags = _lib.BATCH_AUTOTUNE if autotune else 0
        check(lib().hdg_batch_create(self.ctx.h, len(self.ops), arr, flags, C.byref(h)),
              self.ctx.h, "batch_create")
        self.h = h
        for op in self.ops:
            op._batch = self

    def classes(self):
        n = C.c_int(0)
        check(lib().hdg_batch_get_classes(self.h, 0, None, C.byref(n)), self.ctx.h, "classes")
        arr = (_lib.BatchClass * n.value)()
        check(lib().hdg_batch_get_classes(self.h, n.value, arr, C.byref(n)), self.ctx.h,
              "classes")
        return [{"ni": c.ni, "nb": c.nb, "physics": c.physics, "dim": c.dim,
                 "nmesh": c.nmesh, "ncells": c.ncells,
                 "family": _lib.FAMILY_NAMES.get(c.family, str(c.family)),
                 "grid_per_cu": c.grid_per_cu, "tuned_ms": c.tuned_ms} for c in arr]

    def launch(self, precision=FP64):
        """Enqueue one graph replay (no host synchronisation)."""
        check(lib().hdg_batch_launch(self.h, int(precision)), self.ctx.h, "batch_launch")

    def assemble(self, precision=FP64):
        nbad = C.c_int64(0)
        check(lib().hdg_batch_condense_assemble(self.h, int(precision), C.byref(nbad)),
              self.ctx.h, "batch_condense_assemble")

    def recover(self, lams, outs):
        """Device tensors: lams[m] (nfree of member m), outs[m] (nfree_total)."""
        n = len(self.ops)
        la = (C.c_void_p * n)(*[t.data_ptr() if t.numel() else None for t in lams])
        oa = (C.c_void_p * n)(*[t.data_ptr() for t in outs])
        check(lib().hdg_batch_recover(self.h, la, oa), self.ctx.h, "batch_recover")

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().hdg_batch_destroy(self.h)
        self.h = None
        for op in getattr(self, "ops", []):
            op._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class HybridSolution:
    """The MultiFieldFEFunction's free values and its field blocks."""
    free_values: np.ndarray
    ncells: int
    n_flux: int
    n_pot: int
    lam: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def flux(self):
        return self.free_values[: self.ncells * self.n_flux].reshape(self.ncells, self.n_flux)

    @property
    def potential(self):
        a = self.ncells * self.n_flux
        return self.free_values[a: a + self.ncells * self.n_pot].reshape(self.ncells, self.n_pot)


def solve(op: HybridAffineFEOperator, skeleton_solver="host", rtol=1e-12):
    """solve(op) (src/HybridAffineFEOperators.jl:67-100): the skeleton solve,
    then GPU recovery.  skeleton_solver="host": the reference's sparse direct
    solve on the host; "device": hdg_skeleton_solve (iterative, to rtol)."""
    if skeleton_solver == "device":
        lam_d, _, _ = op.skeleton_solve(rtol=rtol)
        out = op.recover(lam_d).cpu().numpy()
        return HybridSolution(out, op.info.ncells, op.info.n_flux, op.info.n_pot,
                              lam_d.cpu().numpy())
    if skeleton_solver != "host":
        raise HDGError(_lib.HDG_ERR_ARG, f"unknown skeleton solver {skeleton_solver!r}")
    import scipy.sparse.linalg as spla
    K, rhs = op.skeleton_system()
    lam = spla.spsolve(K.tocsc(), rhs) if op.info.nfree else np.zeros(0)
    out = op.recover(lam).cpu().numpy()
    return HybridSolution(out, op.info.ncells, op.info.n_flux, op.info.n_pot, lam)


def _problem(weakform, spaces):
    k = max(spaces.k_flux, spaces.k_pot, spaces.k_trace)
    qd = weakform.quad_degree if weakform.quad_degree is not None else 2 * (spaces.k_flux + 1)
    if weakform.physics == DARCY:
        young = 1.0
    else:
        young = weakform.E
    _ = k
    return _lib.Problem(int(weakform.physics), spaces.model.D, int(spaces.space),
                        int(spaces.k_flux), int(spaces.k_pot), int(spaces.k_trace), int(qd),
                        float(weakform.tau), float(young),
                        _lib.SOURCE_QP if weakform.source_qp is not None else int(weakform.source),
                        int(bool(getattr(weakform, "bulk_to_skeleton_projection", False))))


def _create_mesh(ctx, weakform, spaces):
    model = spaces.model
    D = model.D
    nc = model.num_cells
    coef = weakform.kappa if weakform.physics == DARCY else weakform.nu
    coef = np.ascontiguousarray(np.broadcast_to(np.asarray(coef, np.float64), (nc,)))
    prob = _problem(weakform, spaces)
    if isinstance(model, AffineDiscreteModel):
        x0 = np.ascontiguousarray(model.x0, np.float64)
        # [cell][D][D] with J[c][a][e] = d x_a / d xi_e, stored column-major per cell
        jac = np.ascontiguousarray(np.asarray(model.J, np.float64).transpose(0, 2, 1))
        cf = np.ascontiguousarray(model.cell_facets, np.int32)
        if model.facet_dirichlet is None:
            cnt = np.bincount(cf.ravel(), minlength=int(model.nfacets))
            fd = np.ascontiguousarray(cnt == 1, np.int8)
        else:
            fd = np.ascontiguousarray(model.facet_dirichlet, np.int8)
        fp = None if model.facet_perm is None else np.ascontiguousarray(model.facet_perm, np.int8)
        desc = _lib.MeshDesc(nc, int(model.nfacets), x0.ctypes.data, jac.ctypes.data,
                             cf.ctypes.data, fd.ctypes.data, coef.ctypes.data,
                             None if fp is None else fp.ctypes.data)
        h = C.c_void_p()
        check(lib().hdg_mesh_create(ctx.h, C.byref(prob), C.byref(desc), C.byref(h)), ctx.h,
              "mesh_create")
        return h
    cells = (C.c_int64 * 3)(*(list(model.partition) + [1] * (3 - D)))
    dom = (C.c_double * 6)(*(list(model.domain) + [0.0] * (6 - 2 * D)))
    h = C.c_void_p()
    check(lib().hdg_mesh_create_cartesian(ctx.h, C.byref(prob), cells, dom, coef.ctypes.data,
                                          C.byref(h)), ctx.h, "mesh_create")
    return h


from .nonlinear import FEFunction, FESolver, HybridFEOperator, NLSolver, solve_newton  # noqa: E402
