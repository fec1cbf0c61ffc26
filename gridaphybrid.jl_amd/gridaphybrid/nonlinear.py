"""The hybrid Newton path: HybridFEOperator + Newton's method whose linear
solves are hybrid (condense -> skeleton solve -> recover).

Mirrors src/HybridFEOperators.jl (HybridFEOperator, residual / jacobian
evaluation into HybridFEVector / HybridFEMatrix) and
src/HybridLinearSolvers.jl:15-59 (solve!(x, ::HybridBackslashNumericalSetup,
::HybridFEVector): the Jacobian's cell matrices and the residual's cell
vectors are statically condensed, the skeleton system is assembled and solved
on the host (x_skel = A \\ b, :45), and the interior correction is recovered
with zero Dirichlet values, :47-57), driven by a Newton iteration as
NLSolver(method=:newton) drives it in test/DarcyHDGTests.jl:157-165.

The weak forms integrated by libhdg (Darcy and linear elasticity HDG) are
affine, so the Jacobian is the bilinear form and does not change between
iterations: its condensation (factors, S) is computed once on the GPU and
reused by every Newton step -- the reference re-integrates and re-condenses it
at every step.  With r(x) = A x - b:

* the condensed residual of a cell is r_b - A_bi A_ii^-1 r_i = S_c x_b - g_c
  (the interior unknowns cancel exactly), so the assembled skeleton residual
  is K lam - rhs on the free trace dofs, the Dirichlet part of the correction
  being zero;
* the interior correction A_ii^-1 (r_i - A_ib dlam) equals
  x_i - A_ii^-1 (b_i - A_ib (lam - dlam)): one BackwardStaticCondensationMap of
  the updated trace on the GPU.

Every step therefore runs the same GPU kernels as HybridAffineFEOperator +
solve (condensation, assembly, recovery); only the skeleton solve is on the
host, as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._lib import HDGError, HDG_ERR_ARG


@dataclass
class NLSolver:
    """NLSolver(show_trace=..., method=:newton) of NLsolve.jl as the reference
    uses it: Newton's method, stopping when the residual's max norm is below
    ftol (relative to the right-hand side's) or the step below xtol."""
    method: str = "newton"
    ftol: float = 1e-12
    xtol: float = 0.0
    iterations: int = 20
    show_trace: bool = False

    def __post_init__(self):
        if self.method != "newton":
            raise HDGError(HDG_ERR_ARG, f"NLSolver method {self.method!r}: only :newton")


@dataclass
class FESolver:
    """FESolver(nls): the nonlinear FE solver wrapping an NLSolver."""
    nls: NLSolver = field(default_factory=NLSolver)


@dataclass
class FEFunction:
    """FEFunction(X, free_values) of the hybridizable multi-field space: free
    values [flux of all cells; potential of all cells; free trace]."""
    free_values: np.ndarray


@dataclass
class NewtonTrace:
    """Per-iteration residual max norms (skeleton, interior) and step norms."""
    residual_skeleton: list = field(default_factory=list)
    residual_interior: list = field(default_factory=list)
    step: list = field(default_factory=list)
    converged: bool = False

    @property
    def iterations(self):
        return len(self.step)


class HybridFEOperator:
    """HybridFEOperator(residual, trial, test, bulk_fields, skeleton_fields)
    (src/HybridFEOperators.jl:85-145).  ``residual`` is the weak form whose
    residual a(x, v) - l(v) is solved for (DarcyHDG or ElasticityHDG); its
    Jacobian is the bilinear form a."""

    def __init__(self, residual, trial, test, bulk_fields, skeleton_fields, ctx=None):
        from . import HybridAffineFEOperator, FP64
        # the GPU operator integrates and condenses the Jacobian (bilinear form)
        # together with the linear form, which gives the residual's cell vectors
        self.op = HybridAffineFEOperator(residual, trial, test, bulk_fields, skeleton_fields,
                                         precision=FP64, ctx=ctx, assemble=False)
        self.trial, self.test = trial, test
        self.bulk_fields, self.skeleton_fields = list(bulk_fields), list(skeleton_fields)
        self.info = self.op.info
        self._jacobian = None

    def jacobian(self):
        """jacobian!(A, op, uh) for an affine form: the condensed skeleton
        system (host CSC) and the condensed linear form, condensed on the GPU
        once; the factors stay resident for every recovery."""
        if self._jacobian is None:
            self.op.assemble()
            self._jacobian = self.op.skeleton_system()
        return self._jacobian

    def num_free_dofs(self):
        return int(self.info.nfree_total)

    def close(self):
        self.op.close()


def _split(op, x):
    nc, ni = op.info.ncells, op.info.ni
    return x[:nc * ni], x[nc * ni:]


def solve_newton(xh0: FEFunction, solver: FESolver, op: HybridFEOperator):
    """solve!(xh0, solver, op) of test/DarcyHDGTests.jl:157-165 with the
    hybrid linear solver of src/HybridLinearSolvers.jl:15-59.  Returns
    (xh, trace)."""
    import scipy.sparse.linalg as spla
    nls = solver.nls
    x = np.array(xh0.free_values, np.float64, copy=True)
    if x.shape != (op.num_free_dofs(),):
        raise HDGError(HDG_ERR_ARG, f"initial free values of length {x.shape}, expected "
                                    f"{op.num_free_dofs()}")
    K, rhs = op.jacobian()
    lu = spla.splu(K.tocsc()) if K.shape[0] else None
    scale_b = max(float(np.max(np.abs(rhs))) if rhs.size else 0.0, 1e-300)
    trace = NewtonTrace()
    for it in range(nls.iterations + 1):
        xi, lam = _split(op, x)
        # residual: condensed on the skeleton, and the interior defect
        # A_ii^-1 r_i = x_i - A_ii^-1 (b_i - A_ib lam) from one recovery
        G = K @ lam - rhs if lam.size else np.zeros(0)
        xr = op.op.recover(lam).cpu().numpy()
        ei = xi - xr[:xi.size]
        rs = float(np.max(np.abs(G))) / scale_b if G.size else 0.0
        ri = float(np.max(np.abs(ei))) / max(float(np.max(np.abs(xr[:xi.size]))), 1e-300)
        trace.residual_skeleton.append(rs)
        trace.residual_interior.append(ri)
        if nls.show_trace:
            print(f"newton {it}: skeleton residual {rs:.3e}, interior defect {ri:.3e}", flush=True)
        if rs <= nls.ftol and ri <= nls.ftol:
            trace.converged = True
            break
        if it == nls.iterations:
            break
        # hybrid linear solve of J dx = r: skeleton correction (zero at the
        # Dirichlet dofs), then the interior correction by recovery
        dlam = lu.solve(G) if lu is not None else np.zeros(0)
        xn = op.op.recover(lam - dlam).cpu().numpy()
        dx = np.concatenate([xi - xn[:xi.size], dlam])
        x -= dx
        step = float(np.max(np.abs(dx))) if dx.size else 0.0
        trace.step.append(step)
        if nls.xtol > 0 and step <= nls.xtol:
            trace.converged = True
            break
    return FEFunction(x), trace
