"""ctypes binding of libhdg (include/hdg.h).

The library is the product: every numerical call of this package goes through
it.  There is no CPU fallback; when libhdg.so is missing or no HIP device is
present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhdg.so")

HDG_OK = 0
HDG_ERR_ARG = -1
HDG_ERR_HIP = -2
HDG_ERR_ALLOC = -3
HDG_ERR_UNSUPPORTED = -4
HDG_ERR_STATE = -5
HDG_ERR_NO_DEVICE = -6
HDG_SINGULAR = 1

FP64, MIXED = 0, 1
DARCY, ELASTICITY = 0, 1
SPACE_Q, SPACE_P = 0, 1

SOURCE_ZERO = 0
SOURCE_DARCY_KAT = 1
SOURCE_DARCY_LINEAR = 2
SOURCE_DARCY_SINPROD = 3
SOURCE_ELAS_LINEAR = 11
SOURCE_ELAS_SIN = 12
SOURCE_QP = 100

# every symbol include/hdg.h declares
EXPORTS = [
    "hdg_ctx_create", "hdg_ctx_destroy", "hdg_last_error", "hdg_ctx_stream",
    "hdg_ctx_synchronize", "hdg_ctx_memory", "hdg_memcpy_d2h", "hdg_memcpy_h2d", "hdg_static_condensation",
    "hdg_backward_static_condensation", "hdg_mesh_create", "hdg_mesh_create_cartesian",
    "hdg_mesh_destroy", "hdg_mesh_get_info", "hdg_mesh_set_dirichlet",
    "hdg_mesh_set_dirichlet_mms", "hdg_mesh_set_source", "hdg_condense_assemble",
    "hdg_skeleton_system", "hdg_skeleton_solve", "hdg_cell_blocks", "hdg_recover",
    "hdg_cell_systems",
    "hdg_condense", "hdg_assemble", "hdg_mesh_singular_count", "hdg_mesh_interpolate_trace",
    "hdg_local_factors", "hdg_batch_create", "hdg_batch_destroy", "hdg_batch_get_classes",
    "hdg_batch_launch", "hdg_batch_condense_assemble", "hdg_batch_recover",
    "hdg_mesh_refinement_stats", "hdg_profile_start", "hdg_profile_stop",
    "hdg_footprint_cartesian", "hdg_mesh_factor_layout", "hdg_mesh_set_family",
    "hdg_mesh_cell_quadrature", "hdg_static_condensation_noalloc", "hdg_mesh_set_trace_fields",
]

BATCH_AUTOTUNE = 1
FAMILY_NAMES = {0: "wave", 1: "grid8x8", 3: "grid32x16", 4: "blocked",
                5: "panel_mfma", 7: "grid16x16", 8: "flux_schur",
                10: "flux_ref"}
FAMILY_FLUX_SCHUR = 8
FAMILY_FLUX_REF = 10


class HDGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libhdg error {code}: {msg}")
        self.code = code


class SingularError(HDGError):
    """A local A_ii had a zero pivot (the reference's @check info==0)."""


class FactorLayout(C.Structure):
    _fields_ = [("rows", C.c_int), ("family", C.c_int), ("u_stride", C.c_int64),
                ("y_stride", C.c_int64)]


class Fields(C.Structure):
    _fields_ = [
        ("nfields", C.c_int), ("field_sizes", C.POINTER(C.c_int32)),
        ("n_interior", C.c_int), ("interior_fields", C.POINTER(C.c_int32)),
        ("n_boundary", C.c_int), ("boundary_fields", C.POINTER(C.c_int32)),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("physics", C.c_int), ("dim", C.c_int), ("space", C.c_int), ("k_flux", C.c_int),
        ("k_pot", C.c_int), ("k_trace", C.c_int), ("quad_degree", C.c_int),
        ("tau", C.c_double), ("young", C.c_double), ("source", C.c_int),
        ("projection", C.c_int),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("ncells", C.c_int64), ("nfacets", C.c_int64), ("x0", C.c_void_p), ("jac", C.c_void_p),
        ("cell_facets", C.c_void_p), ("facet_dirichlet", C.c_void_p), ("coef", C.c_void_p),
        ("facet_perm", C.c_void_p),
    ]


class MeshInfo(C.Structure):
    _fields_ = [
        ("ni", C.c_int), ("nb", C.c_int), ("n_flux", C.c_int), ("n_pot", C.c_int),
        ("n_trace_facet", C.c_int), ("nq_cell", C.c_int), ("ncells", C.c_int64),
        ("nfacets", C.c_int64), ("nfree", C.c_int64), ("ndir", C.c_int64), ("nnz", C.c_int64),
        ("nfree_total", C.c_int64),
    ]


class BatchClass(C.Structure):
    _fields_ = [
        ("ni", C.c_int), ("nb", C.c_int), ("physics", C.c_int), ("dim", C.c_int),
        ("nmesh", C.c_int), ("ncells", C.c_int64), ("family", C.c_int),
        ("grid_per_cu", C.c_int), ("tuned_ms", C.c_double),
    ]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("total_ms", C.c_double)]


class RefinementStats(C.Structure):
    _fields_ = [
        ("mixed", C.c_int), ("fp64_fallback", C.c_int), ("max_sweeps", C.c_int),
        ("mean_sweeps", C.c_double), ("unconverged", C.c_int64),
    ]


_lib = None


def lib():
    """Load libhdg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HDGError(HDG_ERR_NO_DEVICE,
                           f"{LIB_PATH} not built (run __graft_entry__.build())")
        # torch ships its own HIP runtime under another file name: loaded
        # after libhdg it would be a second runtime in the process, and the
        # first device query of libhdg's fails.  Loaded first, libhdg binds to
        # it.  (A host without torch, such as a Julia ccall, is not affected.)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
        L.hdg_last_error.restype = C.c_char_p
        L.hdg_last_error.argtypes = [vp]
        L.hdg_ctx_stream.restype = vp
        L.hdg_ctx_stream.argtypes = [vp]
        L.hdg_ctx_create.argtypes = [i32, C.POINTER(vp)]
        L.hdg_ctx_destroy.argtypes = [vp]
        L.hdg_ctx_synchronize.argtypes = [vp]
        L.hdg_ctx_memory.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64)]
        L.hdg_memcpy_d2h.argtypes = [vp, vp, vp, i64]
        L.hdg_memcpy_h2d.argtypes = [vp, vp, vp, i64]
        L.hdg_static_condensation.argtypes = [vp, C.POINTER(Fields), i64, vp, vp, vp, vp, vp]
        L.hdg_static_condensation_noalloc.argtypes = [vp, C.POINTER(Fields), i64, vp, vp, vp,
                                                      vp, vp]
        L.hdg_backward_static_condensation.argtypes = [vp, C.POINTER(Fields), i64, vp, vp, vp,
                                                       vp, vp]
        L.hdg_mesh_create.argtypes = [vp, C.POINTER(Problem), C.POINTER(MeshDesc), C.POINTER(vp)]
        L.hdg_mesh_create_cartesian.argtypes = [vp, C.POINTER(Problem), vp, vp, vp, C.POINTER(vp)]
        L.hdg_mesh_destroy.argtypes = [vp]
        L.hdg_mesh_get_info.argtypes = [vp, C.POINTER(MeshInfo)]
        L.hdg_mesh_set_dirichlet.argtypes = [vp, vp]
        L.hdg_mesh_set_dirichlet_mms.argtypes = [vp]
        L.hdg_mesh_set_source.argtypes = [vp, vp]
        L.hdg_mesh_cell_quadrature.argtypes = [vp, vp, vp]
        L.hdg_mesh_set_trace_fields.argtypes = [vp, i32]
        L.hdg_condense_assemble.argtypes = [vp, i32, C.POINTER(i64)]
        L.hdg_skeleton_solve.argtypes = [vp, C.c_double, i32, vp, C.POINTER(C.c_int),
                                         C.POINTER(C.c_double)]
        L.hdg_skeleton_system.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                          C.POINTER(vp)]
        L.hdg_cell_blocks.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
        L.hdg_recover.argtypes = [vp, vp, vp]
        L.hdg_cell_systems.argtypes = [vp, i64, i64, vp, vp]
        L.hdg_condense.argtypes = [vp, i32]
        L.hdg_assemble.argtypes = [vp]
        L.hdg_mesh_singular_count.argtypes = [vp, C.POINTER(i64)]
        L.hdg_mesh_interpolate_trace.argtypes = [vp, vp]
        L.hdg_local_factors.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
        L.hdg_mesh_factor_layout.argtypes = [vp, C.POINTER(FactorLayout)]
        L.hdg_mesh_set_family.argtypes = [vp, i32, i32]
        L.hdg_batch_create.argtypes = [vp, i32, C.POINTER(vp), i32, C.POINTER(vp)]
        L.hdg_batch_destroy.argtypes = [vp]
        L.hdg_batch_get_classes.argtypes = [vp, i32, C.POINTER(BatchClass), C.POINTER(i32)]
        L.hdg_batch_launch.argtypes = [vp, i32]
        L.hdg_batch_condense_assemble.argtypes = [vp, i32, C.POINTER(i64)]
        L.hdg_batch_recover.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
        L.hdg_mesh_refinement_stats.argtypes = [vp, C.POINTER(RefinementStats)]
        L.hdg_profile_start.argtypes = [vp]
        L.hdg_footprint_cartesian.argtypes = [C.POINTER(Problem), vp, i32, C.POINTER(i64)]
        L.hdg_profile_stop.argtypes = [vp, i32, C.POINTER(KernelTime), C.POINTER(i32)]
        _ = dbl
        _lib = L
    return _lib


def check(rc, ctx=None, what=""):
    if rc == HDG_OK:
        return
    msg = lib().hdg_last_error(ctx).decode() if ctx else ""
    if rc == HDG_SINGULAR:
        raise SingularError(rc, f"{what}: {msg}")
    raise HDGError(rc, f"{what}: {msg}")
