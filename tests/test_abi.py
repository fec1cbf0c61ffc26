"""The C ABI library and the host mirror, without a GPU (CPU only)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hdg.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hdg_[a-z0-9_]+)\s*\(", src)))


def test_library_loads_and_exports_every_declared_symbol():
    from gridaphybrid import _lib
    L = _lib.lib()
    decl = declared_functions()
    assert len(decl) >= 20
    missing = [f for f in decl if not hasattr(L, f)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == decl


def test_julia_shim_binds_only_exported_symbols():
    """Every libhdg symbol the Julia binding ccalls is declared and exported."""
    from gridaphybrid import _lib
    shim = os.path.join(ROOT, "gridaphybrid.jl_amd", "julia", "GridapHybridMI355X.jl")
    used = set(re.findall(r"ccall\(\(:(hdg_[a-z0-9_]+), libhdg\)", open(shim).read()))
    assert len(used) >= 8
    assert used <= set(declared_functions())
    L = _lib.lib()
    assert all(hasattr(L, f) for f in used)


def test_library_is_built_for_gfx950():
    from gridaphybrid import _lib
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_no_device_is_reported(monkeypatch):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from gridaphybrid import _lib
    h = C.c_void_p()
    assert _lib.lib().hdg_ctx_create(0, C.byref(h)) == _lib.HDG_ERR_NO_DEVICE


def test_null_arguments_are_rejected():
    from gridaphybrid import _lib
    L = _lib.lib()
    assert L.hdg_last_error(None) == b"null context"
    info = _lib.MeshInfo()
    assert L.hdg_mesh_get_info(None, C.byref(info)) == _lib.HDG_ERR_ARG
    assert L.hdg_mesh_destroy(None) == _lib.HDG_ERR_ARG
    assert L.hdg_ctx_destroy(None) == _lib.HDG_ERR_ARG
    assert L.hdg_condense(None, 0) == _lib.HDG_ERR_ARG
    assert L.hdg_recover(None, None, None) == _lib.HDG_ERR_ARG


def test_host_mirror_checks_field_partition():
    """StaticCondensationMap's _check_preconditions (src/StaticCondensationMap.jl:16-34)."""
    import gridaphybrid as gh
    gh.StaticCondensationMap([1, 2], [3])
    gh.StaticCondensationMap([2], [1, 3])
    for bad in (([1, 2], [2]), ([1, 1], [3]), ([1, 2], [4]), ([1], [3])):
        with pytest.raises(gh.HDGError):
            gh.StaticCondensationMap(*bad)


def test_product_path_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import gridaphybrid as gh
    with pytest.raises(gh.HDGError):
        gh.Context(0)


def test_newton_solver_options():
    """NLSolver accepts :newton only (the reference's NLSolver(method=:newton));
    the Newton path's host types import without a GPU."""
    import gridaphybrid as gh
    assert gh.NLSolver().method == "newton"
    with pytest.raises(gh.HDGError):
        gh.NLSolver(method="trust_region")
    assert gh.FESolver().nls.iterations > 0


def test_scalar2arrayblockmap():
    """test/Scalar2ArrayBlockMapTests.jl:1-19: bs = [8, 16], A rand(24, 24)."""
    import numpy as np
    import gridaphybrid as gh
    bs = [8, 16]
    rng = np.random.default_rng(0)
    A = rng.random((24, 24))
    b = rng.random(24)
    Ab, bb = gh.Scalar2ArrayBlockMap().evaluate((A, b), bs)
    assert np.array_equal(Ab[0][0], A[0:8, 0:8])
    assert np.array_equal(Ab[1][0], A[8:24, 0:8])
    assert np.array_equal(Ab[0][1], A[0:8, 8:24])
    assert np.array_equal(Ab[1][1], A[8:24, 8:24])
    assert np.array_equal(bb[0], b[:8]) and np.array_equal(bb[1], b[8:])
    with pytest.raises(gh.HDGError):
        gh.Scalar2ArrayBlockMap().evaluate((A, b), [8, 8])
