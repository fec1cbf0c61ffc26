"""hdg_static_condensation_noalloc: the allocation-free variant of
hdg_static_condensation (a fixed device scratch array per GPU instead of
library allocations), called through the C ABI and checked against the
existing entry point on the same cells, on two calls in a row (the second
with fewer, smaller cells), plus its refusals."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _random_blocks(rng, sizes, nc):
    n = sum(sizes)
    A = rng.standard_normal((nc, n, n)) + n * np.eye(n)[None]
    b = rng.standard_normal((nc, n))
    return A, b


def _call(ctx, fn, sizes, A, b):
    import torch
    import gridaphybrid as gh
    from gridaphybrid import _lib
    nc, n = A.shape[0], sum(sizes)
    nb = sizes[-1]
    with ctx.stream():
        Ad = torch.as_tensor(np.ascontiguousarray(A.transpose(0, 2, 1)), device="cuda")
        bd = torch.as_tensor(b, device="cuda")
        S = torch.full((nc, nb, nb), np.nan, dtype=torch.float64, device="cuda")
        g = torch.full((nc, nb), np.nan, dtype=torch.float64, device="cuda")
        info = torch.full((nc,), -7, dtype=torch.int32, device="cuda")
    f, keep = gh._fields(list(sizes), [1, 2], [3])
    rc = fn(ctx.h, C.byref(f), nc, Ad.data_ptr(), bd.data_ptr(), S.data_ptr(), g.data_ptr(),
            info.data_ptr())
    ctx.synchronize()
    return rc, S.transpose(1, 2).cpu().numpy(), g.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("big,small", [((81, 27, 54), (18, 9, 12)), ((24, 8, 24), (8, 4, 8))])
def test_noalloc_matches_existing_entry_point(ctx, big, small):
    from gridaphybrid import _lib
    L = _lib.lib()
    rng = np.random.default_rng(sum(big))
    for sizes, nc in ((big, 300), (small, 37)):
        print(f"noalloc condensation sizes={sizes} ncells={nc}", flush=True)
        A, b = _random_blocks(rng, sizes, nc)
        rc0, S0, g0, i0 = _call(ctx, L.hdg_static_condensation, sizes, A, b)
        rc1, S1, g1, i1 = _call(ctx, L.hdg_static_condensation_noalloc, sizes, A, b)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(i0, i1) and not i1.any()
        assert O.rel_err(S1, S0) < 1e-12 and O.rel_err(g1, g0) < 1e-12
        ni = sizes[0] + sizes[1]
        So, go, _ = O.condense(A, b, ni, sizes[2])
        assert O.rel_err(S1, So) < 1e-12 and O.rel_err(g1, go) < 1e-12


def test_noalloc_singular_and_limits(ctx):
    from gridaphybrid import _lib
    L = _lib.lib()
    rng = np.random.default_rng(3)
    sizes = (5, 4, 6)
    A, b = _random_blocks(rng, sizes, 6)
    A[4, :9, :9] = 0.0
    rc, S, g, info = _call(ctx, L.hdg_static_condensation_noalloc, sizes, A, b)
    assert rc == _lib.HDG_SINGULAR and info[4] == 1 and info[[0, 1, 2, 3, 5]].sum() == 0
    # a system beyond the fixed maximum is refused with a clear error
    A, b = _random_blocks(rng, (120, 40, 60), 2)
    rc, *_ = _call(ctx, L.hdg_static_condensation_noalloc, (120, 40, 60), A, b)
    assert rc == _lib.HDG_ERR_UNSUPPORTED
    assert b"allocation-free" in L.hdg_last_error(ctx.h)
