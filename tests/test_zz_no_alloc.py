"""prr_trace_closest_noalloc (no reference counterpart: the closest-hit batch
entry point for callers that cannot allocate) against prr_trace_closest, its
allocating twin, through the C ABI: primitive ids, t, u and v bit for bit, on
two calls in a row, the second with a smaller batch (the static spill array
keeps the first call's contents), on a scene deep enough that the LDS stack
spills. Also: n beyond PRR_NOALLOC_MAX_RAYS is rejected with a message."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(sc, n, seed):
    rng = np.random.default_rng(seed)
    b = np.asarray(sc.tris, np.float32).reshape(-1, 3)
    lo, hi = b.min(0), b.max(0)
    org = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org, d


def _call(L, fn, ctx, scene, org, d):
    import torch
    n = org.shape[0]
    o = torch.as_tensor(org, device="cuda").contiguous()
    dd = torch.as_tensor(d, device="cuda").contiguous()
    prim = torch.empty(n, dtype=torch.int32, device="cuda")
    t, u, v = (torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(3))
    ptr = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
    torch.cuda.synchronize()
    rc = getattr(L, fn)(ctx.h, scene.h, n, ptr(o), ptr(dd), ptr(prim), ptr(t), ptr(u), ptr(v))
    assert rc == 0, L.prr_last_error(ctx.h)
    assert L.prr_synchronize(ctx.h) == 0
    return [x.cpu().numpy() for x in (prim, t, u, v)]


def test_noalloc_equals_trace_closest_twice():
    from photorealisticrender4610_amd import abi, scenes
    L = abi.lib()
    ctx = abi.Context(0)
    sc = scenes.sphere_field(num_spheres=1024, env_size=(64, 32))
    scene = ctx.scene(sc)
    for n, seed in ((300000, 1), (70001, 2)):
        print("case noalloc n=%d seed=%d" % (n, seed), flush=True)
        org, d = _rays(sc, n, seed)
        want = _call(L, "prr_trace_closest", ctx, scene, org, d)
        got = _call(L, "prr_trace_closest_noalloc", ctx, scene, org, d)
        assert (want[0] >= 0).mean() > 0.05
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
    # an empty batch is fine; a batch beyond the maximum is refused with a message
    assert L.prr_trace_closest_noalloc(ctx.h, scene.h, 0, None, None, None, None, None, None) == 0
    rc = L.prr_trace_closest_noalloc(ctx.h, scene.h, 1 << 31, None, None, None, None, None, None)
    assert rc == -1 and b"PRR_NOALLOC_MAX_RAYS" in L.prr_last_error(ctx.h)
    scene.close()
    ctx.close()
