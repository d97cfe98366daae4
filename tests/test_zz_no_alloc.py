"""sg_bounce_noalloc (the allocation-free form of sg_bounce) against
sg_bounce through the C ABI: bounce rays, source indices and counts equal
bit for bit on two calls in a row, the second with a smaller input, and a
capacity above the fixed scratch's limit is rejected before anything is
enqueued."""
import ctypes as C

import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import _lib as L
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu


def _bounce(ctx, fn, scene, rb, prim, t, seed, cap):
    out = sa.RayBatch(ctx, np.full((7, cap), np.nan, np.float32))
    src = sa.DeviceBuffer(ctx, 4 * cap)
    src.upload(np.full(cap, 0xDEADBEEF, np.uint32))
    cnt = sa.DeviceBuffer(ctx, 4)
    cnt.upload(np.array([0xFFFFFFFF], np.uint32))
    r, o = rb.struct(), out.struct()
    st = fn(ctx.h, scene.h, C.byref(r), prim.ptr, t.ptr, rb.n, None, seed, C.byref(o), src.ptr, cnt.ptr,
            ctx.stream)
    assert st == 0, L.STATUS_NAMES.get(st, st)
    ctx.sync()
    m = int(cnt.download(np.uint32, 1)[0])
    assert 0 < m <= rb.n
    return m, out.download(m), src.download(np.uint32, m)


def test_noalloc_bounce_equals_bounce(ctx):
    w = W.heightfield2m(grid=128, res_x=320, res_y=240)
    scene = sa.FaceSoup(ctx, w.scene)
    b = sa.FastTwoLevelGridBuilder(ctx, **w.build)
    b.init(scene)
    g = b.build()
    big = sa.RayBatch(ctx, w.rays)
    # the lower half of the image: rays that hit the field (the top rows are sky)
    small = sa.RayBatch(ctx, w.rays[:, w.rays.shape[1] // 2:].copy())
    lib = ctx._l
    for rep, (rb, seed) in enumerate(((big, 11), (small, 12))):
        prim, t = g.trace(rb)
        ctx.sync()
        ref = _bounce(ctx, lib.sg_bounce, scene, rb, prim, t, seed, rb.n)
        got = _bounce(ctx, lib.sg_bounce_noalloc, scene, rb, prim, t, seed, rb.n)
        assert got[0] == ref[0], (rep, got[0], ref[0])
        np.testing.assert_array_equal(got[1].view(np.uint32), ref[1].view(np.uint32), err_msg=f"call {rep}")
        np.testing.assert_array_equal(got[2], ref[2], err_msg=f"call {rep}")
    # beyond the fixed scratch: rejected with its own status, nothing enqueued
    prim, t = g.trace(small)
    r, o = small.struct(), small.struct()
    cnt = sa.DeviceBuffer(ctx, 4)
    st = lib.sg_bounce_noalloc(ctx.h, scene.h, C.byref(r), prim.ptr, t.ptr, L.SG_BOUNCE_NOALLOC_MAX_RAYS + 1, None,
                               1, C.byref(o), None, cnt.ptr, ctx.stream)
    assert st == L.SG_ERR_SCRATCH_LIMIT
    ctx.sync()
    g.close()
    scene.close()
