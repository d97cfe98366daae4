"""Primary rays generated inside the trace kernel (sg_trace_camera) and the
SimpleRayBuffer output (sg_trace_closest_store), against the host generator
and the oracle bit for bit:

* the generated rays equal sg_gen_camera_rays (Camera::init +
  RegularPixelSampler<1,1> + RegularPrimaryRayGenerator) bit for bit, and
  the float64 evaluation of the same formula within binary32 rounding;
* every pixel is written exactly once whatever the tile shape (outputs are
  prefilled with NaN), including tiles that do not divide the image;
* hit ids and t equal the oracle's trace of the host rays;
* the 8 SimpleRayBuffer words equal the oracle's store (RayBuffers.h:52-77)."""
import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu


def _nan_buffer(ctx, words):
    b = sa.DeviceBuffer(ctx, 4 * words)
    b.upload(np.full(words, np.nan, np.float32))
    return b


def _camera64_check(camera, gen):
    """directions against Camera::getDirection at the exact pixel centres in
    float64, from the binary32 camera vectors of sg_pinhole_init.  The
    reference's sampler takes x as frac(id / resX) * resX in binary32, which
    moves x by up to resX * ulp(id / resX) pixels; the bound allows that plus
    a few ulps of the sum, and still fails for a half-pixel shift, a swapped
    or wrong step vector or a wrong row."""
    eye, fwd, up, fov, rx, ry = camera
    c = sa.pinhole(*camera)
    tl, sx, sy = (np.array(v, np.float64)[:, None] for v in (c.top_left, c.step_x, c.step_y))
    pid = np.arange(rx * ry)
    x = (pid % rx) + 0.5
    y = (pid // rx) + 0.5
    d64 = tl + x * sx + y * sy
    bound = np.abs(sx) * rx * (pid / rx + 1.0) * 2.0 ** -22 + 4e-7 * (np.abs(tl) + x * np.abs(sx) + y * np.abs(sy))
    err = np.abs(gen[3:6].astype(np.float64) - d64)
    assert (err <= bound).all(), float(np.max(err / bound))
    assert float(np.max(bound)) < 0.25 * float(np.min(np.abs(sx).max(axis=0) + np.abs(sy).max(axis=0)))


CASES = [
    ("soup10k", lambda: W.soup10k(res=256), (2, 32)),
    ("heightfield_odd", lambda: W.heightfield2m(grid=128, res_x=203, res_y=117), (2, 32)),
    ("heightfield_odd_7x5", lambda: W.heightfield2m(grid=128, res_x=203, res_y=117), (7, 5)),
    ("heightfield_rows", lambda: W.heightfield2m(grid=128, res_x=160, res_y=90), (1, 1)),
    ("heightfield_wide_tile", lambda: W.heightfield2m(grid=128, res_x=50, res_y=40), (64, 64)),
]


@pytest.mark.parametrize("name,make,tile", CASES, ids=[c[0] for c in CASES])
def test_camera_trace_equals_host_rays_and_oracle(ctx, oracle, name, make, tile):
    w = make()
    rx, ry = w.camera[4], w.camera[5]
    n = rx * ry
    scene = sa.FaceSoup(ctx, w.scene)
    builder = (sa.FastSimpleGridBuilder if w.kind == 0 else sa.FastTwoLevelGridBuilder)(ctx, **w.build)
    builder.init(scene)
    g = builder.build()
    cam = sa.pinhole(*w.camera)
    prim = _nan_buffer(ctx, n)
    t = _nan_buffer(ctx, n)
    raybuf = _nan_buffer(ctx, 8 * n)
    rays_out = sa.RayBatch(ctx, np.full((7, n), np.nan, np.float32))
    g.trace_camera(cam, tile, prim=prim, t=t, raybuf=raybuf, rays_out=rays_out)
    ctx.sync()
    gen = rays_out.download()
    # the device generator is the host generator, bit for bit, at every pixel
    np.testing.assert_array_equal(gen.view(np.uint32), w.rays.view(np.uint32))
    _camera64_check(w.camera, gen)
    og = (oracle.build_uniform if w.kind == 0 else oracle.build_twolevel)(w.scene, **w.oracle)
    po, to = og.trace(w.rays)
    np.testing.assert_array_equal(prim.download(np.uint32, n), po)
    np.testing.assert_array_equal(t.download(np.uint32, n), to.view(np.uint32))
    rb_o = og.trace_store(w.rays)
    np.testing.assert_array_equal(raybuf.download(np.uint32, 8 * n).reshape(8, n), rb_o.view(np.uint32))
    assert (po != oracle.MISS).any() and (po == oracle.MISS).any()
    g.close()
    scene.close()


@pytest.mark.parametrize("kind", ["uniform", "twolevel"])
def test_trace_store_equals_oracle(ctx, oracle, kind):
    """SoA rays in, SimpleRayBuffer out (with a device-side ray count that
    leaves the tail untouched)"""
    if kind == "uniform":
        w = W.soup10k(res=128)
        builder = sa.FastSimpleGridBuilder(ctx, **w.build)
    else:
        w = W.heightfield2m(grid=96, res_x=120, res_y=80)
        builder = sa.FastTwoLevelGridBuilder(ctx, **w.build)
    scene = sa.FaceSoup(ctx, w.scene)
    builder.init(scene)
    g = builder.build()
    n = w.rays.shape[1]
    rb = sa.RayBatch(ctx, w.rays)
    buf = _nan_buffer(ctx, 8 * n)
    prim = _nan_buffer(ctx, n)
    g.trace_store(rb, raybuf=buf, prim=prim)
    ctx.sync()
    og = (oracle.build_uniform if kind == "uniform" else oracle.build_twolevel)(w.scene, **w.oracle)
    np.testing.assert_array_equal(buf.download(np.uint32, 8 * n).reshape(8, n), og.trace_store(w.rays).view(np.uint32))
    np.testing.assert_array_equal(prim.download(np.uint32, n), og.trace(w.rays)[0])
    # device count: only the first m rays are traced and stored
    m = n // 3
    cnt = sa.DeviceBuffer(ctx, 4)
    cnt.upload(np.array([m], np.uint32))
    buf2 = _nan_buffer(ctx, 8 * n)
    g.trace_store(rb, raybuf=buf2, d_count=cnt.ptr)
    ctx.sync()
    got = buf2.download(np.float32, 8 * n).reshape(8, n)
    np.testing.assert_array_equal(got[:, :m].view(np.uint32), og.trace_store(w.rays).view(np.uint32)[:, :m])
    assert np.isnan(got[:, m:]).all()
    g.close()
    scene.close()
