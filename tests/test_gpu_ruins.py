"""The reference's own scene on the GPU: ruins.obj (38,665 triangles) read by
the native OBJ reader, built as a two-level grid with the reference builder's
defaults and as a uniform grid, traced with the 1024x1024 rays of its
camera.cfg and one diffuse bounce; every array, hit id and distance equal to
the oracle's."""
import os

import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ruins():
    return W.scene_file(os.path.join(GOLDEN, "ruins.obj.gz"), os.path.join(GOLDEN, "ruins_camera.cfg"))


def test_ruins_twolevel_build_trace_bounce(ctx, oracle, ruins):
    w = ruins
    print(f"ruins: {w.scene.shape[1]} tris, {w.rays.shape[1]} rays", flush=True)
    scene = sa.FaceSoup(ctx, w.scene)
    b = sa.FastTwoLevelGridBuilder(ctx, keep_debug=1, **w.build)   # keep_debug: the leaf keys too
    b.init(scene)
    g = b.build()
    og = oracle.build_twolevel(w.scene, **w.oracle)
    info = g.info()
    assert info.status == 0
    assert (info.C1, info.P1, info.C2, info.P2) == (og.C1, og.P1, og.C2, og.P2)
    for k in ("top_keys", "top_vals", "top_cells", "leaf_keys", "leaf_vals", "leaves", "indices"):
        np.testing.assert_array_equal(g.download(k), getattr(og, k), err_msg=k)

    rb = sa.RayBatch(ctx, w.rays)
    prim_d, t_d = g.trace(rb)
    prim = prim_d.download(np.uint32, rb.n)
    t = t_d.download(np.float32, rb.n)
    po, to = og.trace(w.rays)
    np.testing.assert_array_equal(prim, po)
    np.testing.assert_array_equal(t.view(np.uint32), to.view(np.uint32))
    assert (prim != sa.MISS).mean() > 0.5   # the camera looks into the ruins

    out = sa.RayBatch(ctx, capacity=rb.n)
    count = sa.bounce(ctx, scene, rb, prim_d, t_d, w.bounce_seed, out)
    m = int(count.download(np.uint32, 1)[0])
    b_o, _ = oracle.bounce(w.scene, w.rays, prim, t, w.bounce_seed)
    b_g = out.download(m)
    np.testing.assert_array_equal(b_g.view(np.uint32), b_o.view(np.uint32))
    out.n = m
    p2, t2 = g.trace(out)
    po2, to2 = og.trace(b_o)
    np.testing.assert_array_equal(p2.download(np.uint32, m), po2)
    np.testing.assert_array_equal(t2.download(np.float32, m).view(np.uint32), to2.view(np.uint32))


def test_ruins_uniform_grid(ctx, oracle, ruins):
    w = ruins
    scene = sa.FaceSoup(ctx, w.scene)
    b = sa.FastSimpleGridBuilder(ctx, lam=5.0)
    b.init(scene)
    g = b.build()
    og = oracle.build_uniform(w.scene, lam=5.0, top_divisor=1)
    np.testing.assert_array_equal(g.download("top_ranges"), og.top_ranges)
    np.testing.assert_array_equal(g.download("indices"), og.indices)
    sub = np.ascontiguousarray(w.rays[:, ::5])
    pg, tg = g.trace_host(sub)
    po, to = og.trace(sub)
    np.testing.assert_array_equal(pg, po)
    np.testing.assert_array_equal(tg.view(np.uint32), to.view(np.uint32))
