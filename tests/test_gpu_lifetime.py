"""Object lifetimes across contexts: closing a context closes what is still
open on it, a scene or grid destroyed after its context through the C ABI
does not read the destroyed context, and a later context works normally."""
import ctypes as C

import numpy as np
import pytest

import shinobi_amd as sa

pytestmark = pytest.mark.gpu


def _trace_once(ctx, scene9, rays7):
    soup = sa.FaceSoup(ctx, scene9)
    b = sa.FastTwoLevelGridBuilder(ctx, lam=5.0, lambda_leaf=2.0)
    b.init(soup)
    g = b.build()
    return soup, g, g.trace_host(rays7)


def test_context_close_then_new_context():
    scene9 = sa.gen_soup(61, 5_000, 0.02)
    rays7 = sa.gen_random_rays(62, 4_000, 0.0, 1.0)
    ctx1 = sa.Context(0)
    soup1, g1, (p1, t1) = _trace_once(ctx1, scene9, rays7)
    ctx1.close()                       # closes g1 and soup1 first
    assert not g1.h and not soup1.h
    ctx2 = sa.Context(0)
    _, _, (p2, t2) = _trace_once(ctx2, scene9, rays7)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(t1.view(np.uint32), t2.view(np.uint32))
    ctx2.close()


def test_grid_destroyed_after_its_context_c_abi():
    lib = sa.lib()
    scene9 = sa.gen_soup(63, 3_000, 0.02)
    rays7 = sa.gen_random_rays(64, 2_000, 0.0, 1.0)
    h = C.c_void_p()
    assert lib.sg_ctx_create(0, C.byref(h)) == 0
    sc = C.c_void_p()
    ptrs = sa._lib.Float9(*[scene9[i].ctypes.data for i in range(9)])
    assert lib.sg_scene_create(h, ptrs, scene9.shape[1], 0, None, C.byref(sc)) == 0
    p = sa._lib.BuildParams()
    lib.sg_build_params_default(C.byref(p), sa._lib.SG_GRID_TWO_LEVEL)
    gr = C.c_void_p()
    assert lib.sg_build(h, sc, C.byref(p), None, C.byref(gr)) == 0
    assert lib.sg_ctx_destroy(h) == 0
    assert lib.sg_grid_destroy(gr) == 0     # uses the device recorded at creation
    assert lib.sg_scene_destroy(sc) == 0
    with sa.Context(0) as ctx:
        _, _, (pg, tg) = _trace_once(ctx, scene9, rays7)
        assert (pg != sa.MISS).any()
