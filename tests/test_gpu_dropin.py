"""The C++ drop-in layer (include/shinobi/grid.hpp) used the way the
reference's engine drives its builder: its grid and hits equal the oracle's
for the same scene and rays."""
import json
import os
import subprocess

import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "drop_in_example")


def _fnv(a):
    h = 1469598103934665603
    for x in np.asarray(a, np.uint32).tolist():
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_cpp_drop_in_matches_oracle(oracle):
    assert os.path.exists(EXE), "build() compiles tests/cpp/drop_in_example.cpp"
    n, nr = 20000, 50000
    out = subprocess.run([EXE, str(n), str(nr)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout.strip().splitlines()[-1])
    soa = sa.gen_soup(17, n, 0.02)
    rays = sa.gen_random_rays(18, nr, 0.0, 1.0)
    og = oracle.build_twolevel(soa, lambda_leaf=2.0)
    prim, _ = og.trace(rays)
    assert got["res"] == list(og.res)
    assert got["leavesCount"] == og.C2 and got["instancesCount"] == og.P2
    assert got["cells"] == _fnv(og.top_cells)
    assert got["leaves"] == _fnv(og.leaves)
    assert got["indices"] == _fnv(og.indices)
    assert got["prim"] == _fnv(prim)
    assert got["hits"] == int((prim != sa.MISS).sum())


# ------------------------------------------------ device concepts (device.hpp)
DEV_LIB = os.path.join(ROOT, "tests", "native", "_build", "libsg_dropin_device.so")


def _dev_lib():
    import ctypes as C
    assert os.path.exists(DEV_LIB), "build() compiles tests/native/dropin_device.hip"
    lib = C.CDLL(DEV_LIB)
    lib.dropin_trace_rays.argtypes = [C.c_void_p, C.c_void_p * 7, C.c_uint32, C.c_void_p]
    lib.dropin_trace_camera.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def _grid(ctx, w):
    scene = sa.FaceSoup(ctx, w.scene)
    b = (sa.FastSimpleGridBuilder if w.kind == 0 else sa.FastTwoLevelGridBuilder)(ctx, **w.build)
    b.init(scene)
    return scene, b.build()


DEVICE_CASES = [
    ("soup10k_uniform", lambda: W.soup10k(res=128)),
    ("heightfield_twolevel", lambda: W.heightfield2m(grid=128, res_x=150, res_y=100)),
    ("cluster_twolevel", lambda: W.cluster(n_small=40_000, n_large=300, res_x=96, res_y=64)),
]


@pytest.mark.parametrize("name,make", DEVICE_CASES, ids=[c[0] for c in DEVICE_CASES])
def test_device_concepts_kernel_matches_oracle(ctx, oracle, name, make):
    """A trace kernel written against the reference's device concepts
    (FaceSoup storage with indices / operator[] / operator(), SimpleIntersector,
    the grid's traverser with rcp(dir) in shared memory, RegularPrimaryRayGenerator,
    SimpleRayBuffer) and compiled with hipcc's default flags writes the
    oracle's SimpleRayBuffer bit for bit -- camera rays and random rays -- and
    equals the library's own sg_trace_closest_store."""
    import ctypes as C
    lib = _dev_lib()
    w = make()
    scene, g = _grid(ctx, w)
    view = g.device_view()
    assert view.status == 0 and view.kind == w.kind
    og = (oracle.build_uniform if w.kind == 0 else oracle.build_twolevel)(w.scene, **w.oracle)
    # camera rays from RegularPrimaryRayGenerator
    n = w.rays.shape[1]
    cam = sa.pinhole(*w.camera)
    buf = sa.DeviceBuffer(ctx, 8 * 4 * n)
    ctx.sync()
    assert lib.dropin_trace_camera(C.byref(view), C.byref(cam), buf.ptr) == 0
    expect = og.trace_store(w.rays)
    np.testing.assert_array_equal(buf.download(np.uint32, 8 * n).reshape(8, n), expect.view(np.uint32))
    # random rays through the SoA generator, against the oracle and the library
    rays = sa.gen_random_rays(77, 20_000, 0.0, 1.0)
    if w.name == "cluster":
        rays[0:3] = 0.5 + (rays[0:3] - 0.5) * 0.04      # origins around the dense cluster
    rb = sa.RayBatch(ctx, rays)
    buf2 = sa.DeviceBuffer(ctx, 8 * 4 * rays.shape[1])
    ptrs = (C.c_void_p * 7)(*[rb.field(k).value for k in range(7)])
    ctx.sync()
    assert lib.dropin_trace_rays(C.byref(view), ptrs, rays.shape[1], buf2.ptr) == 0
    got = buf2.download(np.uint32, 8 * rays.shape[1]).reshape(8, -1)
    expect2 = og.trace_store(rays)
    np.testing.assert_array_equal(got, expect2.view(np.uint32))
    assert (expect2[7].view(np.uint32) != sa.MISS).sum() > 100
    lib_buf = g.trace_store(rb)
    ctx.sync()
    np.testing.assert_array_equal(lib_buf.download(np.uint32, 8 * rays.shape[1]).reshape(8, -1), got)
    # the storage concept: indices as the grid's reference -> triangle map
    assert view.n_refs == len(og.indices)
    g.close()
    scene.close()
