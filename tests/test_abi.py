"""The C-ABI library: it loads without a GPU, exports every entry point that
include/shinobi_grid.h declares, reports errors through status codes, and its
host-side generators are deterministic (no GPU calls here)."""
import ctypes as C
import math
import os
import re

import numpy as np

import shinobi_amd as sa
from shinobi_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("shinobi_grid.h",):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(sg_[a-z0-9_]+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(L.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 35
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    assert declared == set(L.SIGNATURES), declared ^ set(L.SIGNATURES)


def test_status_strings():
    lib = sa.lib()
    for s in range(12):
        assert lib.sg_status_str(s)
    assert lib.sg_status_str(99) == b"unknown status"


def test_no_device_is_an_error_not_a_crash():
    if sa.device_count() > 0:
        return
    h = C.c_void_p()
    assert sa.lib().sg_ctx_create(0, C.byref(h)) == 11      # SG_ERR_NO_DEVICE


def test_null_arguments_are_rejected():
    lib = sa.lib()
    assert lib.sg_grid_get_info(None, None) == 1
    assert lib.sg_trace_closest(None, None, None, 0, None, None, None, None, None) == 1
    assert lib.sg_rebuild(None, None, None, None) == 1


def test_soup_generator():
    a = sa.gen_soup(1, 10_000, 0.01)
    b = sa.gen_soup(1, 10_000, 0.01)
    assert np.array_equal(a, b) and a.shape == (9, 10_000)
    v = a.reshape(3, 3, -1)
    c = v.mean(axis=0)
    assert (np.abs(v - c) <= 0.0201).all()
    assert c.min() > -0.02 and c.max() < 1.02
    assert not np.array_equal(a, sa.gen_soup(2, 10_000, 0.01))


def test_heightfield_generator():
    g = 33
    a = sa.gen_heightfield(2, g, 0.2)
    assert a.shape == (9, 2 * (g - 1) ** 2)
    assert a[2].min() >= 0.0 and a[2].max() <= 0.2
    # the two triangles of quad (0,0) share the diagonal
    assert np.array_equal(a[0:3, 0], a[0:3, 1]) and np.array_equal(a[6:9, 0], a[3:6, 1])
    assert a[0, 0] == 0.0 and a[3, 0] == np.float32(1 / 32)
    assert sa.gen_heightfield(2, 1024, 0.2).shape[1] == 2_093_058


def test_camera_rays_follow_the_camera_formula():
    """Camera::init + getDirection (src/Primitive/Camera.h:41-70) evaluated
    independently in float64"""
    rx, ry, fov = 64, 48, 60.0
    r = sa.gen_camera_rays((0.5, 0.5, -1.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), fov, rx, ry)
    f = np.array([0, 0, 1.0])
    right = np.cross(f, [0, 1.0, 0])
    right /= np.linalg.norm(right)
    up = np.cross(f, right)
    tn = math.tan(math.radians(fov) / 2)
    row = 2 * right * tn * rx / ry
    col = 2 * up * tn
    tl = f - row / 2 - col / 2
    ids = np.arange(rx * ry)
    px, py = ids % rx + 0.5, ids // rx + 0.5
    d = tl[:, None] + np.outer(row / rx, px) + np.outer(col / ry, py)
    assert np.allclose(r[3:6], d, atol=2e-6)
    assert (r[0:3].T == [0.5, 0.5, -1.0]).all()
    assert (r[6] == np.float32(3.402823466e38)).all()


def test_random_rays_are_unit_and_in_the_box():
    r = sa.gen_random_rays(4, 10_000, 0.0, 1.0)
    assert (r[0:3] >= 0).all() and (r[0:3] <= 1).all()
    assert np.allclose(np.linalg.norm(r[3:6], axis=0), 1.0, atol=1e-6)
    assert abs(r[5].mean()) < 0.03


def test_cluster_generator():
    a = sa.gen_cluster(3, 5000, 100)
    v = a.reshape(3, 3, -1).astype(np.float64)
    c = v.mean(axis=0)
    small = np.linalg.norm(c[:, :5000] - 0.5, axis=0)
    assert (small <= 0.0201).all()
    assert (c[:, 5000:] >= -0.31).all() and (c[:, 5000:] <= 10.31).all()
