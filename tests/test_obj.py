"""Scene files on the host: the native OBJ / camera.cfg readers against
hand-derived answers and against the oracle restatement, and the reference's
ruins scene (tests/golden/ruins.obj.gz, ruins_camera.cfg)."""
import gzip
import os
import shutil

import numpy as np
import pytest

import shinobi_amd as sa
import obj_oracle as OO

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# every accepted / rejected form the reader distinguishes; comments give the
# triangles (0-based vertex indices) or why the line is rejected
CASES = b"""# comment line
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0.5 0.5 1
vn 0 0 1
vt 0 0
vt 0.5 0.5 0.5
f 1 2 3
f 1/1 3/1 4/1
f 1//1 2//1 5//1
f 1/1/1 2/1/1 3/1/1 4/1/1
fo 2 3 5
f 1 2
f 1 2 3 # trailing comment
v 1 2
v 1 2 3 4
vp 1 2 3
usemtl
usemtl Stone
mtllib ruins.mtl
g group
s off
o object
bogus entity line
f 5 4 3 2 1
v 0x1p-3 2e0 -0
f 6 1 2
v 1.0000000596046447762579 0 0\r
  f\t7 1 2\r
"""
EXPECTED_TRIS = [(0, 1, 2), (0, 2, 3), (0, 1, 4), (0, 1, 2), (0, 2, 3), (1, 2, 4),
                 (4, 3, 2), (4, 2, 1), (4, 1, 0), (5, 0, 1), (6, 0, 1)]
EXPECTED_BAD = 6   # f 1 2 / f ... # comment / v 1 2 / v 1 2 3 4 / vp / usemtl without name


def _write(tmp_path, data: bytes, name="scene.obj"):
    p = tmp_path / name
    p.write_bytes(data)
    return str(p)


def _expected_vertices():
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1], [0.125, 2.0, -0.0],
                  [1.0, 0, 0]], np.float32)
    return v


def test_obj_forms_hand_derived(tmp_path):
    scene, info = sa.load_obj(_write(tmp_path, CASES))
    assert info == {"n_vertices": 7, "n_faces": 8, "n_triangles": len(EXPECTED_TRIS), "n_bad_lines": EXPECTED_BAD}
    V = _expected_vertices()
    want = np.empty((9, len(EXPECTED_TRIS)), np.float32)
    for i, tri in enumerate(EXPECTED_TRIS):
        for c in range(3):
            want[3 * c:3 * c + 3, i] = V[tri[c]]
    np.testing.assert_array_equal(scene.view(np.uint32), want.view(np.uint32))


def test_obj_coordinates_round_twice_like_the_reference(tmp_path):
    """(float)strtod: the decimal just above the midpoint between 1 and
    1+2^-23 rounds to the midpoint as a double and then to 1.0 (ties to
    even); a direct conversion to float would give 1+2^-23"""
    scene, _ = sa.load_obj(_write(tmp_path, b"v 1.0000000596046447762579 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\n"))
    assert scene[0, 0] == np.float32(1.0)
    assert float(OO._libc.strtof(b"1.0000000596046447762579", None)) == float(np.float32(1.0) + np.float32(2**-23))


def test_obj_matches_oracle_on_forms(tmp_path):
    path = _write(tmp_path, CASES)
    scene, info = sa.load_obj(path)
    o_scene, nv, nf, bad = OO.read_obj(path)
    np.testing.assert_array_equal(scene.view(np.uint32), o_scene.view(np.uint32))
    assert (info["n_vertices"], info["n_faces"], info["n_bad_lines"]) == (nv, nf, bad)


def test_obj_index_outside_vertex_list_is_an_error(tmp_path):
    with pytest.raises(sa.ShinobiError):
        sa.load_obj(_write(tmp_path, b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n"))
    with pytest.raises(sa.ShinobiError):
        sa.load_obj(_write(tmp_path, b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf -1 2 3\n"))
    with pytest.raises(sa.ShinobiError):
        sa.load_obj(str(tmp_path / "missing.obj"))


@pytest.fixture(scope="module")
def ruins_obj(tmp_path_factory):
    d = tmp_path_factory.mktemp("ruins")
    p = d / "ruins.obj"
    with gzip.open(os.path.join(GOLDEN, "ruins.obj.gz"), "rb") as f, open(p, "wb") as g:
        shutil.copyfileobj(f, g)
    return str(p)


def test_ruins_scene(ruins_obj):
    scene, info = sa.load_obj(ruins_obj)
    assert info == {"n_vertices": 24625, "n_faces": 38665, "n_triangles": 38665, "n_bad_lines": 0}
    o_scene, nv, nf, bad = OO.read_obj(ruins_obj)
    np.testing.assert_array_equal(scene.view(np.uint32), o_scene.view(np.uint32))
    # first face of the file: f 1//1 2//1 3//1 over its first three vertices
    np.testing.assert_array_equal(scene[:, 0], np.array(
        [6.687335, -1.340005, 7.707626, 6.651253, -1.380855, 7.705503, 6.693880, -1.425483, 7.669686], np.float32))


def test_ruins_camera():
    path = os.path.join(GOLDEN, "ruins_camera.cfg")
    cam = sa.load_camera(path)
    o = OO.read_camera(path)
    assert (cam["res_x"], cam["res_y"], cam["fov_deg"]) == (1024, 1024, 60.0) == (o["resX"], o["resY"], o["FOV"])
    for k in ("position", "orientation", "up", "rotation"):
        assert np.array_equal(np.array(cam[k], np.float32).view(np.uint32), np.array(o[k], np.float32).view(np.uint32))
    assert cam["position"][0] == np.float32(-3.07111)
