"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle (oracle/sg_oracle.c), the synthetic generators and camera set-up
(sg_gen.cpp), the OBJ / camera.cfg readers (sg_obj.cpp) and the serving
loop's first-fit arena (sg_arena.hpp), built with gcc / g++ (CPU only, no
GPU code) and driven by tests/native/san_driver.cpp over ordinary and
degenerate scenes, NaN / infinite / zero rays, malformed OBJ lines and
200,000 random arena operations.  Any sanitizer finding aborts the driver."""
import gzip
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "shinobi_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1",
       "-ffp-contract=off", "-pthread"]

BAD_OBJ = """# malformed lines the reader must skip or reject without reading out of bounds
v 0 0 0
v 1 0 0
v 0 1 0
v 1 1
v nan 1 2
v 1e40 0 0
vt 0.5 0.5
f 1 2 3
f 1/1 2/2 3/3
f 1//1 2//2 3//3 4//4
f 1 2
f -1 -2 -3
f 1 2 x
f
f 3 2 1 3 2 1 3 2 1 3 2 1
usemtl none
o object
g group
"""


def _have(cmd):
    return shutil.which(cmd) is not None


@pytest.mark.skipif(not (_have("gcc") and _have("g++")), reason="needs gcc and g++")
def test_host_code_under_asan_ubsan(tmp_path):
    out = tmp_path / "san"
    out.mkdir()
    objs = []
    c_src = os.path.join(ROOT, "oracle", "sg_oracle.c")
    o = str(out / "sg_oracle.o")
    subprocess.run(["gcc", "-std=c11", *SAN, "-fexcess-precision=standard", "-c", c_src, "-o", o], check=True)
    objs.append(o)
    for src in ("sg_gen.cpp", "sg_obj.cpp"):
        o = str(out / (src + ".o"))
        subprocess.run(["g++", "-std=c++17", *SAN, "-c", os.path.join(CSRC, src), "-o", o], check=True)
        objs.append(o)
    exe = str(out / "san_driver")
    subprocess.run(["g++", "-std=c++17", *SAN, os.path.join(ROOT, "tests", "native", "san_driver.cpp"), *objs,
                    "-o", exe, "-lm"], check=True)
    scene = out / "ruins.obj"
    with gzip.open(os.path.join(ROOT, "tests", "golden", "ruins.obj.gz"), "rb") as f, open(scene, "wb") as g:
        shutil.copyfileobj(f, g)
    bad = out / "bad.obj"
    bad.write_text(BAD_OBJ)
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=1:strict_string_checks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(scene), str(bad), os.path.join(ROOT, "tests", "golden", "ruins_camera.cfg")],
                       capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "all checks passed" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
