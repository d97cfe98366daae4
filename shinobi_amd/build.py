"""Build libshinobi_grid.so (gfx950) in-tree: shinobi_amd/lib/.

Every HIP translation unit is compiled with -ffp-contract=off, correctly
rounded fp32 divide/sqrt and without fast-math: the build kernels must round
each binary32 operation exactly once, in source order, to stay bit-exact with
the restatement of the reference builder.  Packed-fp32 instructions are
disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libshinobi_grid.so")

HIP_SOURCES = ["sg_build.hip", "sg_sort.hip", "sg_trace.hip", "sg_api.hip", "sg_serve.hip"]
CXX_SOURCES = ["sg_gen.cpp", "sg_obj.cpp"]

ARCH = os.environ.get("SHINOBI_ARCH", "gfx950")

HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
]
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-pthread"]

# Per-source additions.  The traversal kernels are scheduled for
# instruction-level parallelism (LLVM's max-ilp machine scheduler strategy):
# on the height field the camera and bounce launches run 2 % faster (0.502
# -> 0.492 ms, 0.679 -> 0.663 ms) with the same registers and bit-identical
# results (scheduling reorders independent instructions only; contraction
# stays off).  The build sources gained nothing from it.
EXTRA_FLAGS = {"sg_trace.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]}


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = r.stdout.strip()
    # the host pass does not know the packed-fp32 target feature: harmless
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    if lines and os.environ.get("SHINOBI_BUILD_VERBOSE"):
        print("\n".join(lines))


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        objs.append(o)
        jobs.append([hipcc, *HIPFLAGS, *EXTRA_FLAGS.get(s, []), "-c", os.path.join(CSRC, s), "-o", o])
    for s in CXX_SOURCES:
        o = os.path.join(OBJ, s.replace(".cpp", ".o"))
        objs.append(o)
        jobs.append(["g++", *CXXFLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(len(jobs), 8)) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-pthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
