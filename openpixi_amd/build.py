"""Builds the gfx950 shared library ``openpixi_amd/libopx.so`` in-tree with hipcc.

Run as ``python -m openpixi_amd.build``. The library is the product: the C ABI
declared in ``include/opx.h`` over hand-written HIP kernels.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libopx.so")

SOURCES = ["opx_kernels.hip", "opx_stream.hip", "opx_gauss.hip", "opx_adjoint.hip", "opx_particles.hip", "opx_tile.hip",
           "opx_plane.hip", "opx_implicit.hip", "opx_cgc.hip", "opx_capi.cpp",
           "opx_adjoint_capi.cpp", "opx_cgc_capi.cpp"]

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 are generated.
# -amdgpu-disable-unclustered-high-rp-reschedule: the machine scheduler keeps its first schedule of
# kernels at high register pressure instead of re-scheduling them unclustered; the double cell form
# (one wave per SIMD) keeps more loads in flight and runs 4 % faster, the other forms are unchanged.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-disable-unclustered-high-rp-reschedule",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build(verbose: bool = False) -> str:
    """Compile every source and link libopx.so (always recompiles)."""
    cc = hipcc()
    objs = []
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    cmds = []
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"]
        cmds.append([cc, *HIP_FLAGS, *inc, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    # translation units compile independently: up to 8 at a time
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as pool:
        for f in [pool.submit(run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
