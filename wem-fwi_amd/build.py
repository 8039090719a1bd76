"""Builds libwemfwi.so (HIP kernels + C++ host layer + C ABI) for gfx950.

Usage: python wem-fwi_amd/build.py   (writes wem-fwi_amd/libwemfwi.so)
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libwemfwi.so")
SOURCES = ["kernels.hip", "sweep_tile.hip", "born.hip", "host.cpp", "ops.cpp", "params.cpp", "traces.hip", "cmap.cpp", "efwi.hip", "interp.hip", "uneven.hip", "noalloc.hip"]
# Host-only sources compiled as plain C++: the conformal-map geometry must see
# the C library's complex functions (std::abs of std::complex<float> is glibc's
# hypotf there; HIP compile mode substitutes an inline formula that rounds
# differently, and the maps truncate coordinates to grid indices).
PLAIN_CXX = {"cmap.cpp"}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(HERE, "..", "include")]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        if src in PLAIN_CXX:
            cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-x", "c++", "-I", os.path.join(HERE, "..", "include"),
                   "-c", os.path.join(CSRC, src), "-o", obj]
        else:
            cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    # sources compile independently: run them side by side (MAX_JOBS caps it)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1), 16))

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(run, c) for c in cmds]:
            f.result()
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    shutil.rmtree(objdir, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
