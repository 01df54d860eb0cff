"""In-tree build of libmpship.so for gfx950 (hipcc; no JIT caches).

Each HIP source is compiled to its own object in parallel (kernels are
launched only from host code in their own translation unit, so no
relocatable device code is needed), then the objects are linked into one
shared library.
"""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG)
SOURCES = ["gemm.hip", "gemm_capped.hip", "qr.hip", "svd.hip", "misc.hip", "contract.hip", "sweeps.hip", "steady.hip", "twosite.hip", "capi.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stderr[-8000:])


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every HIP source into matrixproductstates.jl_amd/lib/libmpship.so."""
    out_dir = os.path.join(PKG, "lib")
    obj_dir = os.path.join(out_dir, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    out = os.path.join(out_dir, "libmpship.so")
    srcs = [os.path.join(PKG, "csrc", s) for s in SOURCES if os.path.exists(os.path.join(PKG, "csrc", s))]
    objs = [os.path.join(obj_dir, os.path.basename(s).replace(".hip", ".o")) for s in srcs]
    inc = "-I" + os.path.join(REPO, "include")
    cmds = [["hipcc", *FLAGS, inc, "-c", "-o", o, s] for s, o in zip(srcs, objs)]
    jobs = jobs or min(len(cmds), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = out + ".tmp"
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
