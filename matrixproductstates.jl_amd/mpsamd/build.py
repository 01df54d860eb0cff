"""In-tree build of libmpship.so for gfx950 (hipcc; no JIT caches)."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG)
SOURCES = ["gemm.hip", "qr.hip", "svd.hip", "misc.hip", "contract.hip", "sweeps.hip", "capi.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into matrixproductstates.jl_amd/lib/libmpship.so."""
    out_dir = os.path.join(PKG, "lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libmpship.so")
    srcs = [os.path.join(PKG, "csrc", s) for s in SOURCES if os.path.exists(os.path.join(PKG, "csrc", s))]
    tmp = out + ".tmp"
    cmd = ["hipcc", *FLAGS, "-I" + os.path.join(REPO, "include"), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-8000:])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
