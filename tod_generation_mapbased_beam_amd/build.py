"""Build the gfx950 HIP library ``libtodgpu.so`` in-tree with hipcc.

The library is a plain C-ABI shared object (``include/todgpu.h``) loaded with
ctypes; no torch types cross its boundary.  Flags:

* ``--offload-arch=gfx950``: MI355X only.
* ``-ffp-contract=off``: the HEALPix index path must evaluate binary64
  expressions in source order without FMA contraction, so that neighbour
  indices match the reference bit for bit.  Fused multiply-adds that are
  wanted are written explicitly.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 math.
"""

from __future__ import annotations

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_NAME = "libtodgpu.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)

SOURCES = ["tod_kernels.hip", "corrector.hip", "beam_tools.hip"]
HEADERS = ["healpix_ring.hpp", "internal.hpp", "corrector.hpp"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fPIC",
    "-shared",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile libtodgpu.so; returns its path.  Always recompiles."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(REPO_DIR, "include"), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
