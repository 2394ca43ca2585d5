"""Build the gfx950 HIP library ``libtodgpu.so`` in-tree with hipcc.

The library is a plain C-ABI shared object (``include/todgpu.h``) loaded with
ctypes; no torch types cross its boundary.  A second library,
``libtodgpu_checked.so``, is compiled from the same sources with
``-DTODG_CHECKED``: every map gather, weight-grid deposit and ring-slot
lookup also tests its index and counts violations (``todg_check_counts``),
for tests and debugging; the release library's code is unchanged by it.
Flags:

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
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_NAME = "libtodgpu.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
CHECKED_LIB_PATH = os.path.join(PKG_DIR, "libtodgpu_checked.so")

SOURCES = ["tod_fp_bilinear.hip", "tod_fp_nearest.hip", "tod_fp_bicubic.hip", "tod_fp_learned.hip",
           "tod_fp_fcorr.hip", "tod_kernels.hip", "corrector.hip", "beam_tools.hip"]
HEADERS = ["healpix_ring.hpp", "internal.hpp", "corrector.hpp", "tod_device.hpp", "check.hpp"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
OBJ_DIR = os.path.join(PKG_DIR, "_build")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False, jobs: int | None = None, checked: bool = True) -> str:
    """Compile libtodgpu.so and (when checked) libtodgpu_checked.so: one
    object per source and library, compiled in parallel, then linked.
    Always recompiles; returns the release library's path."""
    targets = [(LIB_PATH, OBJ_DIR, [])]
    if checked:
        targets.append((CHECKED_LIB_PATH, os.path.join(OBJ_DIR, "checked"), ["-DTODG_CHECKED"]))
    inc = ["-I", os.path.join(REPO_DIR, "include")]
    cmds, links = [], []
    for lib_path, obj_dir, defs in targets:
        os.makedirs(obj_dir, exist_ok=True)
        objs = [os.path.join(obj_dir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        cmds += [[hipcc(), *HIPCC_FLAGS, *defs, *inc, "-c", "-o", o, os.path.join(CSRC, s)]
                 for s, o in zip(SOURCES, objs)]
        links.append((lib_path, obj_dir, objs))
    jobs = jobs or max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    for lib_path, obj_dir, objs in links:
        # export the C ABI (todg_*) only; the C++ internals stay local
        vscript = os.path.join(obj_dir, "exports.map")
        with open(vscript, "w") as f:
            f.write("{ global: todg_*; local: *; };\n")
        tmp = lib_path + ".tmp"
        _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", f"-Wl,--version-script={vscript}",
              "-o", tmp, *objs], verbose)
        os.replace(tmp, lib_path)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
