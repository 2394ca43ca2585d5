"""ctypes binding of ``libtodgpu.so`` (declared in ``include/todgpu.h``).

torch is imported first on purpose: the library's HIP runtime dependency
(``libamdhip64.so.7``) then resolves to the runtime torch already loaded, so
device pointers and streams from torch tensors are valid inside the library.
There is no fallback: if the library is missing, every call raises.
"""

from __future__ import annotations

import contextlib
import ctypes
import os
import threading

import torch  # noqa: F401  (must be loaded before the HIP library)

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libtodgpu.so")
# the same sources compiled with -DTODG_CHECKED: index checks on every map
# gather, weight-grid deposit and ring-slot lookup (see checked())
CHECKED_LIB_PATH = os.path.join(_PKG_DIR, "libtodgpu_checked.so")
ABI_VERSION = 1
CHECK_KINDS = ("map_pixel", "grid_cell", "ring_slot")

TODG_OK = 0
TODG_ERR_ARG = -1
TODG_ERR_HIP = -2

_lock = threading.Lock()
_lib = None
_handles: dict = {}

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int

# name -> (restype, argtypes); mirrors include/todgpu.h
SIGNATURES = {
    "todg_version": (_i32, []),
    "todg_last_error": (_i32, [ctypes.c_char_p, ctypes.c_size_t]),
    "todg_check_counts": (_i32, [_p, _i32, _i32]),
    "todg_map_slots": (_i32, [_i32, _i32, _i32, _p, _p]),
    "todg_pack_map": (_i32, [_p, _i32, _i64, _i32, _i32, _p, _p]),
    "todg_gather_accum_fused": (
        _i32,
        [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _i32, _p, _i64, _i64, _p, _i32, _i32, _i32, _p],
    ),
    "todg_gather_accum_nearest": (
        _i32,
        [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _i32, _p, _i64, _i64, _p, _i32, _i32, _i32, _p],
    ),
    "todg_gather_accum_bicubic": (
        _i32,
        [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _i32, _p, _i64, _i64, _p, _i32, _i32, _i32, _p],
    ),
    "todg_interp_points": (_i32, [_i32, _i64, _p, _i64, _p, _p, _p, _p]),
    "todg_beam_bell_workspace": (ctypes.c_size_t, [_i64, _i32]),
    "todg_beam_bell": (_i32, [_p, _p, _i64, _i32, _p, _p, _p]),
    "todg_beam_bell_noalloc": (_i32, [_p, _p, _i64, _i32, _p, _p]),
    "todg_ring_sums": (_i32, [_p, _i32, _i32, _p, _p, _i64, _p, ctypes.c_double, ctypes.c_double, _p, _p]),
    "todg_corr_features": (_i32, [_i64, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "todg_corr_workspace_floats": (ctypes.c_size_t, [_i64]),
    "todg_corr_forward": (_i32, [_p, _p, _i64, _p, _p, _p]),
    "todg_corr_loss_grad": (_i32, [_p, _p, _p, _p, _i64, _p, _p, _p, _p, _p]),
    "todg_corr_adam": (_i32, [_p, _p, _p, _p, _i32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, _p]),
    "todg_gather_accum_learned": (
        _i32,
        [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _i32, _p, _i64, _i64, _p, _i32, _i32, _p,
         ctypes.c_float, _i64, _p, _p],
    ),
    "todg_corr_gemm_bf16": (_i32, [_i32, _i32, _i32, _p, _i32, _i32, _p, _i32, _i32, _p, _i32,
                                    ctypes.c_float, ctypes.c_float, _p]),
    "todg_rotation_params": (_i32, [_p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p]),
    "todg_fcorr_channels": (_i32, [_i64, _p, _p, _p, _p, _i64, _p]),
    "todg_gather_accum_fcorr": (
        _i32,
        [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _p, _i32, _p, _i64, _i64, _p, _i32, _i32, _p, _p],
    ),
    "todg_bilinear_indices": (
        _i32,
        [_p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _i64, _p, _p, _p],
    ),
    "todg_nearest_indices": (
        _i32,
        [_p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _i64, _p, _p],
    ),
    "todg_ring_table": (_i32, [_i64, _p, _p, _p, _p, _p, _p]),
    "todg_gather_accum_batch": (
        _i32, [_i32, _p, _i32, _i64, _i64, _p, _i32, _i32, _i32, _i32, _p, _p, _i32, _p],
    ),
    "todg_gather_accum": (
        _i32,
        [_i32, _p, _p, _p, _p, _p, _p, _p, _i64, _p, _i32, _p, _i64, _i64, _p, _i32, _i32, _i32, _p,
         _p],
    ),
    "todg_smlp_channels": (_i32, [_i64, _p, _p, _p, _p, _i64, _p]),
    "todg_smlp_features": (_i32, [_i64, _p, _p, _p, _p]),
    "todg_smlp_workspace_floats": (ctypes.c_size_t, [_i64]),
    "todg_smlp_forward": (_i32, [_p, _p, _i64, _p, _p, _p]),
    "todg_smlp_backward": (_i32, [_p, _p, _i64, _p, _p, _p, _p]),
    "todg_smlp_head": (_i32, [_i64, _p, _p, _p, _p, _p]),
    "todg_smlp_head_backward": (_i32, [_i64, _p, _p, _p, _p, _p]),
    "todg_csr_rows2_parts": (ctypes.c_size_t, [_i64]),
    "todg_csr_rows2": (_i32, [_i64, _p, _p, _p, _p, _p, _p, _p, ctypes.c_double, _p, _p, _p, _p, _p]),
    "todg_csr_adjoint2": (_i32, [_i64, _p, _p, _p, _p, _p, _p, _p, _p]),
    "todg_smlp_adam": (_i32, [_p, _p, _p, _p, _i32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, _p]),
    "todg_fp_dump_bytes": (ctypes.c_size_t, [_i64, _i64]),
    "todg_fp_dump": (
        _i32,
        [_i32, _p, _p, _p, _p, _p, _p, _p, _i64, _p, _i32, _p, _i64, _i64, _p, _i32, _i32, _p, _p],
    ),
}


class TodgError(RuntimeError):
    """Error reported by libtodgpu (status code and message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libtodgpu error {code}: {msg}")
        self.code = code


def _load(path: str) -> ctypes.CDLL:
    with _lock:
        if path in _handles:
            return _handles[path]
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it with "
                "`python -m tod_generation_mapbased_beam_amd.build` (or __graft_entry__.build())"
            )
        handle = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        v = handle.todg_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"{os.path.basename(path)} ABI {v} != expected {ABI_VERSION}")
        _handles[path] = handle
        return handle


def lib() -> ctypes.CDLL:
    """Load (once) and return the library; raises if it is not built."""
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def checked():
    """Within the block every call of the package goes to the checked build
    (libtodgpu_checked.so); read its violation counters with
    check_counts().  Not thread-safe: it swaps the library the whole process
    uses.  Device state the library owns (ring tables, queues) is per
    library, so it is built again on first use."""
    global _lib
    prev = lib()
    _lib = _load(CHECKED_LIB_PATH)
    try:
        yield _lib
    finally:
        _lib = prev


def check_counts(reset: bool = True) -> dict:
    """Index-check violations counted by the checked build since the last
    reset, by kind (CHECK_KINDS); synchronises the device."""
    buf = (ctypes.c_ulonglong * len(CHECK_KINDS))()
    h = _load(CHECKED_LIB_PATH)
    rc = h.todg_check_counts(buf, len(CHECK_KINDS), 1 if reset else 0)
    if rc < 0:
        raise TodgError(rc, _last_error(h))
    return {k: int(buf[i]) for i, k in enumerate(CHECK_KINDS)}


def _last_error(h: ctypes.CDLL) -> str:
    buf = ctypes.create_string_buffer(1024)
    h.todg_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def last_error() -> str:
    return _last_error(lib())


def check(rc: int) -> int:
    """Raise on a negative status: ValueError for argument errors (as the
    reference raises for a bad interp_mode), TodgError for runtime errors."""
    if rc < 0:
        msg = last_error()
        if rc == TODG_ERR_ARG:
            raise ValueError(msg)
        raise TodgError(rc, msg)
    return rc


def ptr(t) -> int:
    """Device (or host) address of a tensor, 0 for None."""
    if t is None:
        return 0
    return int(t.data_ptr())
