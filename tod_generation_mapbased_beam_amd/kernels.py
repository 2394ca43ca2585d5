"""Kernel-level entry points with the reference's signatures.

``_gather_accum_fused_jit`` and ``_gather_accum_nearest_jit`` take the same
arguments in the same order as tod_bilinear.py:205-222 and
tod_nearest.py:21-38 and, like them, ACCUMULATE into ``tod`` (``+=``).
Arguments may be numpy arrays (uploaded; ``tod`` is written back in place)
or torch tensors (device tensors are used in place, no copies).
``mp_stacked`` may also be a :class:`DeviceMap` that is already packed.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .device import device_map_for, stream_handle, to_dev

# variant 0: footprint kernel (per-sample LDS weight grid, 4 waves/SIMD,
# default); 8: the same compiled for 2 waves/SIMD; 1, 2, 4: direct kernel
# with that many samples per wavefront.
DEFAULT_VARIANT = 0


def _map_channels(mp_stacked) -> int:
    """C of a map argument (a DeviceMap, or a (C, npix) / (npix,) array)."""
    if hasattr(mp_stacked, "packed") and hasattr(mp_stacked, "C"):
        return int(mp_stacked.C)
    shape = tuple(mp_stacked.shape)
    return 1 if len(shape) == 1 else int(shape[0])


def check_tod(tod, C: int, B: int, name: str = "tod") -> None:
    """The reference's accumulator contract (tod_bilinear.py:220, (C, B)
    float64, updated in place): a numpy array, or a contiguous CUDA tensor on
    the current device.  Checked before anything touches the device."""
    if isinstance(tod, np.ndarray):
        if tod.dtype != np.float64 or tod.shape != (C, B):
            raise ValueError(f"{name} must be float64 of shape ({C}, {B})")
        return
    if not isinstance(tod, torch.Tensor):
        raise ValueError(f"{name} must be a numpy array or a torch tensor")
    if tod.dtype != torch.float64 or tuple(tod.shape) != (C, B):
        raise ValueError(f"{name} must be a float64 tensor of shape ({C}, {B})")
    if not tod.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if tod.device.type != "cuda":
        raise ValueError(f"{name} must live on the GPU (got a {tod.device.type} tensor); "
                         "pass a numpy array to have it copied in and out")
    if tod.device.index is not None and tod.device.index != torch.cuda.current_device():
        raise ValueError(f"{name} is on cuda:{tod.device.index}, not the current device "
                         f"cuda:{torch.cuda.current_device()}")


def _run(entry, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, mp_stacked,
         beam_vals, B, S, tod, c_q, c_u, variant, mode=None, status=None):
    B = int(B)
    S = int(S)
    c_q = int(c_q)
    c_u = int(c_u)
    check_tod(tod, _map_channels(mp_stacked), B)
    if status is not None:
        check_status(status)
    dmap = device_map_for(mp_stacked, c_q, c_u)
    if dmap.nside != int(nside):
        raise ValueError(f"map nside {dmap.nside} != nside {nside}")
    f32 = torch.float32
    ops = [to_dev(a, f32) for a in (vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, beam_vals)]
    v, ax, ca, sa, bp, cp, sp, bv = ops
    if v.numel() < 3 * S or bv.numel() < S:
        raise ValueError("vec_orig / beam_vals shorter than S")
    if ax.numel() < 3 * B or bp.numel() < 3 * B or min(ca.numel(), sa.numel(), cp.numel(), sp.numel()) < B:
        raise ValueError("rotation operands shorter than B")
    host_tod = None
    if isinstance(tod, np.ndarray):
        host_tod = tod
        t = torch.from_numpy(np.ascontiguousarray(tod)).to(v.device)
    else:
        t = tod
    args = (_lib.ptr(v), _lib.ptr(ax), _lib.ptr(ca), _lib.ptr(sa), _lib.ptr(bp), _lib.ptr(cp),
            _lib.ptr(sp), int(nside), _lib.ptr(dmap.packed), dmap.C, _lib.ptr(bv), B, S,
            _lib.ptr(t), c_q, c_u, int(variant))
    if mode is None:
        _lib.check(entry(*args, stream_handle()))
    else:
        _lib.check(entry(int(mode), *args, _lib.ptr(status), stream_handle()))
    if host_tod is not None:
        host_tod[...] = t.cpu().numpy()
    return tod


def check_status(status) -> None:
    if not (isinstance(status, torch.Tensor) and status.dtype == torch.int32 and status.numel() >= 1
            and status.device.type == "cuda"):
        raise ValueError("status must be a CUDA int32 tensor with at least one element")


MODES = {"bilinear": 0, "nearest": 1, "bicubic": 2}


def gather_accum(interp_mode, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, mp_stacked,
                 beam_vals, B, S, tod, c_q=-1, c_u=-1, variant=DEFAULT_VARIANT, status=None):
    """Any of the three kernels (``todg_gather_accum``), accumulating into
    ``tod``; a sample with a non-finite rotation operand is skipped (its tod
    entries keep their values) and sets bit 0 of ``status`` (a CUDA int32
    tensor, or None)."""
    if interp_mode not in MODES:
        raise ValueError(f"interp_mode {interp_mode!r} not available; use 'bilinear', 'nearest' "
                         "or 'bicubic'")
    return _run(_lib.lib().todg_gather_accum, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p,
                nside, mp_stacked, beam_vals, B, S, tod, c_q, c_u, variant,
                mode=MODES[interp_mode], status=status)


def _gather_accum_fused_jit(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside,
                            mp_stacked, beam_vals, B, S, tod, c_q=-1, c_u=-1, variant=DEFAULT_VARIANT):
    """Bilinear + spin-2 beam TOD on the GPU (tod_bilinear.py:205-477)."""
    return _run(_lib.lib().todg_gather_accum_fused, vec_orig, axes, cos_a, sin_a, ax_pts,
                cos_p, sin_p, nside, mp_stacked, beam_vals, B, S, tod, c_q, c_u, variant)


def _gather_accum_nearest_jit(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside,
                              mp_stacked, beam_vals, B, S, tod, c_q=-1, c_u=-1,
                              variant=DEFAULT_VARIANT):
    """Nearest-pixel + spin-2 beam TOD on the GPU (tod_nearest.py:21-158)."""
    return _run(_lib.lib().todg_gather_accum_nearest, vec_orig, axes, cos_a, sin_a, ax_pts,
                cos_p, sin_p, nside, mp_stacked, beam_vals, B, S, tod, c_q, c_u, variant)


def _gather_accum_bicubic_jit(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside,
                              mp_stacked, beam_vals, B, S, tod, c_q=-1, c_u=-1,
                              variant=DEFAULT_VARIANT):
    """Bicubic-judge + spin-2 beam TOD on the GPU (same arguments as the
    bilinear kernel; README.md:120 describes the reference's bicubic branch)."""
    if int(variant) not in (0, 8):
        variant = DEFAULT_VARIANT
    return _run(_lib.lib().todg_gather_accum_bicubic, vec_orig, axes, cos_a, sin_a, ax_pts,
                cos_p, sin_p, nside, mp_stacked, beam_vals, B, S, tod, c_q, c_u, variant)


INTERP_MODES = {"bilinear": 0, "bicubic": 1}


def interp_points(nside, field, z, phi, mode="bilinear"):
    """Values of one RING-ordered map field (npix,) at directions (z, phi):
    bilinear (numba_healpy.py:358-499) or the bicubic judge.  Returns a
    float64 torch tensor on the device (numpy in, numpy out)."""
    if mode not in INTERP_MODES:
        raise ValueError(f"mode {mode!r} not available; use 'bilinear' or 'bicubic'")
    host = isinstance(z, np.ndarray)
    m = to_dev(field, torch.float32).reshape(-1)
    if m.numel() != 12 * int(nside) ** 2:
        raise ValueError("field must hold 12 * nside**2 values")
    zz = to_dev(z, torch.float64).reshape(-1)
    pp = to_dev(phi, torch.float64).reshape(-1)
    if zz.numel() != pp.numel():
        raise ValueError("z and phi differ in length")
    out = torch.empty(zz.numel(), dtype=torch.float64, device=zz.device)
    _lib.check(_lib.lib().todg_interp_points(INTERP_MODES[mode], int(nside), _lib.ptr(m),
                                             zz.numel(), _lib.ptr(zz), _lib.ptr(pp), _lib.ptr(out),
                                             stream_handle()))
    return out.cpu().numpy() if host else out


def bilinear_indices(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, B, S, b0, nb):
    """(nb, S, 4) neighbour pixels and weights exactly as the fused kernel
    evaluates them on the device; returned as numpy int64 / float64."""
    f32 = torch.float32
    ops = [to_dev(a, f32) for a in (vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p)]
    dev = ops[0].device
    pix = torch.empty((int(nb), int(S), 4), dtype=torch.int64, device=dev)
    wgt = torch.empty((int(nb), int(S), 4), dtype=torch.float64, device=dev)
    _lib.check(
        _lib.lib().todg_bilinear_indices(
            *[_lib.ptr(o) for o in ops], int(nside), int(B), int(S), int(b0), int(nb),
            _lib.ptr(pix), _lib.ptr(wgt), stream_handle(),
        )
    )
    return pix.cpu().numpy(), wgt.cpu().numpy()


def fp_dump(interp_mode, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, mp_stacked,
            beam_vals, B, S, c_q=-1, c_u=-1, host=True):
    """Run the production footprint kernel's index-dump build over all B
    samples: returns (pix (B, S, 4) int64, w (B, S, 4) float64, path (B, S)
    uint8, tod (C, B) float64) -- the pixels each beam pixel was deposited
    into (bilinear neighbours in the reference's order; nearest: column 0),
    the deposited beam value x weight, the path that took it (0 certified
    fast path, 1 exact path) and the TOD of the same launch.  numpy arrays,
    or device tensors (views of one buffer) with host=False."""
    if interp_mode not in ("bilinear", "nearest"):
        raise ValueError("index dumps exist for 'bilinear' and 'nearest'")
    B = int(B)
    S = int(S)
    c_q = int(c_q)
    c_u = int(c_u)
    dmap = device_map_for(mp_stacked, c_q, c_u)
    f32 = torch.float32
    ops = [to_dev(a, f32) for a in (vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, beam_vals)]
    v, ax, ca, sa, bp, cp, sp, bv = ops
    nbytes = int(_lib.lib().todg_fp_dump_bytes(B, S))
    buf = torch.empty(max(nbytes, 8) // 8 + 1, dtype=torch.int64, device=v.device)
    tod = torch.zeros((dmap.C, B), dtype=torch.float64, device=v.device)
    _lib.check(_lib.lib().todg_fp_dump(
        MODES[interp_mode], _lib.ptr(v), _lib.ptr(ax), _lib.ptr(ca), _lib.ptr(sa), _lib.ptr(bp),
        _lib.ptr(cp), _lib.ptr(sp), int(nside), _lib.ptr(dmap.packed), dmap.C, _lib.ptr(bv), B, S,
        _lib.ptr(tod), c_q, c_u, _lib.ptr(buf), stream_handle()))
    n = B * S
    if not host:
        return (buf[: 4 * n].view(B, S, 4), buf[4 * n: 8 * n].view(torch.float64).view(B, S, 4),
                buf.view(torch.uint8)[64 * n: 65 * n].view(B, S), tod)
    raw = buf.cpu().numpy()
    pix = raw[: 4 * n].reshape(B, S, 4)
    w = raw[4 * n: 8 * n].view(np.float64).reshape(B, S, 4)
    path = raw.view(np.uint8)[64 * n: 65 * n].reshape(B, S)
    return pix, w, path, tod.cpu().numpy()


def nearest_indices(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, B, S, b0, nb):
    """(nb, S) nearest pixels exactly as the nearest kernel picks them."""
    f32 = torch.float32
    ops = [to_dev(a, f32) for a in (vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p)]
    pix = torch.empty((int(nb), int(S)), dtype=torch.int64, device=ops[0].device)
    _lib.check(
        _lib.lib().todg_nearest_indices(
            *[_lib.ptr(o) for o in ops], int(nside), int(B), int(S), int(b0), int(nb),
            _lib.ptr(pix), stream_handle(),
        )
    )
    return pix.cpu().numpy()


def rotation_params(theta, phi, psi, centre):
    """The six float32 kernel operands of boresight pointings (float64, numpy
    or torch, B each) computed on the GPU (``todg_rotation_params``: the
    driver's casts, precompute_rotation_vector_batch and _rotation_params,
    tod_rotations.py:219-333).  ``centre``: tod_rotations.beam_centre(...)."""
    sizes = [int(x.numel()) if isinstance(x, torch.Tensor) else int(np.asarray(x).size) for x in (theta, phi, psi)]
    if len(set(sizes)) != 1:
        raise ValueError("theta, phi and psi must have the same length")
    ctr = np.ascontiguousarray(centre, dtype=np.float64).reshape(-1)
    if ctr.shape != (6,) or not np.all(np.isfinite(ctr)):
        raise ValueError("centre must hold 6 finite values (c, n_c)")
    f64 = torch.float64
    th, ph, ps = (to_dev(x, f64).reshape(-1) for x in (theta, phi, psi))
    B = sizes[0]
    f32 = torch.float32
    dev = th.device
    axes = torch.empty((B, 3), dtype=f32, device=dev)
    ax_pts = torch.empty((B, 3), dtype=f32, device=dev)
    cos_a, sin_a, cos_p, sin_p = (torch.empty(B, dtype=f32, device=dev) for _ in range(4))
    _lib.check(_lib.lib().todg_rotation_params(
        _lib.ptr(th), _lib.ptr(ph), _lib.ptr(ps), B, ctr.ctypes.data, _lib.ptr(axes), _lib.ptr(cos_a),
        _lib.ptr(sin_a), _lib.ptr(ax_pts), _lib.ptr(cos_p), _lib.ptr(sin_p), stream_handle()))
    return axes, cos_a, sin_a, ax_pts, cos_p, sin_p
