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


def _run(entry, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, mp_stacked,
         beam_vals, B, S, tod, c_q, c_u, variant):
    B = int(B)
    S = int(S)
    c_q = int(c_q)
    c_u = int(c_u)
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
        if tod.dtype != np.float64 or tod.shape != (dmap.C, B):
            raise ValueError(f"tod must be float64 of shape ({dmap.C}, {B})")
        host_tod = tod
        t = torch.from_numpy(np.ascontiguousarray(tod)).to(v.device)
    else:
        t = tod
        if t.dtype != torch.float64 or not t.is_contiguous() or tuple(t.shape) != (dmap.C, B):
            raise ValueError(f"tod must be a contiguous float64 tensor of shape ({dmap.C}, {B})")
    _lib.check(
        entry(
            _lib.ptr(v), _lib.ptr(ax), _lib.ptr(ca), _lib.ptr(sa), _lib.ptr(bp), _lib.ptr(cp),
            _lib.ptr(sp), int(nside), _lib.ptr(dmap.packed), dmap.C, _lib.ptr(bv), B, S,
            _lib.ptr(t), c_q, c_u, int(variant), stream_handle(),
        )
    )
    if host_tod is not None:
        host_tod[...] = t.cpu().numpy()
    return tod


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
