"""Drop-in replacement for the reference's dispatcher ``beam_tod_batch``
(tod_core.py:52-180): same signature, same channel bookkeeping, same errors,
same return contract (a dict mapping each Stokes component to a float32
``(B,)`` array), with the hot loop on the GPU.

Inputs may be numpy (the reference's types; results come back as numpy) or
torch tensors (results stay on the device as torch tensors).
"""

from __future__ import annotations

import torch

from . import autotune, kernels
from .device import device, to_dev
from .tod_rotations import _rotation_params_split


def _qu_slots(comp_indices):
    """Q and U channel positions (tod_core.py:105-112)."""
    c_q = -1
    c_u = -1
    for ci, comp in enumerate(comp_indices):
        if comp == 1:
            c_q = ci
        elif comp == 2:
            c_u = ci
    return c_q, c_u


def beam_tod_batch(nside, mp, data, rot_vecs, phi_b, theta_b, psis_b, interp_mode="bilinear",
                   variant=None, params=None):
    """Beam-weighted sky samples of one beam entry for a batch of pointings.

    Args mirror tod_core.py:52-96: ``data`` needs ``vec_orig (S,3)``,
    ``beam_vals (S,)``, ``comp_indices`` and (production) ``mp_stacked
    (C, npix)`` float32 (numpy, torch, or a packed ``DeviceMap``).
    ``interp_mode`` is ``"bilinear"`` or ``"nearest"`` (tod_core.py:114-119),
    or ``"bicubic"`` (the reference's bicubic branch, README.md:120, 317);
    anything else raises ``ValueError`` as the reference does.  ``params``:
    the six kernel operands already computed for this batch (e.g. on the GPU
    by ``kernels.rotation_params``); rot_vecs / phi_b / theta_b / psis_b are
    then not used (pass None) and the result rows stay on the device.

    When ``mp_stacked`` is absent the reference falls back to a split
    healpy pipeline with bilinear weights and no spin-2 (tod_core.py:169-180);
    here the maps ``mp[c]`` are stacked and the bilinear kernel runs without
    spin-2, which is the same computation minus the float32 rounding of the
    intermediate rotated vectors.
    """
    vec_orig = data["vec_orig"]
    beam_vals = data["beam_vals"]
    comp_indices = list(data["comp_indices"])
    B = int(phi_b.shape[0]) if params is None else int(params[1].shape[0])
    S = int(vec_orig.shape[0])
    C = len(comp_indices)
    mp_stacked = data.get("mp_stacked")

    c_q, c_u = _qu_slots(comp_indices)
    if interp_mode not in ("nearest", "bilinear", "bicubic"):
        raise ValueError(
            f"interp_mode {interp_mode!r} not available; use 'bilinear', 'nearest' or 'bicubic'"
        )

    on_device = params is not None or isinstance(rot_vecs, torch.Tensor)
    if params is None:
        if on_device:
            params = rotation_params_torch(rot_vecs, phi_b, theta_b, psis_b)
        else:
            params = _rotation_params_split(rot_vecs, phi_b, theta_b, psis_b)

    if mp_stacked is None:
        mp_stacked = torch.stack([to_dev(mp[c], torch.float32) for c in comp_indices])
        c_q = c_u = -1
        if interp_mode == "nearest":
            interp_mode = "bilinear"

    if variant is None:  # the autotuner's choice for this workload shape, if any
        variant = autotune.tuned_variant(int(nside), S, interp_mode, C, c_q >= 0 and c_u >= 0,
                                         kernels.DEFAULT_VARIANT)
    tod_arr = torch.zeros((C, B), dtype=torch.float64, device=device())
    fn = {"nearest": kernels._gather_accum_nearest_jit,
          "bilinear": kernels._gather_accum_fused_jit,
          "bicubic": kernels._gather_accum_bicubic_jit}[interp_mode]
    fn(vec_orig, *params, nside, mp_stacked, beam_vals, B, S, tod_arr, c_q, c_u, variant=variant)
    out32 = tod_arr.to(torch.float32)
    if on_device:
        return {comp: out32[i] for i, comp in enumerate(comp_indices)}
    host = out32.cpu().numpy()
    return {comp: host[i].copy() for i, comp in enumerate(comp_indices)}


def rotation_params_torch(rot_vecs, phi_b, theta_b, psis_b):
    """Device (torch float32) version of ``_rotation_params`` for pointings
    that already live in HBM."""
    dev = device()
    rv = to_dev(rot_vecs, torch.float64)
    ang = torch.linalg.norm(rv, dim=-1).to(torch.float32)
    ok = ang > 1e-10
    axes = (rv / torch.where(ok, ang, torch.ones_like(ang)).to(rv.dtype)[:, None]).to(torch.float32)
    axes = torch.where(ok[:, None], axes, torch.zeros_like(axes))
    ph = to_dev(phi_b, torch.float32)
    th = to_dev(theta_b, torch.float32)
    ps = to_dev(psis_b, torch.float32)
    st = torch.sin(th)
    ax_pts = torch.stack([st * torch.cos(ph), st * torch.sin(ph), torch.cos(th)], dim=-1)
    return (axes.contiguous(), torch.cos(ang), torch.sin(ang), ax_pts.contiguous(),
            torch.cos(ps), torch.sin(ps))
