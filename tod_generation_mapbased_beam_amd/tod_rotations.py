"""Host-side rotation preparation (numpy), the counterpart of the reference's
``precompute_rotation_vector_batch`` and ``_rotation_params``
(tod_rotations.py:219-333).

These run once per batch on the host in the reference too; their outputs (six
float32 arrays per batch) are the per-sample operands of the GPU kernels.
Parity tests feed the *same* arrays to the GPU and to the CPU oracle, so the
bits of numpy's float32 cos/sin do not matter for kernel parity.
"""

from __future__ import annotations

import numpy as np


def _unit(phi, theta):
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], axis=-1)


def beam_centre(ra, dec, center_idx=None) -> np.ndarray:
    """(6,) float64: the beam centre's unit vector c and its local north n_c
    for beam-frame offset grids ``ra``/``dec`` (centre element ``center_idx``,
    default ``(H//2, W//2)``), as precompute_rotation_vector_batch builds
    them (tod_rotations.py:219-303)."""
    ra = np.asarray(ra, dtype=np.float64)
    dec = np.asarray(dec, dtype=np.float64)
    if center_idx is None:
        center_idx = (ra.shape[0] // 2, ra.shape[1] // 2)
    phi_c = ra[center_idx]
    theta_c = np.pi / 2 - dec[center_idx]
    c = _unit(phi_c, theta_c)
    n_c = np.array(
        [np.cos(phi_c) * np.cos(theta_c), np.sin(phi_c) * np.cos(theta_c), -np.sin(theta_c)]
    )
    return np.concatenate([c, n_c])


def precompute_rotation_vector_batch(ra, dec, phi_batch, theta_batch, center_idx=None):
    """Rodrigues vectors taking the beam centre to each boresight, and the
    parallactic offsets beta in [0, 2pi).  Same contract as
    tod_rotations.py:219-303: ``ra``/``dec`` are (H, W) beam-frame offset
    grids whose ``center_idx`` element (default ``(H//2, W//2)``) is the beam
    centre; returns ``rot_vector (B, 3)`` float64 and ``beta (B,)`` float64.

    A HEALPix beam centred on the north pole is ``ra=[[0.]], dec=[[pi/2]]``.
    """
    ctr = beam_centre(ra, dec, center_idx)
    phi_b = np.asarray(phi_batch)
    theta_b = np.asarray(theta_batch)
    c = ctr[None, :3]  # (1, 3) beam centre
    t = _unit(phi_b, theta_b)  # (B, 3) boresight

    # Local north at the beam centre and local north/east at the boresight.
    n_c = ctr[3:]
    n_t = np.stack(
        [np.cos(phi_b) * np.cos(theta_b), np.sin(phi_b) * np.cos(theta_b), -np.sin(theta_b)],
        axis=-1,
    )
    e_t = np.stack([-np.sin(phi_b), np.cos(phi_b), np.zeros_like(phi_b)], axis=-1)

    k = np.cross(c, t)
    kn = np.linalg.norm(k, axis=-1, keepdims=True)
    k = np.where(kn > 1e-10, k / np.where(kn > 0, kn, 1.0), 0.0)
    ang = np.arccos(np.clip(np.sum(c * t, axis=-1), -1.0, 1.0))
    rot_vector = k * ang[:, None]

    # Rotate the centre's north with the same Rodrigues rotation, then read
    # its bearing in the boresight's (north, east) frame.
    ca = np.cos(ang)[:, None]
    v = n_c[None, :]
    w = v * ca + np.cross(k, v) * np.sin(ang)[:, None] + k * np.sum(k * v, axis=-1, keepdims=True) * (1.0 - ca)
    beta = np.arctan2(np.sum(w * e_t, axis=-1), np.sum(w * n_t, axis=-1))
    beta = np.where(beta < 0, beta + 2 * np.pi, beta)
    return rot_vector, beta


def _rotation_params(rot_vecs, phi_b, theta_b, psis_b):
    """Per-sample float32 kernel operands, as tod_rotations.py:306-333:
    ``axes (B,3), cos_a, sin_a (B,), ax_pts (B,3), cos_p, sin_p (B,)``.

    The rotation angle is the float32 norm of the Rodrigues vector; the axis is
    zero where that angle is <= 1e-10.  The boresight vector and the roll
    angle's cos/sin are evaluated from float32 copies of the angles.
    """
    rot_vecs = np.asarray(rot_vecs)
    ang = np.linalg.norm(rot_vecs, axis=-1).astype(np.float32)
    ok = ang > np.float32(1e-10)
    axes = (rot_vecs / np.where(ok[:, None], ang[:, None], np.float32(1.0))).astype(np.float32)
    axes = np.where(ok[:, None], axes, np.float32(0.0))
    cos_a = np.cos(ang)
    sin_a = np.sin(ang)

    ph = np.asarray(phi_b, dtype=np.float32)
    th = np.asarray(theta_b, dtype=np.float32)
    ps = np.asarray(psis_b, dtype=np.float32)
    st = np.sin(th)
    ax_pts = np.stack([st * np.cos(ph), st * np.sin(ph), np.cos(th)], axis=-1)
    return axes, cos_a, sin_a, ax_pts, np.cos(ps), np.sin(ps)


# Samples below which _rotation_params_split does not split over threads.
SPLIT_MIN = 1 << 16
_pool = None
_pool_pid = None


def _rotation_params_split(rot_vecs, phi_b, theta_b, psis_b, workers=None):
    """_rotation_params over pieces of the batch computed by a shared pool
    of up to 16 host threads (numpy's loops release the GIL).  Every
    operation is per sample, so the result equals _rotation_params' bit for
    bit.  The pool belongs to the process that made it: a forked child
    (whose copy has no threads) makes its own."""
    global _pool, _pool_pid
    import os
    from concurrent.futures import ThreadPoolExecutor

    rot_vecs = np.asarray(rot_vecs)
    n = int(rot_vecs.shape[0])
    w = max(1, min(16, os.cpu_count() or 1)) if workers is None else int(workers)
    parts = max(1, min(w, n // SPLIT_MIN))
    if parts == 1:
        return _rotation_params(rot_vecs, phi_b, theta_b, psis_b)
    if _pool is None or _pool_pid != os.getpid():
        _pool = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))
        _pool_pid = os.getpid()
    ph, th, ps = (np.asarray(x) for x in (phi_b, theta_b, psis_b))
    edges = [(n * i) // parts for i in range(parts + 1)]
    pieces = list(_pool.map(lambda i: _rotation_params(rot_vecs[edges[i]:edges[i + 1]], ph[edges[i]:edges[i + 1]],
                                                        th[edges[i]:edges[i + 1]], ps[edges[i]:edges[i + 1]]),
                            range(parts)))
    return tuple(np.concatenate([p[j] for p in pieces]) for j in range(6))


def rotation_params_for_pointing(theta, phi, psi, ra=None, dec=None, center_idx=None):
    """Convenience: boresight (theta, phi, psi) -> the six kernel operands,
    following the driver's casts (sample_based_tod_generation_gridint.py:
    277-281): pointings are cast to float32 first, psi_b - beta combines the
    roll.  Defaults to a HEALPix beam centred on the north pole."""
    if ra is None:
        ra = np.zeros((1, 1))
        dec = np.full((1, 1), np.pi / 2)
        center_idx = (0, 0)
    th = np.asarray(theta).astype(np.float32)
    ph = np.asarray(phi).astype(np.float32)
    ps = np.asarray(psi).astype(np.float32)
    rot, beta = precompute_rotation_vector_batch(ra, dec, ph, th, center_idx=center_idx)
    psis = -beta + ps
    return _rotation_params(rot, ph, th, psis)
