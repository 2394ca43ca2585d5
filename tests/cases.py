"""Seeded input cases shared by the CPU and GPU tests."""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from tod_generation_mapbased_beam_amd import inputs
from tod_generation_mapbased_beam_amd import tod_rotations as R


def random_unit(shape, rng):
    v = rng.standard_normal(shape)
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    return v.astype(np.float32)


def random_rot_params(B, S, rng, pole_safe=True):
    """Random beam vectors and per-sample rotation operands (the shape of the
    reference tests' fixtures, tests/test_tod_bilinear.py:59-96)."""
    vec_orig = random_unit((S, 3), rng)
    a1 = rng.uniform(0.0, math.pi, B).astype(np.float32)
    axes = random_unit((B, 3), rng)
    cos_a = np.cos(a1).astype(np.float32)
    sin_a = np.sin(a1).astype(np.float32)
    if pole_safe:
        th = rng.uniform(0.1, math.pi - 0.1, B)
        ph = rng.uniform(0.0, 2 * math.pi, B)
        ax_pts = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)],
                          axis=-1).astype(np.float32)
    else:
        ax_pts = random_unit((B, 3), rng)
    a2 = rng.uniform(0.0, 2 * math.pi, B).astype(np.float32)
    cos_p = np.cos(a2).astype(np.float32)
    sin_p = np.sin(a2).astype(np.float32)
    return vec_orig, (axes, cos_a, sin_a, ax_pts, cos_p, sin_p)


def recenter_params(ax_pts):
    """Rotation operands taking vec_orig = [0,0,1] onto each ax_pts[b] with
    no roll (tests/test_tod_bilinear.py:114-146)."""
    B = ax_pts.shape[0]
    axes = np.zeros((B, 3), np.float32)
    cos_a = np.ones(B, np.float32)
    sin_a = np.zeros(B, np.float32)
    for b in range(B):
        tx, ty, tz = (float(v) for v in ax_pts[b])
        ang = math.acos(max(-1.0, min(1.0, tz)))
        if ang < 1e-8:
            axes[b] = [1.0, 0.0, 0.0]
        else:
            ax = np.array([-ty, tx, 0.0])
            nrm = np.linalg.norm(ax)
            axes[b] = (ax / nrm).astype(np.float32) if nrm > 1e-10 else [1.0, 0.0, 0.0]
        cos_a[b] = math.cos(ang)
        sin_a[b] = math.sin(ang)
    return axes, cos_a, sin_a, ax_pts.astype(np.float32), np.ones(B, np.float32), np.zeros(B, np.float32)


@lru_cache(maxsize=4)
def small_config(n_samples=10000, nside=128, seed=0):
    """The small production-like case: nside-128 I/Q/U map from this
    package's harmonic synthesis (seed 0), 60' Gaussian beam at nside_beam
    512 with a 0.99 power cut, 10 Hz ring scan."""
    vec, bv, _ = inputs.gaussian_beam(512, 60.0, 0.99)
    th, ph, ps = inputs.ring_scan(n_samples, 10.0)
    params = R.rotation_params_for_pointing(th, ph, ps)
    mp = inputs.synfast_iqu(nside, seed=seed).astype(np.float32)
    return dict(vec=vec, bv=bv, params=params, map=mp, nside=nside, B=n_samples,
                S=vec.shape[0], theta=th, phi=ph, psi=ps)


def rel_err(got, ref):
    """Max |got - ref| per component, normalised by that component's RMS."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    rms = np.sqrt(np.mean(ref ** 2, axis=-1, keepdims=True))
    rms = np.where(rms > 0, rms, 1.0)
    return float(np.max(np.abs(got - ref) / rms))
