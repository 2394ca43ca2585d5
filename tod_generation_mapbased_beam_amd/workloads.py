"""Synthetic multi-detector workload: many detectors in one persistent launch.

64 detectors, each with a seeded elliptical Gaussian beam on an nside-4096
HEALPix grid (FWHM uniform in 25-40', ellipticity uniform in 0-0.2, random
orientation), a power cut uniform in 0.97-0.999 and the tail holding the last
3 % of the power k-means-clustered as in beam_cluster.py:215-397, so kept
pixel counts are uneven.  Detectors sit at focal-plane offsets up to 2 deg
from the boresight (the ring scan's opening angle and spin phase shifted per
detector) and scan chunks of uniform 2e5 - 2e6 samples at 20 Hz on an
nside-2048 I/Q/U map, one shared I/Q/U beam entry per detector (spin-2 on).
"""

from __future__ import annotations

import math
import time

import numpy as np

from . import inputs
from . import tod_rotations as R
from .beam_cluster import cluster_beam_pixels


def batch_detectors(n_det: int = 64, seed: int = 0, nside_beam: int = 4096, tail_fraction: float = 0.03,
                    tail_clusters: int = 64, min_samples: int = 200_000, max_samples: int = 2_000_000,
                    fs: float = 20.0, max_offset_deg: float = 2.0, log=None):
    """Host-side inputs of the batch: a list of (vec_orig, beam_vals,
    rotation params) per detector, and a description dict."""
    rng = np.random.default_rng(seed)
    out = []
    desc = {"detectors": n_det, "S": [], "B": [], "fwhm": [], "ellipticity": [], "cut": []}
    t0 = time.time()
    for d in range(n_det):
        fwhm = rng.uniform(25.0, 40.0)
        ell = rng.uniform(0.0, 0.2)
        ang = rng.uniform(0.0, math.pi)
        cut = rng.uniform(0.97, 0.999)
        vec, bv, _ = inputs.gaussian_beam(nside_beam, fwhm, cut, ellipticity=ell, angle=ang)
        vec, bv, _ = cluster_beam_pixels(vec, bv, tail_clusters, tail_fraction=tail_fraction,
                                         random_state=42 + d)
        n = int(rng.integers(min_samples, max_samples + 1))
        off = rng.uniform(-max_offset_deg, max_offset_deg)
        phase = rng.uniform(0.0, 60.0)  # seconds of spin phase
        th, ph, ps = inputs.ring_scan(n, fs, opening_deg=45.0 + off, t0=3600.0 * d + phase)
        params = R.rotation_params_for_pointing(th, ph, ps)
        out.append((np.ascontiguousarray(vec, np.float32), np.ascontiguousarray(bv, np.float32), params))
        for k, v in (("S", int(bv.shape[0])), ("B", n), ("fwhm", fwhm), ("ellipticity", ell), ("cut", cut)):
            desc[k].append(v)
        if log and (d + 1) % 16 == 0:
            log(f"batch inputs: {d + 1}/{n_det} detectors ({time.time() - t0:.1f}s)")
    desc["bps"] = float(sum(b * s for b, s in zip(desc["B"], desc["S"])))
    return out, desc
