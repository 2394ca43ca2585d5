"""Synthetic multi-detector workload: many detectors in one persistent launch.

64 detectors, each with its own seeded beams on an nside-4096 HEALPix grid:
an elliptical Gaussian co-polar response (FWHM uniform in 25-40',
ellipticity uniform in 0-0.2, random orientation) and Q and U beams carrying
1 % cross-polar leakage of random orientation.  Each Stokes beam is its own
beam entry, as the reference's per-file beam dictionary holds them
(sample_based_tod_generation_gridint.py:92-160): power-cut (uniform 0.97-0.999
per detector) and normalised separately, its tail holding the last 3 % of
the power k-means-clustered (beam_cluster.py:215-397) with the cluster count
from the clustering calibration (tod_calibrate.py:392-530), so kept pixel
counts are uneven.  An entry holding Q alone or U alone gets no spin-2 in the
reference (tod_core.py:105-112, tod_bilinear.py:276): each entry is a
one-field (C = 1) launch on its own Stokes map.  Detectors sit at 2-D
focal-plane offsets up to 2 deg from the boresight of a 20 Hz ring scan, with
scan chunks uniform in 2e5 - 2e6 samples.
"""

from __future__ import annotations

import math
import time

import numpy as np

from . import inputs
from . import tod_rotations as R
from .beam_cluster import cluster_beam_pixels

STOKES = ("I", "Q", "U")


def focal_plane_pointing(theta, phi, psi, dx, dy):
    """(theta, phi, psi) of a detector at focal-plane offset (dx, dy)
    radians from the boresight (theta, phi, psi): the offset is laid out along
    the boresight's polarisation frame u1 = cos psi e_theta + sin psi e_phi,
    u2 = -sin psi e_theta + cos psi e_phi, as a great-circle step of length
    |(dx, dy)|; the detector's psi is the angle of u1, carried to the
    detector's direction, against its own meridian."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    psi = np.asarray(psi, dtype=np.float64)
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    b = np.stack([st * cp, st * sp, ct], -1)
    e_th = np.stack([ct * cp, ct * sp, -st], -1)
    e_ph = np.stack([-sp, cp, np.zeros_like(sp)], -1)
    u1 = np.cos(psi)[:, None] * e_th + np.sin(psi)[:, None] * e_ph
    u2 = -np.sin(psi)[:, None] * e_th + np.cos(psi)[:, None] * e_ph
    rho = math.hypot(dx, dy)
    alpha = math.atan2(dy, dx)
    t = math.cos(alpha) * u1 + math.sin(alpha) * u2  # unit tangent of the step
    d = math.cos(rho) * b + math.sin(rho) * t
    # u1 carried along the great circle b -> d: its component along the step
    # turns with the step, the rest is unchanged
    u1_d = u1 + (u1 * t).sum(-1, keepdims=True) * ((math.cos(rho) - 1.0) * t - math.sin(rho) * b)
    th_d = np.arccos(np.clip(d[:, 2], -1.0, 1.0))
    ph_d = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * math.pi)
    s2, c2 = np.sin(th_d), np.cos(th_d)
    e_th_d = np.stack([c2 * np.cos(ph_d), c2 * np.sin(ph_d), -s2], -1)
    e_ph_d = np.stack([-np.sin(ph_d), np.cos(ph_d), np.zeros_like(ph_d)], -1)
    psi_d = np.mod(np.arctan2((u1_d * e_ph_d).sum(-1), (u1_d * e_th_d).sum(-1)), 2 * math.pi)
    return th_d, ph_d, psi_d


def polarised_beams(nside_beam: int, fwhm: float, ellipticity: float, angle: float, cut: float,
                    leakage: float, rng):
    """I, Q and U beam entries (vec_orig, beam_vals) of one detector, each
    cut and normalised on its own: the co-polar Gaussian, and Q / U with a
    cross-polar leakage of relative amplitude ``leakage`` and random
    orientation, b (1 + leakage cos 2(phi - a))."""
    vals, vecs, _, ph = inputs.gaussian_beam_map(nside_beam, fwhm, ellipticity, angle)
    aq, au = rng.uniform(0.0, math.pi, 2)
    out = []
    for v in (vals, vals * (1.0 + leakage * np.cos(2.0 * (ph - aq))),
              vals * (1.0 + leakage * np.cos(2.0 * (ph - au)))):
        vec, bv, _ = inputs.select_beam(v, vecs, cut)
        out.append((vec, bv))
    return out


def calibrated_tail_clusters(entries, tail_fraction: float, error_threshold: float = 1e-3,
                             bell_lmax: int = 500, log=None, device=None) -> int:
    """Cluster count for the given tail fraction from the clustering
    calibration's (tail x K) sweep (beam_tools.calibrate_beam_clustering,
    the counterpart of tod_calibrate.py:392-530) over a few beam entries."""
    from .beam_tools import calibrate_beam_clustering

    data = {f"b{i}": {"vec_orig": v, "beam_vals": b} for i, (v, b) in enumerate(entries)}
    _, K, table = calibrate_beam_clustering(data, error_threshold=error_threshold, bell_lmax=bell_lmax,
                                            tail_fractions=(tail_fraction,), verbose=False, device=device)
    if log:
        log("clustering calibration: " + ", ".join(f"K={k}: {sp:.2f}x div {dv:.1e}"
                                                   for _, k, _, sp, dv in table) + f" -> K={K}")
    return int(K)


def batch_detectors(n_det: int = 64, seed: int = 0, nside_beam: int = 4096, tail_fraction: float = 0.03,
                    tail_clusters=None, min_samples: int = 200_000, max_samples: int = 2_000_000,
                    fs: float = 20.0, max_offset_deg: float = 2.0, leakage: float = 0.01,
                    calib_beams: int = 4, log=None, device="auto"):
    """Host-side inputs of the batch: per detector a dict with its three beam
    entries ``entries`` [(vec_orig, beam_vals)] * 3 (I, Q, U) and the six
    float32 rotation operands ``params`` of its pointing chunk, and a
    description dict.  ``tail_clusters=None`` takes the cluster count from
    the clustering calibration (needs the GPU); an int fixes it.  The tail
    k-means runs on ``device`` ("auto": the GPU when one is visible)."""
    if device == "auto":
        import torch

        device = "cuda" if torch.cuda.is_available() else None
    rng = np.random.default_rng(seed)
    t0 = time.time()
    draws = []
    for d in range(n_det):
        fwhm = rng.uniform(25.0, 40.0)
        ell = rng.uniform(0.0, 0.2)
        ang = rng.uniform(0.0, math.pi)
        cut = rng.uniform(0.97, 0.999)
        r_off = max_offset_deg * math.sqrt(rng.uniform(0.0, 1.0))  # uniform in the disc
        a_off = rng.uniform(0.0, 2 * math.pi)
        n = int(rng.integers(min_samples, max_samples + 1))
        phase = rng.uniform(0.0, 60.0)  # seconds of spin phase
        beam_rng = np.random.default_rng([seed, d])
        draws.append((fwhm, ell, ang, cut, r_off, a_off, n, phase,
                      polarised_beams(nside_beam, fwhm, ell, ang, cut, leakage, beam_rng)))
    if tail_clusters is None:
        tail_clusters = calibrated_tail_clusters([e for dr in draws[:calib_beams] for e in dr[8][:1]],
                                                 tail_fraction, log=log, device=device)
    out = []
    desc = {"detectors": n_det, "S": [], "B": [], "fwhm": [], "ellipticity": [], "cut": [],
            "offset_deg": [], "tail_clusters": int(tail_clusters), "tail_fraction": tail_fraction}
    for d, (fwhm, ell, ang, cut, r_off, a_off, n, phase, beams) in enumerate(draws):
        entries = []
        for k, (vec, bv) in enumerate(beams):
            v, b, _ = cluster_beam_pixels(vec, bv, tail_clusters, tail_fraction=tail_fraction,
                                          random_state=42 + 3 * d + k, device=device)
            entries.append((np.ascontiguousarray(v, np.float32), np.ascontiguousarray(b, np.float32)))
        th, ph, ps = inputs.ring_scan(n, fs, t0=3600.0 * d + phase)
        off = math.radians(r_off)
        th, ph, ps = focal_plane_pointing(th, ph, ps, off * math.cos(a_off), off * math.sin(a_off))
        out.append({"entries": entries, "params": R.rotation_params_for_pointing(th, ph, ps), "B": n})
        desc["S"].append([int(e[1].shape[0]) for e in entries])
        for k, v in (("B", n), ("fwhm", fwhm), ("ellipticity", ell), ("cut", cut), ("offset_deg", r_off)):
            desc[k].append(v)
        if log and (d + 1) % 16 == 0:
            log(f"batch inputs: {d + 1}/{n_det} detectors ({time.time() - t0:.1f}s)")
    desc["bps"] = float(sum(b * sum(s) for b, s in zip(desc["B"], desc["S"])))
    return out, desc


def batch_plans(dets, maps, chunk: int = 8, variant: int = 0):
    """One BatchPlan per Stokes component (I, Q, U): the component's entry of
    every detector as a one-field launch on that component's map
    (``maps``: three one-field DeviceMaps)."""
    from .batch import BatchPlan, Detector

    return [BatchPlan([Detector(d["entries"][k][0], d["entries"][k][1], d["params"]) for d in dets],
                      maps[k], "bilinear", chunk, variant) for k in range(3)]
