"""Beam transfer function and the clustering calibration sweep.

``compute_bell`` keeps the reference's signature and steps
(tod_utils.py:147-251): dB power-cut selection, normalisation, the
cos-distance cos(dec) cos(ra) of each pixel from the beam centre, the
Legendre sum B_ell = sum_i w_i P_ell(cos theta_i) (on the GPU,
``todg_beam_bell``), B_0 = 1 normalisation and the absolute value.

``calibrate_beam_clustering`` is the reference's (tail_fraction x
n_clusters) sweep (tod_calibrate.py:392-530): the pair with the largest
speedup whose relative-RMS B_ell divergence from the unclustered beam stays
below the threshold, else the minimum-divergence pair.  Pixel directions
(theta, phi) of HEALPix beam vectors enter compute_bell as
(ra, dec) = (phi, pi/2 - theta) exactly as the reference feeds them
(tod_calibrate.py:431-441).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .beam_cluster import cluster_beam_pixels
from .device import device, stream_handle
from .inputs import db_threshold_from_power

TAIL_FRACTIONS = (0.005, 0.01, 0.02, 0.03, 0.05, 0.075, 0.10, 0.15, 0.20, 0.30)
N_CLUSTERS = (10, 20, 50, 100, 200, 500, 1000, 2000)


def bell_sum(cos_theta, weights, lmax: int) -> np.ndarray:
    """sum_i w_i P_ell(x_i) for ell = 0..lmax (float64, computed on the GPU)."""
    dev = device()
    x = torch.as_tensor(np.ascontiguousarray(cos_theta, dtype=np.float64), device=dev)
    w = torch.as_tensor(np.ascontiguousarray(weights, dtype=np.float64), device=dev)
    N = int(x.numel())
    ws = torch.empty(int(_lib.lib().todg_beam_bell_workspace(N, int(lmax))), dtype=torch.float64,
                     device=dev)
    out = torch.empty(int(lmax) + 1, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().todg_beam_bell(_lib.ptr(x), _lib.ptr(w), N, int(lmax), _lib.ptr(ws),
                                         _lib.ptr(out), stream_handle()))
    return out.cpu().numpy()


def compute_bell(ra, dec, pixel_map, lmax=1000, power_cut=0.99, normalise=True, verbose=True):
    """(ell, |B_ell|) of a pixelised beam (tod_utils.py:147-251)."""
    pixel_map = np.asarray(pixel_map, dtype=np.float64)
    ra = np.asarray(ra, dtype=np.float64).ravel()
    dec = np.asarray(dec, dtype=np.float64).ravel()
    flat = pixel_map.ravel()
    if power_cut >= 1.0:
        sel = np.ones(flat.shape, dtype=bool)
        if verbose:
            print(f"  power_cut=1.0: selecting all {len(flat)} pixels")
    else:
        dB_cut = db_threshold_from_power(flat, power_cut)
        log_map = 10.0 * np.log10(np.abs(flat) + 1e-30)
        sel = log_map > dB_cut
        if verbose:
            print(f"  power_cut={power_cut}: {np.sum(sel)}/{len(flat)} pixels selected "
                  f"(dB_cut={dB_cut:.2f})")
    if not np.any(sel):
        raise ValueError("No pixels survive the power-cut selection. "
                         "Check that pixel_map is in linear (not dB) units.")
    beam_vals = flat[sel]
    norm = beam_vals.sum()
    if norm <= 0:
        raise ValueError("Sum of beam values is non-positive after selection.")
    beam_vals = beam_vals / norm
    cos_theta = np.clip(np.cos(dec[sel]) * np.cos(ra[sel]), -1.0, 1.0)
    bell = bell_sum(cos_theta, beam_vals, lmax)
    if normalise and bell[0] != 0.0:
        bell = bell / bell[0]
    return np.arange(lmax + 1, dtype=np.int64), np.abs(bell)


def _bell_from_vecs(vec, bvals, lmax):
    v = np.asarray(vec, dtype=np.float64)
    theta = np.arccos(np.clip(v[:, 2] / np.linalg.norm(v, axis=1), -1.0, 1.0))
    phi = np.mod(np.arctan2(v[:, 1], v[:, 0]), 2.0 * math.pi)
    _, bell = compute_bell(phi, math.pi / 2.0 - theta, np.asarray(bvals, dtype=np.float64),
                           lmax=lmax, power_cut=1.0, verbose=False)
    return bell


def calibrate_beam_clustering(beam_data, error_threshold=1e-3, bell_lmax=500,
                              tail_fractions=TAIL_FRACTIONS, n_clusters=N_CLUSTERS, verbose=True):
    """(tail_fraction, n_clusters) for beam entries ``{name: {'vec_orig',
    'beam_vals'}}`` (tod_calibrate.py:392-530); also returns the table."""
    ref = {bf: _bell_from_vecs(d["vec_orig"], d["beam_vals"], bell_lmax) for bf, d in beam_data.items()}
    S = {bf: len(d["beam_vals"]) for bf, d in beam_data.items()}
    n_tail = {}
    for bf, d in beam_data.items():
        bv = np.asarray(d["beam_vals"], dtype=np.float64)
        cs = np.cumsum(bv[np.argsort(bv)])
        n_tail[bf] = {tf: max(1, min(int(np.searchsorted(cs, tf, side="right")), S[bf] - 1))
                      for tf in tail_fractions}
    results = []
    for tf in tail_fractions:
        for K in n_clusters:
            k_out, divs = {}, []
            for bf, d in beam_data.items():
                if K >= n_tail[bf][tf]:
                    k_out[bf] = S[bf]
                    divs.append(0.0)
                    continue
                vc, bc, _ = cluster_beam_pixels(d["vec_orig"], d["beam_vals"], n_clusters=K,
                                                tail_fraction=tf)
                k_out[bf] = len(bc)
                b = _bell_from_vecs(vc, bc, bell_lmax)
                r = ref[bf]
                divs.append(float(np.sqrt(np.mean((b - r) ** 2)) / (np.sqrt(np.mean(r ** 2)) + 1e-30)))
            speedup = float(np.mean([S[bf] / k_out[bf] for bf in beam_data]))
            results.append((tf, K, int(np.mean(list(k_out.values()))), speedup, float(np.mean(divs))))
    passing = [r for r in results if r[4] <= error_threshold]
    best = max(passing, key=lambda r: r[3]) if passing else min(results, key=lambda r: r[4])
    if verbose:
        for tf, K, ko, sp, dv in results:
            print(f"{tf * 100:>5.1f}%  {K:>5d}  {ko:>6d}  {sp:>8.2f}x  {dv:>10.2e}  "
                  f"{'ok' if dv <= error_threshold else '--'}")
        print(f"recommendation: tail_fraction={best[0]}, n_clusters={best[1]} "
              f"(speedup {best[3]:.2f}x, B_ell div {best[4]:.2e})")
    return float(best[0]), int(best[1]), results
