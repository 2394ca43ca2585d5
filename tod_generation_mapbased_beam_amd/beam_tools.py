"""Beam transfer function B_l and the tail-clustering calibration sweep.

Two host-side front ends around the GPU Legendre sum (``todg_beam_bell``):

* ``compute_bell`` -- the B_l of a pixelised beam, with the reference's call
  signature and rules (tod_utils.py:147-251): keep the pixels above the dB
  level holding ``power_cut`` of the summed beam (``power_cut >= 1`` keeps
  all), normalise their values to unit sum, take each pixel's cosine
  distance from the beam centre as cos(dec) cos(ra), sum w P_l(x) over the
  pixels, divide by B_0 and return |B_l|.
* ``calibrate_beam_clustering`` -- the (tail fraction x cluster count) sweep
  of tod_calibrate.py:392-530: for every pair, cluster each beam's tail,
  measure the relative RMS change of its B_l against the unclustered beam,
  and recommend the pair with the largest mean pixel reduction among those
  within the error threshold (the smallest error when none is).  HEALPix
  beam vectors enter as (ra, dec) = (phi, pi/2 - theta), which is how the
  reference feeds them (tod_calibrate.py:431-441), so B_l is measured about
  the x axis as there.
"""

from __future__ import annotations

import math
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from .beam_cluster import cluster_beam_pixels
from .device import device, stream_handle
from .inputs import db_threshold_from_power

TAIL_FRACTIONS = (0.005, 0.01, 0.02, 0.03, 0.05, 0.075, 0.10, 0.15, 0.20, 0.30)
N_CLUSTERS = (10, 20, 50, 100, 200, 500, 1000, 2000)


def bell_sum(cos_theta, weights, lmax: int) -> np.ndarray:
    """sum_i w_i P_l(x_i) for l = 0..lmax (float64, computed on the GPU)."""
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


def _kept_pixels(values: np.ndarray, power_cut: float) -> np.ndarray:
    """Mask of the pixels a power cut keeps (all of them at power_cut >= 1)."""
    if power_cut >= 1.0:
        return np.ones(values.shape, dtype=bool)
    level = db_threshold_from_power(values, power_cut)
    return 10.0 * np.log10(np.abs(values) + 1e-30) > level


def compute_bell(ra, dec, pixel_map, lmax=1000, power_cut=0.99, normalise=True, verbose=True):
    """(l, |B_l|) of a pixelised beam given per-pixel offsets (ra, dec) from
    its centre (radians) and its linear-unit values; see the module notes."""
    values = np.asarray(pixel_map, dtype=np.float64).ravel()
    keep = _kept_pixels(values, power_cut)
    n_keep = int(np.count_nonzero(keep))
    if verbose:
        print(f"compute_bell: {n_keep} of {values.size} beam pixels in the {power_cut:g} power cut")
    if n_keep == 0:
        raise ValueError("compute_bell: the power cut keeps no pixel (are the beam values in dB "
                         "rather than linear units?)")
    w = values[keep]
    total = float(w.sum())
    if not total > 0.0:
        raise ValueError("compute_bell: the kept beam values do not sum to a positive total")
    x = np.cos(np.asarray(dec, dtype=np.float64).ravel()[keep]) * \
        np.cos(np.asarray(ra, dtype=np.float64).ravel()[keep])
    bell = bell_sum(np.clip(x, -1.0, 1.0), w / total, lmax)
    if normalise and bell[0] != 0.0:
        bell = bell / bell[0]
    return np.arange(lmax + 1, dtype=np.int64), np.abs(bell)


def beam_bell(vec, beam_vals, lmax: int) -> np.ndarray:
    """|B_l| of a beam given as unit vectors and values (every pixel kept)."""
    v = np.asarray(vec, dtype=np.float64)
    theta = np.arccos(np.clip(v[:, 2] / np.linalg.norm(v, axis=1), -1.0, 1.0))
    phi = np.mod(np.arctan2(v[:, 1], v[:, 0]), 2.0 * math.pi)
    return compute_bell(phi, math.pi / 2.0 - theta, beam_vals, lmax=lmax, power_cut=1.0, verbose=False)[1]


class SweepPoint(NamedTuple):
    """One row of the calibration table (averages over the beams)."""
    tail_fraction: float
    n_clusters: int
    kept_pixels: int      # beam pixels after clustering
    speedup: float        # original pixels / kept pixels
    divergence: float     # relative RMS change of B_l


def _tail_size(beam_vals, tail_fraction: float) -> int:
    """Pixels in the weakest part of the beam holding tail_fraction of it."""
    cum = np.cumsum(np.sort(np.asarray(beam_vals, dtype=np.float64)))
    return int(np.searchsorted(cum, tail_fraction, side="right"))


def calibrate_beam_clustering(beam_data, error_threshold=1e-3, bell_lmax=500,
                              tail_fractions=TAIL_FRACTIONS, n_clusters=N_CLUSTERS, verbose=True,
                              device=None):
    """Recommended (tail_fraction, n_clusters) for the beam entries
    ``{name: {"vec_orig", "beam_vals"}}`` and the sweep table (a list of
    SweepPoint); see the module notes.  ``device``: where the tail k-means
    runs (None: the CPU)."""
    beams = list(beam_data.values())
    ref = [beam_bell(b["vec_orig"], b["beam_vals"], bell_lmax) for b in beams]
    sizes = [len(b["beam_vals"]) for b in beams]
    table = []
    for tf in tail_fractions:
        for K in n_clusters:
            kept, div = [], []
            for b, r, S in zip(beams, ref, sizes):
                if K >= _tail_size(b["beam_vals"], tf):  # nothing to merge
                    kept.append(S)
                    div.append(0.0)
                    continue
                vc, bc, _ = cluster_beam_pixels(b["vec_orig"], b["beam_vals"], n_clusters=K,
                                                tail_fraction=tf, device=device)
                kept.append(len(bc))
                diff = beam_bell(vc, bc, bell_lmax) - r
                div.append(float(np.sqrt(np.mean(diff ** 2)) / (np.sqrt(np.mean(r ** 2)) + 1e-30)))
            table.append(SweepPoint(tf, K, int(np.mean(kept)),
                                    float(np.mean([S / k for S, k in zip(sizes, kept)])),
                                    float(np.mean(div))))
    within = [p for p in table if p.divergence <= error_threshold]
    pick = max(within, key=lambda p: p.speedup) if within else min(table, key=lambda p: p.divergence)
    if verbose:
        print("tail fraction  clusters  pixels  speed-up  B_l change")
        for p in table:
            mark = "*" if p is pick else ("" if p.divergence <= error_threshold else "  (over)")
            print(f"{p.tail_fraction:13.3f}  {p.n_clusters:8d}  {p.kept_pixels:6d}  {p.speedup:7.2f}x  "
                  f"{p.divergence:10.2e}{mark}")
    return float(pick.tail_fraction), int(pick.n_clusters), table
