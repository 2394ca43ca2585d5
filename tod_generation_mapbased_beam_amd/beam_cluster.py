"""Beam-pixel reduction by weighted spherical k-means of the low-power tail.

Counterpart of the reference's ``cluster_beam_pixels`` (beam_cluster.py:215-397)
with the same contract, so a clustered beam drops into ``beam_tod_batch``:

* full mode (``tail_fraction=None``): all S pixels -> K centroids;
* hybrid mode: pixels are split by cumulative power from the weakest one;
  the weakest pixels holding ``tail_fraction`` of the power are clustered,
  the rest (the main lobe) is kept exactly;
* output layout ``[main pixels (ascending weight)..., tail centroids...]``
  with beam weights renormalised to unit sum, and ``labels`` mapping every
  input pixel to its output row (beam_cluster.py:378-389).  The layout and
  the labels are kept identical to the reference because downstream code
  indexes by them.

The clustering itself is weighted spherical k-means: weighted k-means++
seeding with probability ~ w (1 - cos d)^2 to the nearest chosen centre,
then assignment by maximum cosine, weighted mean direction renormalised to
the sphere, empty clusters reseeded with the worst-fitted pixel, stop when
no centre moves more than ``tol`` radians.  It runs on torch tensors, so the
S x K similarity products go to the GPU when one is given.
"""

from __future__ import annotations

import math

import numpy as np
import torch


def _seed_centres(v: torch.Tensor, w: torch.Tensor, K: int, gen: np.random.Generator):
    S = v.shape[0]
    base = (w / w.sum()).cpu().numpy().astype(np.float64)
    base = base / base.sum()
    idx = [int(gen.choice(S, p=base))]
    best = v @ v[idx[0]]
    for _ in range(1, K):
        d2 = torch.clamp(1.0 - best, min=0.0) ** 2
        p = (w * d2).cpu().numpy().astype(np.float64)
        tot = p.sum()
        p = base if tot <= 0.0 else p / tot
        j = int(gen.choice(S, p=p / p.sum()))
        idx.append(j)
        best = torch.maximum(best, v @ v[j])
    return v[idx].clone()


def kmeans_sphere(vec, weights, K: int, max_iter: int = 150, tol: float = 1e-5,
                  seed: int = 42, device=None):
    """Weighted spherical k-means -> (centres (K,3) float64, labels (S,) int32)."""
    dev = torch.device(device) if device is not None else torch.device("cpu")
    v = torch.as_tensor(np.asarray(vec, dtype=np.float64), device=dev)
    w = torch.as_tensor(np.asarray(weights, dtype=np.float64), device=dev)
    gen = np.random.default_rng(seed)
    C = _seed_centres(v, w, K, gen)
    cos_tol = math.cos(tol)
    for _ in range(max_iter):
        sim = v @ C.T
        lab = torch.argmax(sim, dim=1)
        mass = torch.zeros(K, dtype=v.dtype, device=dev).index_add_(0, lab, w)
        newc = torch.zeros_like(C).index_add_(0, lab, v * w[:, None])
        empty = torch.nonzero(mass == 0).flatten()
        if empty.numel():
            fit = sim.gather(1, lab[:, None]).flatten()
            for k in empty.tolist():
                j = int(torch.argmin(fit))
                newc[k] = v[j]
                fit[j] = 2.0  # do not reuse the same pixel
        nrm = torch.linalg.norm(newc, dim=1, keepdim=True)
        newc = newc / torch.where(nrm > 0, nrm, torch.ones_like(nrm))
        moved = torch.clamp((C * newc).sum(dim=1), -1.0, 1.0)
        C = newc
        if float(moved.min()) > cos_tol:
            break
    lab = torch.argmax(v @ C.T, dim=1)
    return C.cpu().numpy(), lab.cpu().numpy().astype(np.int32)


def cluster_beam_pixels(vec_orig, beam_vals, n_clusters: int, tail_fraction=None,
                        max_iter: int = 150, tol: float = 1e-5, random_state: int = 42,
                        device=None):
    """Same signature and return contract as beam_cluster.py:215-397:
    ``(vec_out (K_out,3) float32, bvals_out (K_out,) float32, labels (S,) int32)``."""
    vec_orig = np.asarray(vec_orig)
    beam_vals = np.asarray(beam_vals)
    S = beam_vals.shape[0]
    if tail_fraction is None:
        K = min(int(n_clusters), S)
        if K >= S:
            return vec_orig.copy(), beam_vals.copy(), np.arange(S, dtype=np.int32)
        C, lab = kmeans_sphere(vec_orig, beam_vals, K, max_iter, tol, random_state, device)
        cw = np.bincount(lab, weights=beam_vals.astype(np.float64), minlength=K)
        cw /= cw.sum()
        return C.astype(np.float32), cw.astype(np.float32), lab
    if not 0.0 < tail_fraction < 1.0:
        raise ValueError(f"tail_fraction must be in (0, 1), got {tail_fraction!r}")

    order = np.argsort(beam_vals)  # weakest first
    cum = np.cumsum(beam_vals[order])
    n_tail = int(np.searchsorted(cum, tail_fraction, side="right"))
    n_tail = max(1, min(n_tail, S - 1))
    tail = order[:n_tail]
    main = order[n_tail:]
    n_main = main.shape[0]
    K = min(int(n_clusters), n_tail)
    if K >= n_tail:
        v_t = vec_orig[tail]
        b_t = beam_vals[tail]
        lab_t = np.arange(n_tail, dtype=np.int32)
    else:
        C, lab_t = kmeans_sphere(vec_orig[tail], beam_vals[tail], K, max_iter, tol,
                                 random_state, device)
        v_t = C.astype(np.float32)
        b_t = np.bincount(lab_t, weights=beam_vals[tail].astype(np.float64),
                          minlength=K).astype(np.float32)
    v_out = np.concatenate([vec_orig[main], v_t], axis=0).astype(np.float32)
    b_out = np.concatenate([beam_vals[main], b_t]).astype(np.float32)
    b_out = (b_out / b_out.sum()).astype(np.float32)
    labels = np.empty(S, dtype=np.int32)
    labels[main] = np.arange(n_main, dtype=np.int32)
    labels[tail] = n_main + lab_t
    return np.ascontiguousarray(v_out), b_out, labels
