"""CPU oracle of the beam transfer function -- TEST INFRASTRUCTURE.

Only tests/ may import this module.  It restates compute_bell
(tod_utils.py:147-251) in float64 numpy: dB power-cut selection,
normalisation, cos theta = cos(dec) cos(ra), the Legendre three-term
recurrence summed over pixels, B_0 = 1 normalisation, absolute value.
Pinned in tests/test_bell.py by closed forms: a single on-centre pixel has
B_ell = 1, a two-point beam equals scipy's Legendre values, and a narrow
Gaussian beam follows exp(-l(l+1) sigma^2 / 2).
"""

import numpy as np


def db_threshold(vals, power_cut):
    """tod_utils.py:254-297: ascending dB order; the suffix sums of the linear
    values give the power kept above each level; the threshold is the dB value
    of the last position whose suffix sum still reaches power_cut of the total."""
    p = np.asarray(vals, dtype=np.float64).ravel()
    goal = p.sum() * power_cut
    with np.errstate(divide="ignore"):
        d = 10.0 * np.log10(p)
    o = np.argsort(d)
    suffix = np.cumsum(p[o][::-1])[::-1]
    k = len(suffix) - 1 - np.searchsorted(suffix[::-1], goal, side="left")
    return d[o][max(0, k)]


def compute_bell(ra, dec, pixel_map, lmax=1000, power_cut=0.99, normalise=True):
    flat = np.asarray(pixel_map, dtype=np.float64).ravel()
    ra = np.asarray(ra, dtype=np.float64).ravel()
    dec = np.asarray(dec, dtype=np.float64).ravel()
    if power_cut >= 1.0:
        sel = np.ones(flat.shape, dtype=bool)
    else:
        sel = 10.0 * np.log10(np.abs(flat) + 1e-30) > db_threshold(flat, power_cut)
    w = flat[sel] / flat[sel].sum()
    x = np.clip(np.cos(dec[sel]) * np.cos(ra[sel]), -1.0, 1.0)
    bell = np.empty(lmax + 1)
    p2 = np.ones_like(x)
    p1 = x.copy()
    bell[0] = np.dot(w, p2)
    if lmax >= 1:
        bell[1] = np.dot(w, p1)
    for l in range(1, lmax):
        p = ((2.0 * l + 1.0) * x * p1 - l * p2) / (l + 1.0)
        bell[l + 1] = np.dot(w, p)
        p2, p1 = p1, p
    if normalise and bell[0] != 0.0:
        bell = bell / bell[0]
    return np.arange(lmax + 1), np.abs(bell)
