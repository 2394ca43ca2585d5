"""Vectorised HEALPix RING-scheme geometry on the host (numpy) for input
preparation: pixel centres, ring layout and pixel->vector conversion.  The
same closed forms as the device table (Gorski et al. 2005, ApJ 622, 759).
"""

from __future__ import annotations

import numpy as np


def npix(nside: int) -> int:
    return 12 * nside * nside


def ring_layout(nside: int):
    """Per ring ir = 1..4n-1: (n_pix, first_pix, phi0, z) as arrays."""
    ir = np.arange(1, 4 * nside, dtype=np.int64)
    n = np.where(ir < nside, 4 * ir, np.where(ir <= 3 * nside, 4 * nside, 4 * (4 * nside - ir)))
    i2 = 4 * nside - ir
    first = np.where(
        ir < nside,
        2 * ir * (ir - 1),
        np.where(ir <= 3 * nside, 2 * nside * (nside - 1) + (ir - nside) * 4 * nside,
                 npix(nside) - 2 * i2 * (i2 + 1)),
    )
    shifted = np.where((ir < nside) | (ir > 3 * nside), 1, ((ir - nside) % 2 == 0).astype(np.int64))
    dphi = 2.0 * np.pi / n
    phi0 = shifted * dphi * 0.5
    z = np.where(
        ir < nside,
        1.0 - ir.astype(np.float64) ** 2 / (3.0 * nside * nside),
        np.where(ir <= 3 * nside, (2.0 / 3.0) * (2.0 - ir / nside),
                 -(1.0 - i2.astype(np.float64) ** 2 / (3.0 * nside * nside))),
    )
    return n, first, phi0, z


def pix2ring(nside: int, pix):
    """Ring index (1-based) and position within the ring for RING pixels."""
    pix = np.asarray(pix, dtype=np.int64)
    ncap = 2 * nside * (nside - 1)
    np_ = npix(nside)
    ir = np.empty_like(pix)
    ip = np.empty_like(pix)
    north = pix < ncap
    south = pix >= np_ - ncap
    belt = ~(north | south)
    pn = pix[north]
    irn = ((1 + np.sqrt(1 + 2 * pn.astype(np.float64))) * 0.5).astype(np.int64)
    # guard the float sqrt at ring boundaries
    irn = np.where(2 * irn * (irn - 1) > pn, irn - 1, irn)
    irn = np.where(2 * (irn + 1) * irn <= pn, irn + 1, irn)
    ir[north] = irn
    ip[north] = pn - 2 * irn * (irn - 1)
    pb = pix[belt] - ncap
    ir[belt] = pb // (4 * nside) + nside
    ip[belt] = pb % (4 * nside)
    ps = np_ - 1 - pix[south]
    irs = ((1 + np.sqrt(1 + 2 * ps.astype(np.float64))) * 0.5).astype(np.int64)
    irs = np.where(2 * irs * (irs - 1) > ps, irs - 1, irs)
    irs = np.where(2 * (irs + 1) * irs <= ps, irs + 1, irs)
    ir[south] = 4 * nside - irs
    ip[south] = pix[south] - (np_ - 2 * irs * (irs + 1))
    return ir, ip


def pix2zphi(nside: int, pix):
    n, first, phi0, z = ring_layout(nside)
    ir, ip = pix2ring(nside, pix)
    k = ir - 1
    return z[k], phi0[k] + ip * (2.0 * np.pi / n[k])


def pix2ang(nside: int, pix):
    z, phi = pix2zphi(nside, pix)
    return np.arccos(np.clip(z, -1.0, 1.0)), phi


def pix2vec(nside: int, pix):
    z, phi = pix2zphi(nside, pix)
    st = np.sqrt(np.maximum(0.0, (1.0 - z) * (1.0 + z)))
    return np.stack([st * np.cos(phi), st * np.sin(phi), z], axis=-1)


def cap_pixels(nside: int, theta_max: float):
    """All pixels of the rings whose centre colatitude is <= theta_max
    (north polar region), in RING order."""
    n, first, phi0, z = ring_layout(nside)
    zmin = np.cos(theta_max)
    last = int(np.searchsorted(-z, -zmin, side="right"))  # rings with z >= zmin
    if last == 0:
        return np.zeros(0, dtype=np.int64)
    return np.arange(0, first[last - 1] + n[last - 1], dtype=np.int64)
