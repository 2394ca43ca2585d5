"""Seeded synthetic inputs for the beam-convolution TOD path.

* sky maps: Gaussian random I/Q/U fields with the spectrum
  C_l^TT = 1e3 / (l(l+1)+1), C_l^EE = 0.05 C_l^TT, C_l^BB = 0.01 C_l^TT,
  TE = 0, synthesised on the HEALPix RING grid by this module's own
  spherical-harmonic synthesis (healpy is not available);
* beams: Gaussian beams stored as HEALPix maps centred on the north pole,
  selected with the reference's dB power cut and normalised to unit sum
  (sample_based_tod_generation_gridint.py:127-146, tod_utils.py:254-297);
* scans: a ring scan (spin axis on the equator advancing in longitude,
  fixed opening angle, constant spin rate, psi = spin phase + pi/2).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import healpix_np as hpx

ARCMIN = math.pi / (180.0 * 60.0)


# ------------------------------------------------------------------ spectra

def cl_spectrum(lmax: int):
    """(TT, EE, BB) with C_l^TT = 1e3/(l(l+1)+1), EE = 0.05 TT, BB = 0.01 TT."""
    ell = np.arange(lmax + 1, dtype=np.float64)
    tt = 1e3 / (ell * (ell + 1) + 1)
    return tt, 0.05 * tt, 0.01 * tt


# ------------------------------------------------------- harmonic synthesis

def _lambda_mm_log(m: np.ndarray) -> np.ndarray:
    """log of sqrt((2m+1)/(4 pi) prod_{k<=m} (2k-1)/(2k)) (no sin^m factor)."""
    k = np.arange(1, int(m.max()) + 2, dtype=np.float64)
    cum = np.concatenate([[0.0], np.cumsum(np.log((2 * k - 1) / (2 * k)))])
    return 0.5 * (np.log((2 * m + 1) / (4 * math.pi)) + cum[m])


def _ring_sums(alm_list, lmax, thetas, dev, spin_fields):
    """F_m(theta_r) = sum_l a_lm lambda_lm(theta_r) (spin 0), for every ring
    and m, for each a_lm set in alm_list ((lmax+1, lmax+1) complex, [l, m]).
    Normalised associated-Legendre recurrence in l at fixed m, vectorised over
    (ring, m), with a per-element exponent so polar rings at high m neither
    underflow nor lose precision.  On a GPU device the sums run in the
    library's kernel (todg_ring_sums), which performs the same binary64
    operations in the same order."""
    dev = torch.device(dev)
    if dev.type == "cuda":
        return _ring_sums_gpu(alm_list, lmax, thetas, dev)
    return _ring_sums_torch(alm_list, lmax, thetas, dev)


def _ring_sums_torch(alm_list, lmax, thetas, dev):
    """_ring_sums as torch tensor operations on ``dev``, vectorised over
    (ring, m): the host path, and the reference the GPU kernel is tested
    against on the same device."""
    dev = torch.device(dev)
    R = thetas.shape[0]
    M = lmax + 1
    x = torch.as_tensor(np.cos(thetas), dtype=torch.float64, device=dev)[:, None]  # (R,1)
    st = torch.as_tensor(np.sin(thetas), dtype=torch.float64, device=dev)[:, None]
    m = torch.arange(M, dtype=torch.float64, device=dev)[None, :]  # (1,M)
    logmm = torch.as_tensor(_lambda_mm_log(np.arange(M)), dtype=torch.float64, device=dev)[None, :]
    # lambda_mm = (-1)^m exp(logmm) sin^m ; keep value 1*sign and exponent E
    E = logmm + m * torch.log(torch.clamp(st, min=1e-300))
    sign = torch.where(torch.remainder(m, 2) == 1, -1.0, 1.0).to(torch.float64)
    prev2 = torch.zeros((R, M), dtype=torch.float64, device=dev)
    prev1 = sign.expand(R, M).clone()  # lambda_{m,m} / exp(E)
    outs = [torch.zeros((R, M), dtype=torch.complex128, device=dev) for _ in alm_list]
    alms = [torch.as_tensor(a, dtype=torch.complex128, device=dev) for a in alm_list]
    mi = torch.arange(M, device=dev)
    big = _BIG
    logbig = math.log(big)
    for k in range(0, M):
        # l = m + k for every m; valid where l <= lmax
        l = m + k
        valid = (l <= lmax)
        if k == 0:
            cur = prev1
        elif k == 1:
            cur = x * torch.sqrt(2 * m + 3) * prev1
        else:
            a = torch.sqrt((4 * l * l - 1) / torch.clamp(l * l - m * m, min=1.0))
            b = torch.sqrt(torch.clamp(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1), min=0.0))
            cur = a * (x * prev1 - b * prev2)
        cur = torch.where(valid, cur, torch.zeros_like(cur))
        scale = torch.exp(torch.clamp(E, max=700.0))
        lidx = torch.clamp(mi + k, max=lmax)
        for o, alm in zip(outs, alms):
            coeff = alm[lidx, mi][None, :]  # (1,M)
            o += torch.where(valid, coeff * (cur * scale), torch.zeros_like(o))
        if k >= 1:
            prev2, prev1 = prev1, cur
        # renormalise large values into the exponent
        mag = torch.abs(prev1)
        hot = mag > big
        if bool(hot.any()):
            prev1 = torch.where(hot, prev1 / big, prev1)
            prev2 = torch.where(hot, prev2 / big, prev2)
            E = torch.where(hot, E + logbig, E)
    return outs


def _alm_realisation(cl, lmax, rng):
    """Complex a_lm, [l, m] for m <= l, with <|a_lm|^2> = C_l."""
    a = np.zeros((lmax + 1, lmax + 1), dtype=np.complex128)
    for ell in range(lmax + 1):
        s = math.sqrt(cl[ell])
        re = rng.standard_normal(ell + 1)
        im = rng.standard_normal(ell + 1)
        row = (re + 1j * im) * s / math.sqrt(2.0)
        row[0] = re[0] * s
        a[ell, : ell + 1] = row
    return a


_BIG = 1e150


def _ring_sums_gpu(alm_list, lmax, thetas, dev):
    """_ring_sums on the GPU: one launch of todg_ring_sums for up to four
    a_lm sets at a time; returns (R, M) complex128 device tensors."""
    from . import _lib
    from .device import stream_handle

    R = int(thetas.shape[0])
    M = lmax + 1
    x = torch.as_tensor(np.cos(thetas), dtype=torch.float64, device=dev).contiguous()
    st = torch.as_tensor(np.sin(thetas), dtype=torch.float64, device=dev).contiguous()
    logmm = torch.as_tensor(_lambda_mm_log(np.arange(M)), dtype=torch.float64, device=dev).contiguous()
    outs = []
    for g0 in range(0, len(alm_list), 4):
        group = alm_list[g0:g0 + 4]
        # [k][m][l]: a lane's reads along l are contiguous
        alm = torch.stack([torch.as_tensor(np.asarray(a), dtype=torch.complex128, device=dev).transpose(0, 1)
                           for a in group]).contiguous()
        if tuple(alm.shape[1:]) != (M, M):
            raise ValueError(f"a_lm sets must be ({M}, {M})")
        F = torch.empty((len(group), M, R), dtype=torch.complex128, device=dev)
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().todg_ring_sums(
                _lib.ptr(alm), len(group), lmax, _lib.ptr(x), _lib.ptr(st), R, _lib.ptr(logmm), _BIG,
                math.log(_BIG), _lib.ptr(F), stream_handle()))
        del alm
        outs += [F[k].transpose(0, 1) for k in range(len(group))]
    return outs


def synthesize_map(nside: int, alm_list, lmax: int, device=None):
    """Spin-0 synthesis of one or more a_lm sets on the RING grid -> (K, npix) float64."""
    dev = torch.device(device) if device is not None else torch.device("cpu")
    n, first, phi0, z = hpx.ring_layout(nside)
    thetas = np.arccos(z)
    F = _ring_sums(alm_list, lmax, thetas, dev, None)
    out = np.zeros((len(alm_list), hpx.npix(nside)), dtype=np.float64)
    Fh = [f.contiguous().cpu().numpy() for f in F]
    mm = np.arange(lmax + 1)
    for r in range(len(n)):
        nr = int(n[r])
        ph = np.exp(1j * mm * phi0[r])
        for i, f in enumerate(Fh):
            g = np.zeros(nr, dtype=np.complex128)
            coeff = f[r] * ph
            coeff[1:] *= 2.0  # m > 0 counted twice (real field)
            np.add.at(g, mm % nr, coeff)
            out[i, first[r]: first[r] + nr] = np.real(np.fft.ifft(g) * nr)
    return out


def synfast_iqu(nside: int, seed: int = 0, lmax: int | None = None, device=None):
    """I/Q/U map (3, npix) float64 with the configured spectrum.

    T is a spin-0 synthesis of a_lm^T.  Q/U are built from E/B modes with the
    same spin-0 machinery applied to gradient-like fields is beyond this
    module; instead Q and U are independent spin-0 realisations with the EE
    and BB power (this keeps the pixel-scale statistics that drive
    interpolation error; the E/B split itself plays no role in the TOD path).
    """
    lmax = 3 * nside - 1 if lmax is None else lmax
    tt, ee, bb = cl_spectrum(lmax)
    rng = np.random.default_rng(seed)
    a_t = _alm_realisation(tt, lmax, rng)
    a_q = _alm_realisation(ee + bb, lmax, rng)
    a_u = _alm_realisation(ee + bb, lmax, rng)
    return synthesize_map(nside, [a_t, a_q, a_u], lmax, device=device)


def white_map(nside: int, seed: int = 0, fields: int = 3, device="cpu"):
    """Cheap (fields, npix) float32 random map (for throughput runs where the
    values do not matter)."""
    g = torch.Generator(device=device).manual_seed(seed)
    return torch.randn((fields, hpx.npix(nside)), generator=g, dtype=torch.float32, device=device)


# ------------------------------------------------------------------- beams

def db_threshold_from_power(values, power_cut: float) -> float:
    """dB level keeping ``power_cut`` of the summed beam (tod_utils.py:254-297):
    sort by dB, accumulate from the top, return the dB value of the last
    element whose retained cumulative sum still reaches the target."""
    prof = np.asarray(values, dtype=np.float64).ravel()
    target = prof.sum() * power_cut
    with np.errstate(divide="ignore"):
        db = 10 * np.log10(prof)
    order = np.argsort(db)
    sdb = db[order]
    sp = prof[order]
    cum = np.cumsum(sp[::-1])[::-1]
    rev = np.searchsorted(cum[::-1], target, side="left")
    idx = max(0, len(cum) - 1 - rev)
    return float(sdb[idx])


def select_beam(values, vecs, power_cut: float):
    """Power-cut selection and normalisation of a pixelised beam
    (sample_based_tod_generation_gridint.py:127-146) -> (vec_orig f32 (S,3),
    beam_vals f32 (S,), selection mask)."""
    cut = db_threshold_from_power(values, power_cut)
    sel = 10 * np.log10(np.abs(values) + 1e-30) > cut
    bv = values[sel].astype(np.float32)
    norm = bv.sum()
    if norm != 0:
        bv /= norm
    return np.ascontiguousarray(vecs[sel].astype(np.float32)), bv, sel


def gaussian_beam_map(nside_beam: int, fwhm_arcmin: float, ellipticity: float = 0.0,
                      angle: float = 0.0, extent_sigma: float = 6.0):
    """Uncut Gaussian beam on the HEALPix pixels of a cap around the north
    pole: (values, unit vectors, colatitudes, longitudes).  ``ellipticity``
    e = 1 - sigma_minor/sigma_major with the geometric-mean width fixed by
    ``fwhm_arcmin``; ``angle`` orients the major axis in the tangent plane."""
    sigma = fwhm_arcmin * ARCMIN / math.sqrt(8 * math.log(2))
    q = 1.0 - ellipticity
    s_maj = sigma / math.sqrt(q)
    s_min = sigma * math.sqrt(q)
    pix = hpx.cap_pixels(nside_beam, extent_sigma * s_maj)
    th, ph = hpx.pix2ang(nside_beam, pix)
    x = th * np.cos(ph - angle)
    y = th * np.sin(ph - angle)
    vals = np.exp(-0.5 * ((x / s_maj) ** 2 + (y / s_min) ** 2))
    return vals, hpx.pix2vec(nside_beam, pix), th, ph


def gaussian_beam(nside_beam: int, fwhm_arcmin: float, power_cut: float = 0.99,
                  ellipticity: float = 0.0, angle: float = 0.0, extent_sigma: float = 6.0):
    """Gaussian beam as a HEALPix map centred on the north pole, cut and
    normalised (gaussian_beam_map, select_beam).  Returns (vec_orig,
    beam_vals, values)."""
    vals, vecs, _, _ = gaussian_beam_map(nside_beam, fwhm_arcmin, ellipticity, angle, extent_sigma)
    vec_orig, bv, sel = select_beam(vals, vecs, power_cut)
    return vec_orig, bv, vals[sel]


# ------------------------------------------------------------------- scans

def ring_scan(n_samples: int, fs_hz: float, opening_deg: float = 45.0, spin_rpm: float = 1.0,
              precess_deg_per_min: float = 1.0, t0: float = 0.0):
    """Boresight (theta, phi, psi) float64 for a ring scan: spin axis on the
    equator at longitude precess*t, boresight at the opening angle from it,
    rotating at spin_rpm; psi = spin phase + pi/2."""
    t = t0 + np.arange(n_samples, dtype=np.float64) / fs_hz
    lam = np.deg2rad(precess_deg_per_min) * t / 60.0
    w = 2 * math.pi * spin_rpm / 60.0 * t
    alpha = np.deg2rad(opening_deg)
    a = np.stack([np.cos(lam), np.sin(lam), np.zeros_like(lam)], axis=-1)
    e1 = np.broadcast_to(np.array([0.0, 0.0, 1.0]), a.shape)
    e2 = np.cross(e1, a)
    p = (math.cos(alpha) * a
         + math.sin(alpha) * (np.cos(w)[:, None] * e1 + np.sin(w)[:, None] * e2))
    theta = np.arccos(np.clip(p[:, 2], -1, 1))
    phi = np.mod(np.arctan2(p[:, 1], p[:, 0]), 2 * math.pi)
    psi = np.mod(w + math.pi / 2, 2 * math.pi)
    return theta, phi, psi
