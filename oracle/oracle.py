"""numpy/ctypes front end of the CPU oracle (``oracle/tod_oracle.c``).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- always as the checker or the
reported CPU baseline, never as part of the product path.

Every function cites the reference code it restates (the C source holds the
line-by-line restatement).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None
_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_f64 = ctypes.c_double

_SIGS = {
    "or_ring_above": (_i64, [_i64, _f64]),
    "or_ring_info": (None, [_i64, _i64, _i64, _p, _p, _p, _p]),
    "or_ring_z": (_f64, [_i64, _i64]),
    "or_ring_interp_with_angles": (None, [_i64, _f64, _f64, _i64, _p, _p, _p, _p]),
    "or_get_interp_weights": (None, [_i64, _i64, _p, _p, _p, _p]),
    "or_pix2zphi": (None, [_i64, _i64, _p, _p]),
    "or_nearest_pix": (None, [_i64, _f64, _f64, _f64, _i64, _p, _p, _p]),
    "or_gather_ring_stencil": (_i64, [_i64, _f64, _f64, _p, _p, _p]),
    "or_rodrigues_apply_one": (None, [_p, _p]),
    "or_rodrigues_batch": (None, [_p, _p, _p, _p, _p, _p, _p, _i64, _i64, _p]),
    "or_spin2": (None, [_f64, _f64, _f64, _f64, _f64, _f64, _p, _p]),
    "or_gather_accum_fused": (
        None, [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _i64, _p, _i64, _i64, _p, _i64, _i64, ctypes.c_int]),
    "or_gather_accum_nearest": (
        None, [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _i64, _p, _i64, _i64, _p, _i64, _i64, ctypes.c_int]),
    "or_gather_accum": (None, [_p, _p, _p, _p, _i64, _i64, _i64, _i64, _p]),
    "or_fused_indices": (None, [_p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _p, _p]),
    "or_nearest_indices": (None, [_p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _p]),
    "or_rotated_dirs": (None, [_p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _p]),
    "or_max_threads": (ctypes.c_int, []),
    "or_bicubic_weights": (_i64, [_i64, _f64, _f64, _p, _p]),
    "or_gather_accum_bicubic": (
        None, [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _i64, _p, _i64, _i64, _p, _i64, _i64, ctypes.c_int]),
    "or_interp_values": (None, [_i64, _p, _i64, _i64, _p, _p, ctypes.c_int, _p]),
    "or_pixel_ring": (_i64, [_i64, _i64]),
    "or_fcorr_features": (None, [_i64, _p, _i64, _p]),
    "or_fcorr_channels": (None, [_i64, _p, _p, _p, _p]),
    "or_gather_accum_fcorr": (
        None, [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _p, _i64, _p, _i64, _i64, _p, _i64, _i64, ctypes.c_int]),
}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(_p)


# ----------------------------------------------------------- ring geometry

def ring_above(nside, z):
    """numba_healpy.py:53-70"""
    return int(lib().or_ring_above(nside, float(z)))


def ring_info(nside, ir):
    """numba_healpy.py:73-101 -> (n_pix, first_pix, phi0, dphi)"""
    n = ctypes.c_int64()
    fp = ctypes.c_int64()
    p0 = ctypes.c_double()
    dp = ctypes.c_double()
    lib().or_ring_info(nside, ir, 12 * nside * nside, ctypes.byref(n), ctypes.byref(fp),
                       ctypes.byref(p0), ctypes.byref(dp))
    return n.value, fp.value, p0.value, dp.value


def ring_z(nside, ir):
    """numba_healpy.py:104-114"""
    return float(lib().or_ring_z(nside, ir))


def ring_interp(nside, z, phi_w):
    """numba_healpy.py:358-499 -> (pix[4], w[4], z_n[4], phi_n[4])"""
    p = np.zeros(4, np.int64)
    w = np.zeros(4)
    zn = np.zeros(4)
    pn = np.zeros(4)
    lib().or_ring_interp_with_angles(nside, float(z), float(phi_w), 12 * nside * nside,
                                     _ptr(p), _ptr(w), _ptr(zn), _ptr(pn))
    return p, w, zn, pn


def get_interp_weights(nside, theta, phi):
    """numba_healpy.py:120-253 -> (pix (4,N) int64, wgt (4,N) float64)"""
    th = _c(np.ravel(theta), np.float64)
    ph = _c(np.ravel(phi), np.float64)
    N = th.shape[0]
    pix = np.empty((4, N), np.int64)
    wgt = np.empty((4, N), np.float64)
    lib().or_get_interp_weights(nside, N, _ptr(th), _ptr(ph), _ptr(pix), _ptr(wgt))
    return pix, wgt


def pix2zphi(nside, pix):
    """numba_healpy.py:553-580"""
    z = ctypes.c_double()
    ph = ctypes.c_double()
    lib().or_pix2zphi(nside, int(pix), ctypes.byref(z), ctypes.byref(ph))
    return z.value, ph.value


def nearest_pix(nside, z, sin_th, phi_w):
    """tod_nearest.py:112-141 -> (pix, z_c, phi_c)"""
    bp = ctypes.c_int64()
    bz = ctypes.c_double()
    bph = ctypes.c_double()
    lib().or_nearest_pix(nside, float(z), float(sin_th), float(phi_w), 12 * nside * nside,
                         ctypes.byref(bp), ctypes.byref(bz), ctypes.byref(bph))
    return bp.value, bz.value, bph.value


def gather_ring_stencil(nside, vz, ph):
    """numba_healpy.py:892-977 -> (pix[M], z[M], phi[M])"""
    out = np.zeros(64, np.int64)
    zb = np.zeros(64)
    pb = np.zeros(64)
    m = lib().or_gather_ring_stencil(nside, float(vz), float(ph), _ptr(out), _ptr(zb), _ptr(pb))
    return out[:m].copy(), zb[:m].copy(), pb[:m].copy()


# ------------------------------------------------------------- rotations

def rodrigues_apply_one(*args13):
    """tod_rotations.py:83-134 (13 scalars -> (vx, vy, vz)), binary64"""
    a = _c(args13, np.float64)
    o = np.zeros(3)
    lib().or_rodrigues_apply_one(_ptr(a), _ptr(o))
    return tuple(o)


def rodrigues_batch(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p):
    """tod_rotations.py:33-80 -> (B, S, 3) float32 (Numba's mixed typing)"""
    v = _c(vec_orig, np.float32)
    ops = [_c(a, np.float32) for a in (axes, cos_a, sin_a, ax_pts, cos_p, sin_p)]
    B = ops[1].shape[0]
    S = v.shape[0]
    out = np.empty((B, S, 3), np.float32)
    lib().or_rodrigues_batch(_ptr(v), *[_ptr(o) for o in ops], B, S, _ptr(out))
    return out


def rotated_dirs(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, b0=0, nb=None):
    """Binary64 rotated directions of the fused kernels, (nb, S, 3)."""
    v = _c(vec_orig, np.float32)
    ops = [_c(a, np.float32) for a in (axes, cos_a, sin_a, ax_pts, cos_p, sin_p)]
    S = v.shape[0]
    if nb is None:
        nb = ops[1].shape[0] - b0
    out = np.empty((nb, S, 3))
    lib().or_rotated_dirs(_ptr(v), *[_ptr(o) for o in ops], S, b0, nb, _ptr(out))
    return out


# ---------------------------------------------------------------- spin-2

def spin2(z_pix, sth_pix, phi_pix, z_pts, sth_pts, phi_pts):
    """tod_bilinear.py:84-145 -> (cos 2d, sin 2d)"""
    c = ctypes.c_double()
    s = ctypes.c_double()
    lib().or_spin2(float(z_pix), float(sth_pix), float(phi_pix), float(z_pts), float(sth_pts),
                   float(phi_pts), ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


# ------------------------------------------------------------ TOD kernels

def _kernel(fn, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, mp_stacked,
            beam_vals, B, S, tod, c_q=-1, c_u=-1, nthreads=0):
    v = _c(vec_orig, np.float32)
    ops = [_c(a, np.float32) for a in (axes, cos_a, sin_a, ax_pts, cos_p, sin_p)]
    m = _c(mp_stacked, np.float32)
    if m.ndim == 1:
        m = m[None, :]
    bv = _c(beam_vals, np.float32)
    assert tod.dtype == np.float64 and tod.flags.c_contiguous and tod.shape == (m.shape[0], B)
    fn(_ptr(v), *[_ptr(o) for o in ops], int(nside), _ptr(m), m.shape[0], _ptr(bv), int(B),
       int(S), _ptr(tod), int(c_q), int(c_u), int(nthreads))
    return tod


def gather_accum_fused(*args, **kw):
    """tod_bilinear.py:205-477 (accumulates into tod, (C, B) float64)"""
    return _kernel(lib().or_gather_accum_fused, *args, **kw)


def gather_accum_nearest(*args, **kw):
    """tod_nearest.py:21-158 (accumulates into tod, (C, B) float64)"""
    return _kernel(lib().or_gather_accum_nearest, *args, **kw)


def gather_accum(pixels, weights, beam_vals, mp_stacked, B, Sc, tod):
    """tod_bilinear.py:41-78 (split pipeline)"""
    px = _c(pixels, np.int64)
    w = _c(weights, np.float64)
    bv = _c(beam_vals, np.float32)
    m = _c(mp_stacked, np.float32)
    lib().or_gather_accum(_ptr(px), _ptr(w), _ptr(bv), _ptr(m), m.shape[0], m.shape[1], int(B),
                          int(Sc), _ptr(tod))
    return tod


def fused_indices(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, b0, nb):
    """(nb, S, 4) pixels and weights of the fused bilinear kernel."""
    v = _c(vec_orig, np.float32)
    ops = [_c(a, np.float32) for a in (axes, cos_a, sin_a, ax_pts, cos_p, sin_p)]
    S = v.shape[0]
    pix = np.empty((nb, S, 4), np.int64)
    wgt = np.empty((nb, S, 4), np.float64)
    lib().or_fused_indices(_ptr(v), *[_ptr(o) for o in ops], int(nside), S, int(b0), int(nb),
                           _ptr(pix), _ptr(wgt))
    return pix, wgt


def nearest_indices(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, b0, nb):
    """(nb, S) nearest pixels of the nearest kernel."""
    v = _c(vec_orig, np.float32)
    ops = [_c(a, np.float32) for a in (axes, cos_a, sin_a, ax_pts, cos_p, sin_p)]
    S = v.shape[0]
    pix = np.empty((nb, S), np.int64)
    lib().or_nearest_indices(_ptr(v), *[_ptr(o) for o in ops], int(nside), S, int(b0), int(nb),
                             _ptr(pix))
    return pix


def bicubic_weights(nside, z, phi):
    """Reconstructed bicubic judge: (stencil pixels, weights), README.md:120 +
    numba_healpy.py:892-977 (parity unpinned: the reference module is absent)."""
    p = np.zeros(64, np.int64)
    w = np.zeros(64)
    m = lib().or_bicubic_weights(nside, float(z), float(phi), _ptr(p), _ptr(w))
    return p[:m].copy(), w[:m].copy()


def gather_accum_bicubic(*args, **kw):
    """Beam TOD with the bicubic judge (accumulates into tod)."""
    return _kernel(lib().or_gather_accum_bicubic, *args, **kw)


def interp_values(nside, mp, z, phi, mode):
    """Interpolate one map field at directions (z, phi): mode 'bilinear' or 'bicubic'."""
    m = _c(mp, np.float32)
    zz = _c(z, np.float64)
    pp = _c(phi, np.float64)
    out = np.empty(zz.shape[0])
    lib().or_interp_values(nside, _ptr(m), m.shape[0], zz.shape[0], _ptr(zz), _ptr(pp),
                           0 if mode == "bilinear" else 1, _ptr(out))
    return out


def max_threads():
    return int(lib().or_max_threads())


# ------------------------------------------- factorised corrector (own)

def pixel_ring(nside, p):
    """RING ring index (1 .. 4 nside - 1) of pixel p."""
    return int(lib().or_pixel_ring(int(nside), int(p)))


def fcorr_features(nside, field, p):
    """The factorised corrector's 16 stencil features of pixel p (the
    library's own definition; no reference counterpart)."""
    m = _c(field, np.float32)
    F = np.zeros(16)
    lib().or_fcorr_features(int(nside), _ptr(m), int(p), _ptr(F))
    return F


def fcorr_channels(nside, field, coef):
    """(A0, A1) float64 maps of one field for the 32 coefficients."""
    m = _c(field, np.float32)
    c = _c(coef, np.float64)
    o0 = np.zeros(m.shape[0])
    o1 = np.zeros(m.shape[0])
    lib().or_fcorr_channels(int(nside), _ptr(m), _ptr(c), _ptr(o0), _ptr(o1))
    return o0, o1


def gather_accum_fcorr(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, A0, A1, beam_vals,
                       B, S, tod, c_q=-1, c_u=-1, nthreads=0):
    """Factorised-corrector TOD (accumulates into tod (C, B) float64)."""
    a0 = _c(A0, np.float32)
    a1 = _c(A1, np.float32)
    if a0.ndim == 1:
        a0 = a0[None]
        a1 = a1[None]
    ops = [_c(x, np.float32) for x in (vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, beam_vals)]
    t = np.ascontiguousarray(tod, dtype=np.float64)
    lib().or_gather_accum_fcorr(*[_ptr(o) for o in ops[:7]], int(nside), _ptr(a0), _ptr(a1), a0.shape[0],
                                _ptr(ops[7]), int(B), int(S), _ptr(t), int(c_q), int(c_u), int(nthreads))
    tod[...] = t
    return tod
