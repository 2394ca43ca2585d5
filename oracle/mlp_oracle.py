"""CPU oracle of the learned interpolation corrector -- TEST INFRASTRUCTURE.

Only tests/ may import this module.  It restates, in float64 numpy, the
corrector the GPU library trains (csrc/corrector.hip):

    x  = [fa - 1/2, v - 1/2, fb - 1/2, fm - 1/2, fp - 1/2, xhat_0..15]   (21)
    h1 = gelu(W1 x + b1), h2 = gelu(W2 h1 + b2), o = W3 h2 + b3
    w  = w_bilinear + o - mean(o)
    p  = sum_k w_k xhat_k,  loss = sum_rows (p - that)^2

with gelu(z) = z * s(z), s the quintic smoothstep 1/2 + (15/16) t - (5/8) t^3 +
(3/16) t^5 of t = clamp(z / 2.64, -1, 1) (GELU's Gaussian CDF replaced by a
polynomial gate; max |gelu(z) - z Phi(z)| = 0.019).  ``bf16=True`` rounds every matrix
product operand to bfloat16 (round to nearest even) as the MFMA kernels do,
so gradients can be compared with the GPU's to FP32-accumulation accuracy;
``bf16=False`` is the exact float64 model, pinned against central finite
differences in tests/test_corrector_oracle.py.  The 4 x 4 neighbourhood
features follow numba_healpy.py:443-480 for the bilinear cell (restated by
oracle/tod_oracle.c) and take rings ira-1 .. ira+2 x slots j-1 .. j+2;
fa, fb, fm, fp are the query's slot fractions in rings ira, ira+1, ira-1,
ira+2 and v the bilinear ring weight.
"""

from __future__ import annotations

import math

import numpy as np

GEO = 5
N_IN, N_H, N_OUT = GEO + 16, 64, 16
SIZES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
         ("W3", (N_OUT, N_H)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(flat):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


GATE_C = 2.64


def gelu(z):
    t = np.clip(z / GATE_C, -1.0, 1.0)
    return z * (0.5 + t * (15.0 / 16.0 + t * t * (-5.0 / 8.0 + t * t * (3.0 / 16.0))))


def gelu_grad(z):
    t = np.clip(z / GATE_C, -1.0, 1.0)
    s = 0.5 + t * (15.0 / 16.0 + t * t * (-5.0 / 8.0 + t * t * (3.0 / 16.0)))
    ds = np.where(np.abs(z) < GATE_C,
                  (15.0 / 16.0 - (15.0 / 8.0) * t * t + (15.0 / 16.0) * t ** 4) / GATE_C, 0.0)
    return s + z * ds


def _mm(a, b, rnd):
    if rnd:
        return bf16(a) @ bf16(b)
    return a @ b


def forward(flat, X, bf16_ops=False):
    p = unpack(flat)
    z1 = _mm(X, p["W1"].T, bf16_ops) + p["b1"]
    a1 = gelu(z1)
    z2 = _mm(a1, p["W2"].T, bf16_ops) + p["b2"]
    a2 = gelu(z2)
    o = _mm(a2, p["W3"].T, bf16_ops) + p["b3"]
    return o, (z1, a1, z2, a2)


def head(O, X, wbil):
    xh = X[:, GEO:]
    w = O - O.mean(axis=1, keepdims=True)
    w[:, 5] += wbil[:, 0]
    w[:, 6] += wbil[:, 1]
    w[:, 9] += wbil[:, 2]
    w[:, 10] += wbil[:, 3]
    return np.sum(w * xh, axis=1)


def loss_grad(flat, X, wbil, that, bf16_ops=False):
    """(loss sum, gradient of loss/N, predictions) -- the quantities of
    todg_corr_loss_grad."""
    X = np.asarray(X, dtype=np.float64)
    N = X.shape[0]
    p = unpack(flat)
    O, (z1, a1, z2, a2) = forward(flat, X, bf16_ops)
    pred = head(O, X, np.asarray(wbil, dtype=np.float64))
    r = pred - np.asarray(that, dtype=np.float64)
    loss = float(np.sum(r * r))
    xh = X[:, GEO:]
    dO = (2.0 / N) * r[:, None] * (xh - xh.mean(axis=1, keepdims=True))
    g = {}
    g["W3"] = _mm(dO.T, a2, bf16_ops)
    g["b3"] = dO.sum(axis=0)
    dz2 = _mm(dO, p["W3"], bf16_ops) * gelu_grad(z2)
    g["W2"] = _mm(dz2.T, a1, bf16_ops)
    g["b2"] = dz2.sum(axis=0)
    dz1 = _mm(dz2, p["W2"], bf16_ops) * gelu_grad(z1)
    g["W1"] = _mm(dz1.T, X, bf16_ops)
    g["b1"] = dz1.sum(axis=0)
    return loss, pack(g), pred


def init_params(seed):
    """He-style initial parameters (the GPU trainer uses the same draw)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if name.startswith("W"):
            fan_in = shape[1]
            d[name] = rng.standard_normal(shape) * math.sqrt(2.0 / fan_in)
        else:
            d[name] = np.zeros(shape)
    d["W3"] *= 0.1  # start near the bilinear weights
    return pack(d).astype(np.float32)


def stencil16(O, nside, z, phi):
    """4 x 4 neighbourhood pixels and the sub-pixel coordinates (fa, fb, v,
    fm, fp) of one query, from the restated ring geometry
    (oracle/tod_oracle.c); None outside rings 2 .. 4 nside - 3."""
    z = max(-1.0, min(1.0, z))
    ira = O.ring_above(nside, z)
    if not (2 <= ira <= 4 * nside - 3):
        return None
    pix, w, _, _ = O.ring_interp(nside, z, phi)
    wa = w[0] + w[1]
    v = w[2] + w[3]
    fa = w[1] / wa if wa > 0 else 0.0
    fb = w[3] / v if v > 0 else 0.0
    p16 = []
    fr = [0.0, 0.0]
    for r in range(4):
        ir = ira - 1 + r
        n, first, phi0, dphi = O.ring_info(nside, ir)
        if r == 1:
            j = int(pix[0] - first)
        elif r == 2:
            j = int(pix[2] - first)
        else:
            tw = math.fmod((phi - phi0) / dphi, n)
            if tw != 0 and tw < 0:
                tw += n
            j = int(tw)
            fr[0 if r == 0 else 1] = tw - j
        for s in range(4):
            p16.append(first + ((j - 1 + s) % n))
    return np.array(p16, dtype=np.int64), fa, fb, v, fr[0], fr[1]
