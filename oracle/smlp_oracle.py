"""CPU oracle of the MLP stencil corrector -- TEST INFRASTRUCTURE.

Only tests/ may import this module.  It restates, in float64 numpy, the
per-pixel model that the GPU library evaluates (csrc/tod_fp_fcorr.hip,
k_smlp_channels) and trains (csrc/corrector.hip):

    F_0 = m(p), F_1 .. F_15  the factorised corrector's stencil features
                             (restated by oracle/tod_oracle.c, or_fcorr_*)
    sigma = sqrt(sum_{j>=1} F_j^2 / 15)
    x  = [2 |z_p| - 1, F_1 / sigma, .., F_15 / sigma]    (0 where sigma = 0)
    o  = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3           16 -> 64 -> 64 -> 32
    A0 = F_0 + sum_{j>=1} o_j F_j,   A1 = sum_{j>=1} o_{16+j} F_j

gelu is the MLP corrector's polynomial-gate GELU (oracle/mlp_oracle.py).
The TOD is T = K0 A0 + K1 A1 for the footprint operator K of a set of
samples (``dense_operator`` builds it column by column from the oracle's
factorised-corrector TOD, or_gather_accum_fcorr); the training loss of one
sky is sum_b (T_b - t_b)^2 / norm.  ``bf16=True`` rounds every matrix
product operand (inputs, weights, activations) to bfloat16 as the MFMA
kernels do (the inputs first to float32, as the GPU stores them), with the
biases added in float64; ``bf16=False`` is the exact model, whose gradient
tests/test_smlp_oracle.py pins against central finite differences.  No
reference counterpart: the corrector is this library's own.
"""

from __future__ import annotations

import math

import numpy as np

import mlp_oracle as M
import oracle as O

N_IN, N_H, N_OUT = 16, 64, 32
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


def features(nside, field):
    """(F (npix, 16), z (npix)): the stencil features of every pixel (unit
    coefficient vectors through the oracle's channel builder) and each
    pixel's ring z."""
    npix = 12 * nside * nside
    F = np.zeros((npix, 16))
    for j in range(16):
        e = np.zeros(32)
        e[j] = 1.0
        F[:, j] = O.fcorr_channels(nside, field, e)[0]
    z = np.array([O.ring_z(nside, O.pixel_ring(nside, p)) for p in range(npix)])
    return F, z


def inputs(F, z):
    """MLP inputs x (npix, 16)."""
    s2 = np.sum(F[:, 1:] ** 2, axis=1)
    sg = np.sqrt(s2 / 15.0)
    inv = np.where(sg > 0.0, 1.0 / np.where(sg > 0.0, sg, 1.0), 0.0)
    x = np.empty_like(F)
    x[:, 0] = 2.0 * np.abs(z) - 1.0
    x[:, 1:] = F[:, 1:] * inv[:, None]
    return x


def _op(a, rnd):
    return M.bf16(a) if rnd else np.asarray(a, dtype=np.float64)


def forward(flat, x, bf16=False):
    """MLP outputs o (n, 32) and the pre-activations / activations."""
    p = unpack(flat)
    x = np.asarray(x, dtype=np.float64)
    if bf16:
        x = x.astype(np.float32)
    z1 = _op(x, bf16) @ _op(p["W1"], bf16).T + p["b1"]
    a1 = M.gelu(z1)
    z2 = _op(a1, bf16) @ _op(p["W2"], bf16).T + p["b2"]
    a2 = M.gelu(z2)
    o = _op(a2, bf16) @ _op(p["W3"], bf16).T + p["b3"]
    return o, (z1, a1, z2, a2)


def head(F, o):
    A0 = F[:, 0] + np.sum(o[:, 1:16] * F[:, 1:], axis=1)
    A1 = np.sum(o[:, 17:32] * F[:, 1:], axis=1)
    return A0, A1


def channels(nside, field, flat, bf16=False):
    F, z = features(nside, field)
    o, _ = forward(flat, inputs(F, z), bf16)
    return head(F, o)


def dense_operator(vec_orig, params, nside, beam_vals, B, S):
    """(K0, K1) as dense (B, npix) float64 matrices: column p is the oracle's
    factorised-corrector TOD of a unit channel at pixel p (tiny nside only)."""
    npix = 12 * nside * nside
    K0 = np.zeros((B, npix))
    K1 = np.zeros((B, npix))
    for p in range(npix):
        e = np.zeros(npix, np.float32)
        e[p] = 1.0
        z = np.zeros(npix, np.float32)
        t = np.zeros((1, B))
        O.gather_accum_fcorr(vec_orig, *params, nside, e, z, beam_vals, B, S, t)
        K0[:, p] = t[0]
        t = np.zeros((1, B))
        O.gather_accum_fcorr(vec_orig, *params, nside, z, e, beam_vals, B, S, t)
        K1[:, p] = t[0]
    return K0, K1


def loss_grad(flat, F, x, K0, K1, target, norm, bf16=False):
    """(loss, gradient) of sum_b (K0 A0 + K1 A1 - target)_b^2 / norm."""
    p = unpack(flat)
    o, (z1, a1, z2, a2) = forward(flat, x, bf16)
    A0, A1 = head(F, o)
    r = K0 @ A0 + K1 @ A1 - target
    loss = float(r @ r) / norm
    rs = (2.0 / norm) * r
    G0 = K0.T @ rs
    G1 = K1.T @ rs
    dO = np.zeros_like(o)
    dO[:, 1:16] = G0[:, None] * F[:, 1:]
    dO[:, 17:32] = G1[:, None] * F[:, 1:]
    g = {}
    g["W3"] = _op(dO, bf16).T @ _op(a2, bf16)
    g["b3"] = dO.sum(axis=0)
    dz2 = (_op(dO, bf16) @ _op(p["W3"], bf16)) * M.gelu_grad(z2)
    g["W2"] = _op(dz2, bf16).T @ _op(a1, bf16)
    g["b2"] = dz2.sum(axis=0)
    dz1 = (_op(dz2, bf16) @ _op(p["W2"], bf16)) * M.gelu_grad(z1)
    g["W1"] = _op(dz1, bf16).T @ _op(x, bf16)
    g["b1"] = dz1.sum(axis=0)
    return loss, pack(g)


def init_params(seed, coef=None, w3_scale=0.0):
    """The package's initial draw (tod_generation_mapbased_beam_amd.smlp
    init_params), restated: He W1, W2; W3 = w3_scale x He; b3 = coef with
    entries 0 and 16 zeroed."""
    rng = np.random.default_rng(seed)
    d = {}
    d["W1"] = rng.standard_normal((N_H, N_IN)) * math.sqrt(2.0 / N_IN)
    d["b1"] = np.zeros(N_H)
    d["W2"] = rng.standard_normal((N_H, N_H)) * math.sqrt(2.0 / N_H)
    d["b2"] = np.zeros(N_H)
    d["W3"] = rng.standard_normal((N_OUT, N_H)) * math.sqrt(2.0 / N_H) * w3_scale
    b3 = np.zeros(N_OUT) if coef is None else np.asarray(coef, dtype=np.float64).copy()
    b3[0] = 0.0
    b3[16] = 0.0
    d["b3"] = b3
    return pack(d).astype(np.float32)
