"""MLP stencil corrector: the learned interpolation corrector as an MLP
evaluated once per map pixel, feeding the factorised corrector's TOD kernel.

Model.  For a rotated beam pixel with bilinear neighbours p_k, weights w_k
(the reference's, numba_healpy.py:358-499) and lower-ring weight g, the
corrected value is, as in ``fcorr``,

    sum_k w_k [A0(p_k) + (1 - g) g A1(p_k)],

but the two channels are per-pixel stencils whose weights an MLP emits from
the pixel's own 4 x 4 ring neighbourhood: with F_0 = m(p), F_1 .. F_15 the
stencil features of ``fcorr`` (ring-interpolated means at ring offsets
dr = j // 4 and ring-step offsets ds = j % 4, minus m(p)) and sigma their
local RMS,

    x  = [2 |z_p| - 1, F_1 / sigma, .., F_15 / sigma]
    o  = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3            16 -> 64 -> 64 -> 32
    A0 = F_0 + sum_{j>=1} o_j F_j,   A1 = sum_{j>=1} o_{16+j} F_j.

The MLP runs once per map pixel and sky (``todg_smlp_channels``: features,
BF16-MFMA layers with FP32 accumulation and channels in one kernel), so the
TOD costs what the factorised corrector's does; the beam pixel's sub-pixel
position enters through its bilinear weights and (1 - g) g.

Training.  The loss is the TOD residual against the bicubic judge,
sum_b (T_b - T_bic,b)^2 / sum_b (T_bil,b - T_bic,b)^2, on synthetic skies and
random pointings.  The corrected TOD is linear in the channels,
T = K0 A0 + K1 A1, with K0 (K1) the footprint operator that deposits each
beam pixel's bilinear weights (times (1 - g) g); K is assembled from the
production kernel's own index dumps (``kernels.fp_dump``) and kept on the
GPU as sparse rows, together with its transpose.  One training step per
sky: the MLP over all its pixels (``todg_smlp_forward``), the channels
(``todg_smlp_head``), the TOD and its residual (``todg_csr_rows2``), the
gradient with respect to the channels through the transpose
(``todg_csr_adjoint2``), back through the head and the MLP
(``todg_smlp_head_backward``, ``todg_smlp_backward``), Adam
(``todg_smlp_adam``).  Every product runs on the library's BF16-MFMA GEMM;
the host only sequences launches.

Initialisation.  W3 = 0 and b3 = the least-squares stencil over the same
training operators (``least_squares_stencil``): the network starts as the
best fixed linear stencil and training moves it from there;
``init="linear"`` takes the factorised corrector's shipped coefficients
(``fcorr_coef.json``) instead, ``init="bilinear"`` plain bilinear
interpolation (b3 = 0).
"""

from __future__ import annotations

import json
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import fcorr
from . import inputs
from .device import device, stream_handle, to_dev
from .kernels import fp_dump, gather_accum

N_IN, N_H, N_OUT = 16, 64, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_H * N_IN
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_OUT * N_H
N_PARAMS = OFF_B3 + N_OUT
assert N_PARAMS == 7328

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "smlp_params.json")


def init_params(seed: int = 0, coef=None, w3_scale: float = 0.0) -> np.ndarray:
    """He-initialised W1, W2 (zero biases); W3 = w3_scale x He; b3 = the
    stencil ``coef`` (32 values as fcorr's: A0's weights 0..15 with the
    pixel's own fixed at 1, A1's 16..31), its entries 0 and 16 unused (0).
    float32, flat layout W1 b1 W2 b2 W3 b3."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS)
    p[OFF_W1:OFF_B1] = (rng.standard_normal((N_H, N_IN)) * math.sqrt(2.0 / N_IN)).ravel()
    p[OFF_W2:OFF_B2] = (rng.standard_normal((N_H, N_H)) * math.sqrt(2.0 / N_H)).ravel()
    p[OFF_W3:OFF_B3] = (rng.standard_normal((N_OUT, N_H)) * math.sqrt(2.0 / N_H) * w3_scale).ravel()
    if coef is not None:
        c = np.asarray(coef, dtype=np.float64)
        if c.shape != (32,):
            raise ValueError("coef must hold 32 values")
        b3 = c.copy()
        b3[0] = 0.0
        b3[16] = 0.0
        p[OFF_B3:] = b3
    return p.astype(np.float32)


def _check_params(params) -> torch.Tensor:
    n = int(params.numel()) if isinstance(params, torch.Tensor) else int(np.asarray(params).size)
    if n != N_PARAMS:
        raise ValueError(f"expected {N_PARAMS} parameters")
    return to_dev(params, torch.float32).reshape(-1).contiguous()


def _check_nside(nside) -> int:
    nside = int(nside)
    if nside < 1 or nside > 8192 or nside & (nside - 1):
        raise ValueError("nside must be a power of two <= 8192")
    return nside


def _check_dev(t, name: str, dtype=torch.float32, min_numel: int = 0, cols: int | None = None) -> None:
    """t must be a contiguous CUDA tensor of dtype on the current device with
    at least min_numel elements (and, with cols, a 2-D (rows, cols) shape):
    the kernels take its raw pointer.  Checked before any device work."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor on the GPU")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must live on the GPU (got a {t.device.type} tensor)")
    if t.device.index is not None and t.device.index != torch.cuda.current_device():
        raise ValueError(f"{name} is on cuda:{t.device.index}, not the current device")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < min_numel:
        raise ValueError(f"{name} holds {t.numel()} elements, needs {min_numel}")
    if cols is not None and (t.dim() != 2 or int(t.shape[1]) != cols):
        raise ValueError(f"{name} must be (rows, {cols})")


def channels(params, nside: int, field, out0, out1, stride: int = 1) -> None:
    """A0 and A1 of one map field (device float32, npix) into out0[p * stride]
    and out1[p * stride] (device float32 views): one fused kernel."""
    nside = _check_nside(nside)
    npix = 12 * nside * nside
    stride = int(stride)
    if stride < 1:
        raise ValueError("stride must be >= 1")
    nfield = int(field.numel()) if isinstance(field, torch.Tensor) else int(np.asarray(field).size)
    if nfield != npix:
        raise ValueError(f"field holds {nfield} values, nside {nside} has {npix} pixels")
    _check_dev(out0, "out0", min_numel=(npix - 1) * stride + 1)
    _check_dev(out1, "out1", min_numel=(npix - 1) * stride + 1)
    P = _check_params(params)
    f = to_dev(field, torch.float32).reshape(-1).contiguous()
    _lib.check(_lib.lib().todg_smlp_channels(int(nside), _lib.ptr(f), _lib.ptr(P), _lib.ptr(out0),
                                              _lib.ptr(out1), int(stride), stream_handle()))


def features(nside: int, field):
    """(F, X): the per-pixel features (npix x 16) and MLP inputs (npix x 16)."""
    nside = _check_nside(nside)
    nfield = int(field.numel()) if isinstance(field, torch.Tensor) else int(np.asarray(field).size)
    if nfield != 12 * nside * nside:
        raise ValueError(f"field holds {nfield} values, nside {nside} has {12 * nside * nside} pixels")
    f = to_dev(field, torch.float32).reshape(-1).contiguous()
    npix = int(f.numel())
    F = torch.empty((npix, 16), dtype=torch.float32, device=f.device)
    X = torch.empty((npix, 16), dtype=torch.float32, device=f.device)
    _lib.check(_lib.lib().todg_smlp_features(int(nside), _lib.ptr(f), _lib.ptr(F), _lib.ptr(X),
                                              stream_handle()))
    return F, X


class Workspace:
    """Device buffers of one training step over N pixel rows."""

    def __init__(self, N: int, dev):
        self.N = int(N)
        nf = int(_lib.lib().todg_smlp_workspace_floats(self.N))
        self.ws = torch.empty(nf, dtype=torch.float32, device=dev)
        self.O = torch.empty((self.N, N_OUT), dtype=torch.float32, device=dev)
        self.dO = torch.empty_like(self.O)
        self.A0 = torch.empty(self.N, dtype=torch.float32, device=dev)
        self.A1 = torch.empty_like(self.A0)
        self.G0 = torch.empty_like(self.A0)
        self.G1 = torch.empty_like(self.A0)
        self.grad = torch.empty(N_PARAMS, dtype=torch.float32, device=dev)
        self.sumsq = torch.zeros(1, dtype=torch.float64, device=dev)


def _check_rows(X, ws: Workspace) -> int:
    _check_dev(X, "X", cols=N_IN)
    N = int(X.shape[0])
    if N < 1 or N > ws.N:
        raise ValueError(f"{N} rows, the workspace holds 1 .. {ws.N}")
    return N


def mlp_forward(params, X, ws: Workspace):
    """O = MLP(X) over the rows of X (activations kept in ws for backward)."""
    N = _check_rows(X, ws)
    _check_dev(params, "params", min_numel=N_PARAMS)
    _lib.check(_lib.lib().todg_smlp_forward(_lib.ptr(params), _lib.ptr(X), N, _lib.ptr(ws.ws), _lib.ptr(ws.O),
                                             stream_handle()))
    return ws.O


def mlp_backward(params, X, ws: Workspace, dO):
    """Parameter gradient (ws.grad) of sum_i dO_i . o_i over the rows of the
    preceding mlp_forward, whose activations in ws it consumes."""
    N = _check_rows(X, ws)
    _check_dev(params, "params", min_numel=N_PARAMS)
    _check_dev(dO, "dO", min_numel=N * N_OUT)
    _lib.check(_lib.lib().todg_smlp_backward(_lib.ptr(params), _lib.ptr(X), N, _lib.ptr(ws.ws), _lib.ptr(dO),
                                              _lib.ptr(ws.grad), stream_handle()))
    return ws.grad


def head(F, O, A0, A1):
    """Channels A0, A1 from the MLP outputs O and the features F."""
    _check_dev(F, "F", cols=16)
    N = int(F.shape[0])
    _check_dev(O, "O", min_numel=N * N_OUT)
    _check_dev(A0, "A0", min_numel=N)
    _check_dev(A1, "A1", min_numel=N)
    _lib.check(_lib.lib().todg_smlp_head(int(F.shape[0]), _lib.ptr(F), _lib.ptr(O), _lib.ptr(A0), _lib.ptr(A1),
                                          stream_handle()))


def head_backward(F, G0, G1, dO):
    """dO from the loss gradients G0, G1 with respect to the channels."""
    _check_dev(F, "F", cols=16)
    N = int(F.shape[0])
    _check_dev(G0, "G0", min_numel=N)
    _check_dev(G1, "G1", min_numel=N)
    _check_dev(dO, "dO", min_numel=N * N_OUT)
    _lib.check(_lib.lib().todg_smlp_head_backward(int(F.shape[0]), _lib.ptr(F), _lib.ptr(G0), _lib.ptr(G1),
                                                   _lib.ptr(dO), stream_handle()))


@dataclass
class TodOperator:
    """The corrected TOD of one sky's training samples as a linear map of the
    two channels: rows b (CSR: rowptr, col = pixel, v0 = K0, v1 = K1) and the
    transpose stored by pixel (rowptr_t, col_t = sample, v0_t, v1_t); the
    bicubic judge's TOD ``target``, the plain bilinear residual's sum of
    squares ``norm``, and the sky's features F and MLP inputs X."""
    nside: int
    B: int
    rowptr: torch.Tensor
    col: torch.Tensor
    v0: torch.Tensor
    v1: torch.Tensor
    rowptr_t: torch.Tensor
    col_t: torch.Tensor
    v0_t: torch.Tensor
    v1_t: torch.Tensor
    target: torch.Tensor
    norm: float
    F: torch.Tensor
    X: torch.Tensor

    @property
    def npix(self) -> int:
        return 12 * self.nside * self.nside

    def apply(self, A0, A1, out=None):
        """y = K0 A0 + K1 A1 (float64, B)."""
        _check_dev(A0, "A0", min_numel=self.npix)
        _check_dev(A1, "A1", min_numel=self.npix)
        if out is not None:
            _check_dev(out, "out", dtype=torch.float64, min_numel=self.B)
        y = torch.empty(self.B, dtype=torch.float64, device=self.v0.device) if out is None else out
        _lib.check(_lib.lib().todg_csr_rows2(self.B, _lib.ptr(self.rowptr), _lib.ptr(self.col), _lib.ptr(self.v0),
                                              _lib.ptr(self.v1), _lib.ptr(A0), _lib.ptr(A1), None, 1.0,
                                              _lib.ptr(y), None, None, None, stream_handle()))
        return y

    def residual(self, A0, A1, scale, resid, part, sumsq):
        """resid = scale (K0 A0 + K1 A1 - target) (float32), sumsq = its
        unscaled sum of squares."""
        _check_dev(A0, "A0", min_numel=self.npix)
        _check_dev(A1, "A1", min_numel=self.npix)
        _check_dev(resid, "resid", min_numel=self.B)
        _check_dev(part, "part", dtype=torch.float64, min_numel=int(_lib.lib().todg_csr_rows2_parts(self.B)))
        _check_dev(sumsq, "sumsq", dtype=torch.float64, min_numel=1)
        _lib.check(_lib.lib().todg_csr_rows2(self.B, _lib.ptr(self.rowptr), _lib.ptr(self.col), _lib.ptr(self.v0),
                                              _lib.ptr(self.v1), _lib.ptr(A0), _lib.ptr(A1), _lib.ptr(self.target),
                                              float(scale), None, _lib.ptr(resid), _lib.ptr(part), _lib.ptr(sumsq),
                                              stream_handle()))

    def adjoint(self, r, G0, G1):
        """G0 = K0^T r, G1 = K1^T r (float32, npix)."""
        _check_dev(r, "r", min_numel=self.B)
        _check_dev(G0, "G0", min_numel=self.npix)
        _check_dev(G1, "G1", min_numel=self.npix)
        _lib.check(_lib.lib().todg_csr_adjoint2(self.npix, _lib.ptr(self.rowptr_t), _lib.ptr(self.col_t),
                                                 _lib.ptr(self.v0_t), _lib.ptr(self.v1_t), _lib.ptr(r), _lib.ptr(G0),
                                                 _lib.ptr(G1), stream_handle()))


def build_operator(nside: int, field, vec_orig, beam_vals, params, chunk: int = 256) -> TodOperator:
    """The footprint operator of one sky's samples from the production
    kernel's index dumps: per sample, every beam pixel's four deposits
    (beam value x bilinear weight, and the same times (1 - g) g, g the lower
    ring's share of the four) summed per map pixel.  Deterministic: the
    per-pixel sums are differences of a binary64 prefix sum in sorted order."""
    nside = _check_nside(nside)
    nfield = int(field.numel()) if isinstance(field, torch.Tensor) else int(np.asarray(field).size)
    if nfield != 12 * nside * nside:
        raise ValueError("field size does not match nside")
    if len(params) != 6:
        raise ValueError("params must be the six rotation operand arrays")
    sizes = [int(p.numel()) if isinstance(p, torch.Tensor) else int(np.asarray(p).size) for p in params]
    B = sizes[1]
    if sizes != [3 * B, B, B, 3 * B, B, B] or B < 1:
        raise ValueError("rotation operands: axes (B, 3), cos_a, sin_a (B), ax_pts (B, 3), cos_p, sin_p (B)")
    nvec = int(vec_orig.numel()) if isinstance(vec_orig, torch.Tensor) else int(np.asarray(vec_orig).size)
    nbv = int(beam_vals.numel()) if isinstance(beam_vals, torch.Tensor) else int(np.asarray(beam_vals).size)
    if nbv < 1 or nvec != 3 * nbv:
        raise ValueError("vec_orig must be (S, 3) for beam_vals (S,)")
    if int(chunk) < 1:
        raise ValueError("chunk must be >= 1")
    dev = device()
    f = to_dev(field, torch.float32).reshape(-1).contiguous()
    npix = int(f.numel())
    vec = to_dev(vec_orig, torch.float32).contiguous()
    bv = to_dev(beam_vals, torch.float32).contiguous()
    P = [to_dev(p, torch.float32).contiguous() for p in params]
    S = int(bv.numel())
    counts, cols, vals0, vals1 = [], [], [], []
    for b0 in range(0, B, chunk):
        nb = min(chunk, B - b0)
        sub = [p[b0:b0 + nb] for p in P]
        pix, w, _, _ = fp_dump("bilinear", vec, *sub, nside, f[None], bv, nb, S, -1, -1, host=False)
        wsum = w.sum(-1)
        g = torch.where(wsum != 0, (w[..., 2] + w[..., 3]) / torch.where(wsum != 0, wsum, 1.0), 0.0)
        w1 = w * (g * (1.0 - g))[..., None]
        bidx = torch.arange(nb, device=dev, dtype=torch.int64)[:, None, None]
        key = (bidx * npix + pix).reshape(-1)
        key, order = torch.sort(key, stable=True)
        cs0 = torch.cumsum(w.reshape(-1)[order], 0)
        cs1 = torch.cumsum(w1.reshape(-1)[order], 0)
        uk, cnt = torch.unique_consecutive(key, return_counts=True)
        end = torch.cumsum(cnt, 0) - 1
        s0 = cs0[end] - torch.cat([cs0.new_zeros(1), cs0[end[:-1]]])
        s1 = cs1[end] - torch.cat([cs1.new_zeros(1), cs1[end[:-1]]])
        counts.append(torch.bincount(uk // npix, minlength=nb))
        cols.append((uk % npix).to(torch.int32))
        vals0.append(s0.to(torch.float32))
        vals1.append(s1.to(torch.float32))
        del pix, w, g, w1, key, order, cs0, cs1
    cnt = torch.cat(counts)
    col = torch.cat(cols)
    v0 = torch.cat(vals0)
    v1 = torch.cat(vals1)
    rowptr = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev), torch.cumsum(cnt, 0)])
    # transpose: entries by pixel (stable, so by sample within a pixel)
    col_sorted, ordt = torch.sort(col.to(torch.int64), stable=True)
    samp = torch.repeat_interleave(torch.arange(B, device=dev, dtype=torch.int64), cnt)
    col_t = samp[ordt].to(torch.int32)
    rowptr_t = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev),
                          torch.cumsum(torch.bincount(col_sorted, minlength=npix), 0)])
    v0_t = v0[ordt].contiguous()
    v1_t = v1[ordt].contiguous()
    # the judge's TOD and the plain bilinear residual
    tc = torch.zeros((1, B), dtype=torch.float64, device=dev)
    gather_accum("bicubic", vec, *P, nside, f[None], bv, B, S, tc)
    target = tc[0].contiguous()
    F, X = features(nside, f)
    op = TodOperator(nside, B, rowptr, col, v0, v1, rowptr_t, col_t, v0_t, v1_t, target, 0.0, F, X)
    zero = torch.zeros(npix, dtype=torch.float32, device=dev)
    y = op.apply(f, zero)
    op.norm = float(torch.sum((y - target) ** 2))
    return op


@dataclass
class TrainReport:
    n_maps: int
    samples_per_map: int
    steps: int
    seconds: float
    start_ratio: float          # RMS residual ratio at the initial parameters (training TODs)
    train_ratio: float          # ... after training
    history: list = field(default_factory=list)   # (step, mean ratio over the last sweep of skies)


class StencilMLPCorrector:
    """FP32 master parameters on the device and their use."""

    def __init__(self, params=None):
        p = init_params(0, fcorr.FactorisedCorrector.load_default().coef) if params is None else params
        self.params = _check_params(p).clone()

    # ------------------------------------------------------------ use
    def channels(self, nside, field, out0, out1, stride=1):
        channels(self.params, nside, field, out0, out1, stride)

    def maps(self, mp_stacked, c_q: int = -1, c_u: int = -1) -> fcorr.FcorrMaps:
        return fcorr.FcorrMaps(mp_stacked, None, c_q, c_u, builder=self.channels)

    def tod(self, maps: fcorr.FcorrMaps, vec_orig, beam_vals, params, tod=None):
        B = int(params[1].numel()) if isinstance(params[1], torch.Tensor) else int(np.asarray(params[1]).shape[0])
        S = int(beam_vals.numel()) if isinstance(beam_vals, torch.Tensor) else int(np.asarray(beam_vals).shape[0])
        if tod is None:
            tod = torch.zeros((maps.C, B), dtype=torch.float64, device=device())
        return fcorr.gather_accum_fcorr(vec_orig, *params, maps.nside, maps, beam_vals, B, S, tod)

    # ------------------------------------------------------------ training
    def ratio(self, op: TodOperator, ws: Workspace) -> float:
        """RMS residual ratio of the current parameters on one operator."""
        mlp_forward(self.params, op.X, ws)
        head(op.F, ws.O, ws.A0, ws.A1)
        y = op.apply(ws.A0, ws.A1)
        return math.sqrt(float(torch.sum((y - op.target) ** 2)) / op.norm)

    def train(self, ops, epochs: int, lr: float = 1e-3, lr_out: float = 1e-6, seed: int = 0,
              log=None) -> TrainReport:
        """Adam over the skies' operators, one step per sky, the skies in a
        seeded order each epoch; the hidden layers at lr, the output layer at
        lr_out (both warmed up linearly over the first tenth of the steps,
        then decayed to 0 on a cosine).  The output layer gets its own,
        much smaller rate because the TOD residual is some 1e-3 of the
        pixel-scale signal: a stencil weight off by 1e-3 already doubles it."""
        t0 = time.time()
        dev = self.params.device
        npix = max(op.npix for op in ops)
        ws = Workspace(npix, dev)
        m = torch.zeros(N_PARAMS, dtype=torch.float32, device=dev)
        v = torch.zeros_like(m)
        nb = max(op.B for op in ops)
        resid = torch.empty(nb, dtype=torch.float32, device=dev)
        part = torch.empty(int(_lib.lib().todg_csr_rows2_parts(nb)), dtype=torch.float64, device=dev)
        start = float(np.mean([self.ratio(op, ws) for op in ops]))
        total = epochs * len(ops)
        warm = max(1, total // 10)  # linear warm-up over the first tenth of the steps
        rng = np.random.default_rng(seed)
        step = 0
        hist = []
        for ep in range(epochs):
            acc = 0.0
            for i in rng.permutation(len(ops)):
                op = ops[i]
                n = op.npix
                mlp_forward(self.params, op.X, ws)
                head(op.F, ws.O, ws.A0, ws.A1)
                op.residual(ws.A0, ws.A1, 2.0 / op.norm, resid, part, ws.sumsq)
                op.adjoint(resid, ws.G0, ws.G1)
                head_backward(op.F, ws.G0[:n], ws.G1[:n], ws.dO)
                mlp_backward(self.params, op.X, ws, ws.dO)
                step += 1
                decay = min(1.0, step / warm) * 0.5 * (1.0 + math.cos(math.pi * (step - 1) / total))
                _lib.check(_lib.lib().todg_smlp_adam(_lib.ptr(self.params), _lib.ptr(ws.grad), _lib.ptr(m),
                                                      _lib.ptr(v), step, lr * decay, lr_out * decay, 0.9, 0.999,
                                                      1e-8, stream_handle()))
                acc += float(ws.sumsq) / op.norm
            hist.append((step, math.sqrt(acc / len(ops))))
            if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
                log(f"smlp train: epoch {ep + 1}/{epochs} ratio {hist[-1][1]:.4f} ({time.time() - t0:.1f}s)")
        final = float(np.mean([self.ratio(op, ws) for op in ops]))
        return TrainReport(len(ops), ops[0].B, step, time.time() - t0, start, final, hist)

    @staticmethod
    def least_squares_stencil(ops, ridge: float = 1e-10) -> np.ndarray:
        """The 32 stencil coefficients (fcorr's layout: A0's weight of the
        pixel itself fixed at 1, A1's at 0) that minimise the TOD residual
        over the operators' samples: each free coefficient's feature is the
        operator applied to that stencil feature, sum_p K(b, p) F_j(p); the
        30 x 30 normal equations are solved in binary64."""
        cols = [(0, j) for j in range(1, 16)] + [(1, j) for j in range(1, 16)]
        G = np.zeros((len(cols), len(cols)))
        h = np.zeros(len(cols))
        for op in ops:
            zero = torch.zeros(op.npix, dtype=torch.float32, device=op.F.device)
            X = np.empty((op.B, len(cols)))
            for k, (ch, j) in enumerate(cols):
                fj = op.F[:, j].contiguous()
                X[:, k] = (op.apply(fj, zero) if ch == 0 else op.apply(zero, fj)).cpu().numpy()
            r = (op.target - op.apply(op.F[:, 0].contiguous(), zero)).cpu().numpy()
            G += X.T @ X
            h += X.T @ r
        x = np.linalg.solve(G + ridge * np.trace(G) / G.shape[0] * np.eye(G.shape[0]), h)
        coef = np.zeros(32)
        coef[0] = 1.0
        coef[1:16] = x[:15]
        coef[17:32] = x[15:]
        return coef

    @classmethod
    def fit(cls, nside: int = 512, n_maps: int = 8, samples_per_map: int = 40_000, epochs: int = 40,
            lr: float = 1e-3, lr_out: float = 1e-6, init: str = "lsq", seed: int = 100,
            fwhm_arcmin: float | None = None, power_cut: float = 0.99, log=None):
        """Operators of n_maps synthetic skies (seeds seed .. seed + n_maps - 1,
        spectrum of inputs.cl_spectrum, I field) at ``nside`` with
        ``samples_per_map`` random pointings each and a Gaussian beam of FWHM
        ``fwhm_arcmin`` (default: 30' scaled to nside 2048's pixel size) at
        nside_beam 2 nside, then ``train``.  Returns (corrector, report,
        operators)."""
        if init not in ("lsq", "linear", "bilinear"):
            raise ValueError("init must be 'lsq', 'linear' or 'bilinear'")
        dev = device()
        if fwhm_arcmin is None:
            fwhm_arcmin = 30.0 * 2048 / nside
        vec, bv, _ = inputs.gaussian_beam(2 * nside, fwhm_arcmin, power_cut)
        dvec = torch.as_tensor(np.ascontiguousarray(vec, np.float32), device=dev)
        dbv = torch.as_tensor(np.ascontiguousarray(bv, np.float32), device=dev)
        t0 = time.time()
        ops = []
        for i in range(n_maps):
            params = [torch.as_tensor(p, device=dev) for p in fcorr.random_pointings(samples_per_map, seed + 1000 + i)]
            m = inputs.synfast_iqu(nside, seed=seed + i, device=dev)[0]
            fld = torch.as_tensor(m, device=dev, dtype=torch.float32).contiguous()
            ops.append(build_operator(nside, fld, dvec, dbv, params))
            if log:
                log(f"smlp operators: sky {i + 1}/{n_maps} ({time.time() - t0:.1f}s)")
        if init == "lsq":
            coef = cls.least_squares_stencil(ops)
        elif init == "linear":
            coef = fcorr.FactorisedCorrector.load_default().coef
        else:
            coef = fcorr.identity_coef()
        corr = cls(init_params(seed, coef))
        rep = corr.train(ops, epochs, lr=lr, lr_out=lr_out, seed=seed, log=log)
        return corr, rep, ops

    # ------------------------------------------------------------ storage
    def to_json(self, **meta) -> str:
        p = self.params.detach().cpu().numpy()
        return json.dumps({"kind": "mlp-stencil-corrector", "layers": [N_IN, N_H, N_H, N_OUT],
                           "params": [float(x) for x in p], **meta})

    @classmethod
    def from_json(cls, text: str) -> "StencilMLPCorrector":
        d = json.loads(text)
        if d.get("kind") != "mlp-stencil-corrector" or d.get("layers") != [N_IN, N_H, N_H, N_OUT]:
            raise ValueError("not an MLP stencil corrector parameter file")
        return cls(np.asarray(d["params"], dtype=np.float32))

    @classmethod
    def load_default(cls) -> "StencilMLPCorrector":
        """The parameters shipped with the package (DEFAULT_PATH, written by
        ``python -m tod_generation_mapbased_beam_amd.smlp --fit``)."""
        with open(DEFAULT_PATH) as f:
            return cls.from_json(f.read())


def main(argv=None):
    import argparse
    import sys

    ap = argparse.ArgumentParser(description="train the MLP stencil corrector on the GPU")
    ap.add_argument("--fit", action="store_true")
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--nside", type=int, default=512)
    ap.add_argument("--maps", type=int, default=8)
    ap.add_argument("--samples", type=int, default=40_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-out", type=float, default=1e-6)
    ap.add_argument("--init", default="lsq", choices=["lsq", "linear", "bilinear"])
    args = ap.parse_args(argv)
    if not args.fit:
        ap.error("nothing to do (use --fit)")
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    corr, rep, _ = StencilMLPCorrector.fit(nside=args.nside, n_maps=args.maps, samples_per_map=args.samples,
                                           epochs=args.epochs, lr=args.lr, lr_out=args.lr_out, init=args.init,
                                           log=log)
    with open(args.out, "w") as f:
        f.write(corr.to_json(trained_on=f"{rep.n_maps} synthetic nside-{args.nside} skies x {rep.samples_per_map} "
                                        f"random pointings, 30' beam scaled to the pixel size",
                             init=args.init, epochs=args.epochs, lr=args.lr, lr_out=args.lr_out,
                             start_ratio=rep.start_ratio,
                             train_rms_ratio=rep.train_ratio) + "\n")
    log(f"training RMS ratio {rep.start_ratio:.4f} -> {rep.train_ratio:.4f} ({rep.seconds:.1f}s) -> {args.out}")


if __name__ == "__main__":
    main()
