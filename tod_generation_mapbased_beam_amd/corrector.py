"""Learned interpolation corrector: training data, training loop and point
prediction on the GPU (kernels in csrc/corrector.hip).

The corrector is an MLP 21 -> 64 -> 64 -> 16 (GELU) that looks at the
sub-pixel position of a direction in the 4 x 4 ring neighbourhood of its
bilinear cell (the slot fractions in the four rings and the bilinear ring
weight) and at the neighbourhood's 16 map values (normalised by their mean
and RMS), and emits corrections to the 16 interpolation weights (mean
removed, so constant maps stay exact).  It is trained on (bilinear input, bicubic target) pairs: the
target is the bicubic judge (Keys/Catmull-Rom weights on the gnomonic tangent
plane, README.md:120 of the reference) at uniformly random directions on
synthetic skies with the spectrum C_l^TT = 1e3/(l(l+1)+1).

Training runs on the library's own kernels: BF16-MFMA matrix products with
FP32 accumulation, FP32 master weights, Adam.  Nothing here falls back to a
CPU path.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, inputs, kernels
from .device import device, stream_handle, to_dev

N_PARAMS = 6608
GEO = 5
N_IN, N_H, N_OUT = GEO + 16, 64, 16


def init_params(seed: int) -> np.ndarray:
    """He-style initial parameters (W3 scaled by 0.1 so training starts near
    plain bilinear weights); float32, flat layout W1 b1 W2 b2 W3 b3."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape in ((N_H, N_IN), (N_H,), (N_H, N_H), (N_H,), (N_OUT, N_H), (N_OUT,)):
        if len(shape) == 2:
            w = rng.standard_normal(shape) * math.sqrt(2.0 / shape[1])
            if shape[0] == N_OUT:
                w *= 0.1
        else:
            w = np.zeros(shape)
        parts.append(w.ravel())
    return np.concatenate(parts).astype(np.float32)


@dataclass
class Features:
    X: torch.Tensor      # (N, 21) float32
    wbil: torch.Tensor   # (N, 4) float32
    mu: torch.Tensor     # (N,) float32
    sigma: torch.Tensor  # (N,) float32
    pix: torch.Tensor    # (N, 16) int32
    ok: torch.Tensor     # (N,) int32


def features(nside: int, field, z, phi) -> Features:
    """Corrector inputs of directions (z, phi) on one RING map field."""
    f = to_dev(field, torch.float32).reshape(-1).contiguous()
    zz = to_dev(z, torch.float64).reshape(-1).contiguous()
    pp = to_dev(phi, torch.float64).reshape(-1).contiguous()
    N = zz.numel()
    dev = zz.device
    X = torch.empty((N, N_IN), dtype=torch.float32, device=dev)
    wb = torch.empty((N, 4), dtype=torch.float32, device=dev)
    mu = torch.empty(N, dtype=torch.float32, device=dev)
    sg = torch.empty(N, dtype=torch.float32, device=dev)
    pix = torch.empty((N, 16), dtype=torch.int32, device=dev)
    ok = torch.empty(N, dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().todg_corr_features(int(nside), _lib.ptr(f), N, _lib.ptr(zz), _lib.ptr(pp),
                                             _lib.ptr(X), _lib.ptr(wb), _lib.ptr(mu), _lib.ptr(sg),
                                             _lib.ptr(pix), _lib.ptr(ok), stream_handle()))
    return Features(X, wb, mu, sg, pix, ok)


def random_directions(n: int, gen: torch.Generator, dev):
    """Uniform directions on the sphere: z ~ U(-1, 1), phi ~ U(0, 2 pi)."""
    z = torch.rand(n, dtype=torch.float64, device=dev, generator=gen) * 2.0 - 1.0
    ph = torch.rand(n, dtype=torch.float64, device=dev, generator=gen) * (2.0 * math.pi)
    return z, ph


def make_pairs(nside: int, seeds, n_per_map: int, lmax: int | None = None, dir_seed: int = 0):
    """Training pairs on synthetic skies: for each map seed, n_per_map random
    directions with their features and normalised bicubic targets
    ((bicubic - mu) / sigma).  Only directions where the corrector applies
    are kept.  Returns (X, wbil, that) on the device."""
    dev = device()
    lmax = 3 * nside - 1 if lmax is None else lmax
    tt, _, _ = inputs.cl_spectrum(lmax)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(dir_seed))
    Xs, Ws, Ts = [], [], []
    for seed in seeds:
        rng = np.random.default_rng(int(seed))
        alm = inputs._alm_realisation(tt, lmax, rng)
        field = torch.as_tensor(inputs.synthesize_map(nside, [alm], lmax, device=dev)[0],
                                dtype=torch.float32, device=dev)
        z, ph = random_directions(n_per_map, gen, dev)
        F = features(nside, field, z, ph)
        target = kernels.interp_points(nside, field, z, ph, "bicubic")
        keep = (F.ok != 0) & (F.sigma > 0)
        that = ((target - F.mu.double()) / F.sigma.double().clamp_min(1e-30)).float()
        Xs.append(F.X[keep])
        Ws.append(F.wbil[keep])
        Ts.append(that[keep])
    return torch.cat(Xs), torch.cat(Ws), torch.cat(Ts)


class Corrector:
    """FP32 master parameters on the device and the kernels that use them."""

    def __init__(self, params=None, seed: int = 0):
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters")
        self.params = torch.as_tensor(p, device=device()).contiguous()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._ws = None

    def _workspace(self, N: int) -> torch.Tensor:
        need = int(_lib.lib().todg_corr_workspace_floats(int(N)))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.float32, device=self.params.device)
        return self._ws

    def loss_grad(self, X, wbil, that, want_grad=True, want_pred=False):
        """(loss sum, gradient of loss/N or None, normalised predictions or None)."""
        N = int(X.shape[0])
        ws = self._workspace(N)
        grad = torch.empty_like(self.params) if want_grad else None
        loss = torch.zeros(1, dtype=torch.float64, device=self.params.device)
        pred = torch.empty(N, dtype=torch.float32, device=self.params.device) if want_pred else None
        _lib.check(_lib.lib().todg_corr_loss_grad(
            _lib.ptr(self.params), _lib.ptr(X), _lib.ptr(wbil), _lib.ptr(that), N, _lib.ptr(ws),
            _lib.ptr(grad), _lib.ptr(loss), _lib.ptr(pred), stream_handle()))
        return loss, grad, pred

    def adam(self, grad, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _lib.check(_lib.lib().todg_corr_adam(_lib.ptr(self.params), _lib.ptr(grad), _lib.ptr(self.m),
                                             _lib.ptr(self.v), self.step, lr, beta1, beta2, eps,
                                             stream_handle()))

    def train(self, X, wbil, that, epochs=1, batch=8192, lr=1e-3, seed=0, log=None,
              schedule="constant"):
        """Adam over shuffled mini-batches; returns the mean loss per epoch.
        schedule: "constant" (lr throughout) or "cosine" (lr decayed to zero
        over all steps along a half cosine)."""
        if schedule not in ("constant", "cosine"):
            raise ValueError("schedule must be 'constant' or 'cosine'")
        N = int(X.shape[0])
        gen = torch.Generator(device=X.device)
        gen.manual_seed(int(seed))
        steps_per_epoch = max(1, N // batch)
        total = epochs * steps_per_epoch
        hist = []
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(N, generator=gen, device=X.device)
            tot = torch.zeros((), dtype=torch.float64, device=X.device)
            for i0 in range(0, N - batch + 1, batch):
                idx = perm[i0:i0 + batch]
                loss, grad, _ = self.loss_grad(X[idx].contiguous(), wbil[idx].contiguous(),
                                               that[idx].contiguous())
                step_lr = lr if schedule == "constant" else 0.5 * lr * (1.0 + math.cos(math.pi * k / total))
                self.adam(grad, lr=step_lr)
                tot += loss[0]
                k += 1
            mean = float(tot) / max(1, (N // batch) * batch)
            hist.append(mean)
            if log:
                log(f"epoch {ep}: mean loss {mean:.4e}")
        return hist

    def eval_loss(self, X, wbil, that, chunk=1 << 20):
        """Mean normalised squared residual of the corrector on (X, that)."""
        N = int(X.shape[0])
        tot = 0.0
        for i0 in range(0, N, chunk):
            loss, _, _ = self.loss_grad(X[i0:i0 + chunk].contiguous(), wbil[i0:i0 + chunk].contiguous(),
                                        that[i0:i0 + chunk].contiguous(), want_grad=False)
            tot += float(loss[0])
        return tot / N

    def predict_points(self, nside: int, field, z, phi) -> torch.Tensor:
        """Corrected interpolation of one field at (z, phi) (float64); plain
        bilinear where the corrector does not apply."""
        F = features(nside, field, z, phi)
        N = int(F.X.shape[0])
        zero = torch.zeros(N, dtype=torch.float32, device=F.X.device)
        _, _, pred = self.loss_grad(F.X, F.wbil, zero, want_grad=False, want_pred=True)
        val = F.mu.double() + F.sigma.double() * pred.double()
        bil = kernels.interp_points(nside, field, z, phi, "bilinear")
        return torch.where(F.ok != 0, val, to_dev(bil, torch.float64))


def gather_accum_learned(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, mp_stacked,
                         beam_vals, B, S, tod, c_q=-1, c_u=-1, params=None, bound=float("inf"),
                         stats=None, power=1.0):
    """Beam TOD with the corrector fused into the gather (same arguments as
    ``kernels._gather_accum_fused_jit`` plus the corrector's parameters);
    accumulates into ``tod``.  The corrector reads the first field outside
    the Q/U pair (I) and its corrected weights serve every field (Q/U
    rotated by spin-2 per pixel).  ``bound``: largest |correction| applied,
    in units of the neighbourhood RMS (plain bilinear beyond it).
    ``power`` < 1: only the strongest beam pixels holding that fraction of
    the beam's total |value| take the corrector, the weaker tail plain
    bilinear (the beam is reordered by value for the call).
    ``stats``: optional int64 device tensor (2,) that receives (corrected,
    plain bilinear) beam-pixel-sample counts."""
    from .device import device_map_for
    from .kernels import _map_channels, check_tod

    B, S, c_q, c_u = int(B), int(S), int(c_q), int(c_u)
    if not 0.0 < float(power) <= 1.0:
        raise ValueError("power must be in (0, 1]")
    check_tod(tod, _map_channels(mp_stacked), B)
    if stats is not None and not (isinstance(stats, torch.Tensor) and stats.device.type == "cuda"
                                  and stats.dtype == torch.int64 and stats.numel() >= 2
                                  and stats.is_contiguous()):
        raise ValueError("stats must be a contiguous CUDA int64 tensor with at least 2 elements")
    dmap = device_map_for(mp_stacked, c_q, c_u)
    if dmap.nside != int(nside):
        raise ValueError(f"map nside {dmap.nside} != nside {nside}")
    f32 = torch.float32
    ops = [to_dev(a, f32) for a in (vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, beam_vals)]
    p = to_dev(params, f32)
    if p.numel() != N_PARAMS:
        raise ValueError(f"expected {N_PARAMS} corrector parameters")
    host = isinstance(tod, np.ndarray)
    t = torch.from_numpy(np.ascontiguousarray(tod)).to(ops[0].device) if host else tod
    v, ax, ca, sa, bp, cp, sp, bv = ops
    n_corr = S
    if power < 1.0:
        order = torch.argsort(torch.abs(bv.reshape(-1)), descending=True, stable=True)
        v = v.reshape(-1, 3)[order].contiguous()
        bv = bv.reshape(-1)[order].contiguous()
        cum = torch.cumsum(torch.abs(bv).double(), 0)
        n_corr = int(torch.searchsorted(cum, power * float(cum[-1])).item()) + 1
    _lib.check(_lib.lib().todg_gather_accum_learned(
        _lib.ptr(v), _lib.ptr(ax), _lib.ptr(ca), _lib.ptr(sa), _lib.ptr(bp), _lib.ptr(cp), _lib.ptr(sp),
        int(nside), _lib.ptr(dmap.packed), dmap.C, _lib.ptr(bv), B, S, _lib.ptr(t), c_q, c_u,
        _lib.ptr(p), float(bound), min(n_corr, S), _lib.ptr(stats), stream_handle()))
    if host:
        tod[...] = t.cpu().numpy()
    return tod


def bilinear_loss(X, wbil, that) -> float:
    """Mean normalised squared residual of plain bilinear weights."""
    xh = X[:, GEO:]
    p = wbil[:, 0] * xh[:, 5] + wbil[:, 1] * xh[:, 6] + wbil[:, 2] * xh[:, 9] + wbil[:, 3] * xh[:, 10]
    return float(torch.mean((p.double() - that.double()) ** 2))
