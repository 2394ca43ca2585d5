"""Factorised learned interpolation corrector: bicubic-judged accuracy at
close to the plain bilinear kernel's cost.

Model.  For a rotated beam pixel with bilinear neighbours p_k, weights w_k
(the reference's, numba_healpy.py:358-499) and lower-ring weight g, the
corrected value is

    sum_k w_k [A0(p_k) + (1 - g) g A1(p_k)],

where A0 and A1 are learned linear stencils of the map (16 features per
pixel: the pixel's value and 15 ring-interpolated second differences out to
3 rings and 3 ring steps, ``todg_fcorr_channels``).  The correction factors
into a map-side part, built once per map, and a geometry part that enters
each beam pixel through one extra weight, so the TOD kernel
(``todg_gather_accum_fcorr``) deposits two weight sets per beam pixel
instead of one and gathers two channels per footprint pixel.  A0 absorbs the
smoothing that bilinear interpolation applies on average; A1, modulated by
each beam pixel's (1 - g) g, absorbs the part that varies with where the beam
pixel falls between two rings, which no map-side filter can see.

Training.  The model is linear in its 31 free coefficients (A0's coefficient
of the pixel itself is fixed at 1, which makes the plain bilinear TOD the
zero-correction start), so the fit is a least-squares problem over whole
TODs: each coefficient's feature is the TOD of the corresponding unit
stencil, computed by the production kernel itself, and the target is the
bicubic judge's TOD minus the bilinear one, on synthetic skies and random
pointings.  The 31 x 31 normal equations are solved in binary64.

The per-beam-pixel MLP corrector (``corrector.py``) reaches a lower residual
at several times the cost; this one is the fast form.
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
from . import inputs
from . import tod_rotations as R
from .device import device, map_slots, nside_of, stream_handle, to_dev
from .kernels import check_status, check_tod, gather_accum

N_FEATURES = 16
# (ring offset, ring-step offset) of features 1..15; feature 0 is the pixel
FEATURE_OFFSETS = [(dr, ds) for dr in range(4) for ds in range(4) if (dr, ds) != (0, 0)]


def identity_coef() -> np.ndarray:
    """Coefficients of the plain bilinear kernel: A0 = map, A1 = 0."""
    c = np.zeros(2 * N_FEATURES)
    c[0] = 1.0
    return c


def channels(nside: int, field, coef, out0, out1, stride: int = 1) -> None:
    """A0 and A1 of one map field (device float32, npix) into out0[p * stride]
    and out1[p * stride] (device float32 views)."""
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    if coef.shape != (2 * N_FEATURES,):
        raise ValueError(f"coef must hold {2 * N_FEATURES} values")
    _lib.check(_lib.lib().todg_fcorr_channels(int(nside), _lib.ptr(field), coef.ctypes.data,
                                               _lib.ptr(out0), _lib.ptr(out1), int(stride),
                                               stream_handle()))


class FcorrMaps:
    """The two packed channel maps of a ``(C, npix)`` sky for a Q/U channel
    assignment, in the slot layout of the plain kernels' packed map.  The
    channels come from the stencil coefficients ``coef``, or from
    ``builder(nside, field, out0, out1, stride)`` when one is given (the MLP
    stencil corrector's)."""

    def __init__(self, mp_stacked, coef, c_q: int = -1, c_u: int = -1, builder=None):
        src = to_dev(mp_stacked, torch.float32)
        if src.dim() == 1:
            src = src[None, :]
        if src.dim() != 2:
            raise ValueError("mp_stacked must be (C, npix)")
        self.C = int(src.shape[0])
        self.npix = int(src.shape[1])
        self.nside = nside_of(self.npix)
        self.c_q = int(c_q)
        self.c_u = int(c_u)
        self.coef = None if coef is None else np.array(coef, dtype=np.float64)
        self.cs, self.perm, self.row = map_slots(self.C, self.c_q, self.c_u)
        self.packed0 = torch.zeros(self.npix * self.cs, dtype=torch.float32, device=src.device)
        self.packed1 = torch.zeros(self.npix * self.cs, dtype=torch.float32, device=src.device)
        for k, c in enumerate(self.perm):
            if c < 0:
                continue
            if builder is not None:
                builder(self.nside, src[c].contiguous(), self.packed0[k:], self.packed1[k:], self.cs)
            else:
                channels(self.nside, src[c].contiguous(), self.coef, self.packed0[k:], self.packed1[k:],
                         self.cs)


def gather_accum_fcorr(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, maps: FcorrMaps,
                       beam_vals, B, S, tod, status=None):
    """Corrected beam TOD (accumulated into the CUDA float64 (C, B) ``tod``):
    the arguments of ``_gather_accum_fused_jit`` with the sky replaced by its
    FcorrMaps."""
    B = int(B)
    S = int(S)
    check_tod(tod, maps.C, B)
    if not isinstance(tod, torch.Tensor):
        raise ValueError("tod must be a CUDA float64 tensor")
    if status is not None:
        check_status(status)
    if maps.nside != int(nside):
        raise ValueError(f"map nside {maps.nside} != nside {nside}")
    f32 = torch.float32
    v, ax, ca, sa, bp, cp, sp, bv = [to_dev(a, f32) for a in (vec_orig, axes, cos_a, sin_a, ax_pts, cos_p,
                                                               sin_p, beam_vals)]
    if v.numel() < 3 * S or bv.numel() < S:
        raise ValueError("vec_orig / beam_vals shorter than S")
    if ax.numel() < 3 * B or bp.numel() < 3 * B or min(ca.numel(), sa.numel(), cp.numel(), sp.numel()) < B:
        raise ValueError("rotation operands shorter than B")
    _lib.check(_lib.lib().todg_gather_accum_fcorr(
        _lib.ptr(v), _lib.ptr(ax), _lib.ptr(ca), _lib.ptr(sa), _lib.ptr(bp), _lib.ptr(cp), _lib.ptr(sp),
        int(nside), _lib.ptr(maps.packed0), _lib.ptr(maps.packed1), maps.C, _lib.ptr(bv), B, S,
        _lib.ptr(tod), maps.c_q, maps.c_u, _lib.ptr(status), stream_handle()))
    return tod


def random_pointings(n: int, seed: int):
    """n pointings uniform on the sphere with uniform position angles, as the
    six float32 rotation operands of the kernels."""
    rng = np.random.default_rng(seed)
    theta = np.arccos(rng.uniform(-1.0, 1.0, n))
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    psi = rng.uniform(0.0, 2.0 * math.pi, n)
    return R.rotation_params_for_pointing(theta, phi, psi)


@dataclass
class FitReport:
    n_maps: int
    samples_per_map: int
    train_ratio: float          # RMS(corrected - bicubic) / RMS(bilinear - bicubic), training TODs
    seconds: float
    per_map_ratio: list = field(default_factory=list)


class FactorisedCorrector:
    """The 32 stencil coefficients (A0: 0..15, A1: 16..31) and their use."""

    def __init__(self, coef=None):
        self.coef = identity_coef() if coef is None else np.array(coef, dtype=np.float64)
        if self.coef.shape != (2 * N_FEATURES,):
            raise ValueError(f"coef must hold {2 * N_FEATURES} values")

    # ------------------------------------------------------------ use
    def maps(self, mp_stacked, c_q: int = -1, c_u: int = -1) -> FcorrMaps:
        return FcorrMaps(mp_stacked, self.coef, c_q, c_u)

    def tod(self, maps: FcorrMaps, vec_orig, beam_vals, params, tod=None):
        B = int(np.asarray(params[1]).shape[0]) if not isinstance(params[1], torch.Tensor) else int(params[1].numel())
        S = int(np.asarray(beam_vals).shape[0]) if not isinstance(beam_vals, torch.Tensor) else int(beam_vals.numel())
        if tod is None:
            tod = torch.zeros((maps.C, B), dtype=torch.float64, device=device())
        return gather_accum_fcorr(vec_orig, *params, maps.nside, maps, beam_vals, B, S, tod)

    # ------------------------------------------------------------ training
    @staticmethod
    def features(nside: int, field, vec_orig, beam_vals, params) -> np.ndarray:
        """(B, 32) TODs of the 32 unit stencils (column j < 16: feature j in
        A0; j >= 16: feature j - 16 in A1) for one map field."""
        B = int(params[1].numel()) if isinstance(params[1], torch.Tensor) else len(params[1])
        S = int(beam_vals.numel()) if isinstance(beam_vals, torch.Tensor) else len(beam_vals)
        fld = to_dev(field, torch.float32)
        npix = int(fld.numel())
        A0 = torch.empty(npix, dtype=torch.float32, device=fld.device)
        A1 = torch.empty(npix, dtype=torch.float32, device=fld.device)
        out = np.empty((B, 2 * N_FEATURES))
        maps = FcorrMaps.__new__(FcorrMaps)
        maps.C, maps.npix, maps.nside, maps.c_q, maps.c_u, maps.cs = 1, npix, int(nside), -1, -1, 1
        for j in range(2 * N_FEATURES):
            e = np.zeros(2 * N_FEATURES)
            e[j] = 1.0
            channels(nside, fld, e, A0, A1, 1)
            maps.packed0, maps.packed1 = A0, A1
            t = torch.zeros((1, B), dtype=torch.float64, device=fld.device)
            gather_accum_fcorr(vec_orig, *params, nside, maps, beam_vals, B, S, t)
            out[:, j] = t[0].cpu().numpy()
        return out

    @classmethod
    def fit(cls, nside: int = 512, n_maps: int = 8, samples_per_map: int = 40_000, fwhm_arcmin: float | None = None,
            power_cut: float = 0.99, seed: int = 100, lmax: int | None = None, ridge: float = 1e-10, log=None):
        """Least-squares fit on n_maps synthetic skies (spectrum of
        inputs.cl_spectrum, seeds seed .. seed + n_maps - 1, I field) at
        ``nside``, ``samples_per_map`` random pointings each, with a Gaussian
        beam of FWHM ``fwhm_arcmin`` (default: 30' scaled to nside 2048's
        pixel size) at nside_beam 2 nside.  Returns (corrector, FitReport)."""
        t0 = time.time()
        dev = device()
        if fwhm_arcmin is None:
            fwhm_arcmin = 30.0 * 2048 / nside
        vec, bv, _ = inputs.gaussian_beam(2 * nside, fwhm_arcmin, power_cut)
        dvec = torch.as_tensor(np.ascontiguousarray(vec, np.float32), device=dev)
        dbv = torch.as_tensor(np.ascontiguousarray(bv, np.float32), device=dev)
        S = int(dbv.numel())
        G = np.zeros((2 * N_FEATURES - 1, 2 * N_FEATURES - 1))
        h = np.zeros(2 * N_FEATURES - 1)
        rr = 0.0
        per_map = []
        keep = []
        for i in range(n_maps):
            params = [torch.as_tensor(p, device=dev) for p in random_pointings(samples_per_map, seed + 1000 + i)]
            m = inputs.synfast_iqu(nside, seed=seed + i, lmax=lmax, device=dev)[0]
            fld = torch.as_tensor(m, device=dev, dtype=torch.float32).contiguous()
            tb = torch.zeros((1, samples_per_map), dtype=torch.float64, device=dev)
            tc = torch.zeros_like(tb)
            gather_accum("bilinear", dvec, *params, nside, fld[None], dbv, samples_per_map, S, tb)
            gather_accum("bicubic", dvec, *params, nside, fld[None], dbv, samples_per_map, S, tc)
            F = cls.features(nside, fld, dvec, dbv, params)
            r = (tc - tb)[0].cpu().numpy()
            X = F[:, 1:]  # feature 0 of A0 is the bilinear TOD itself (coefficient 1)
            G += X.T @ X
            h += X.T @ r
            rr += float(r @ r)
            keep.append((X, r))
            if log:
                log(f"fcorr fit: map {i + 1}/{n_maps} ({time.time() - t0:.1f}s)")
        G_reg = G + ridge * np.trace(G) / G.shape[0] * np.eye(G.shape[0])
        x = np.linalg.solve(G_reg, h)
        coef = np.concatenate([[1.0], x])
        for X, r in keep:
            per_map.append(float(np.std(r - X @ x) / np.std(r)))
        res = sum(float(np.sum((r - X @ x) ** 2)) for X, r in keep)
        ratio = math.sqrt(res / rr)
        return cls(coef), FitReport(n_maps, samples_per_map, ratio, time.time() - t0, per_map)

    # ------------------------------------------------------------ storage
    def to_json(self, **meta) -> str:
        return json.dumps({"kind": "factorised-corrector", "features": N_FEATURES,
                           "coef": [float(v) for v in self.coef], **meta}, indent=1)

    @classmethod
    def from_json(cls, text: str) -> "FactorisedCorrector":
        d = json.loads(text)
        if d.get("kind") != "factorised-corrector" or int(d.get("features", 0)) != N_FEATURES:
            raise ValueError("not a factorised-corrector parameter file")
        return cls(d["coef"])

    @classmethod
    def load_default(cls) -> "FactorisedCorrector":
        """The coefficients shipped with the package (DEFAULT_PATH, written by
        ``python -m tod_generation_mapbased_beam_amd.fcorr --fit``)."""
        with open(DEFAULT_PATH) as f:
            return cls.from_json(f.read())


DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fcorr_coef.json")


def judge(tod, ref_bilinear, ref_bicubic):
    """Per-field RMS and max residual ratios against the bicubic judge:
    |tod - bicubic| over |bilinear - bicubic|."""
    d0 = ref_bilinear - ref_bicubic
    d1 = tod - ref_bicubic
    rms = (torch.sqrt(torch.mean(d1 ** 2, dim=-1)) / torch.sqrt(torch.mean(d0 ** 2, dim=-1))).tolist()
    mx = (torch.max(torch.abs(d1), dim=-1).values / torch.max(torch.abs(d0), dim=-1).values).tolist()
    return {"rms_ratio": rms, "max_ratio": mx}


def main(argv=None):
    import argparse
    import sys

    ap = argparse.ArgumentParser(description="fit the factorised corrector's coefficients on the GPU")
    ap.add_argument("--fit", action="store_true")
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--nside", type=int, default=512)
    ap.add_argument("--maps", type=int, default=8)
    ap.add_argument("--samples", type=int, default=40_000)
    args = ap.parse_args(argv)
    if not args.fit:
        ap.error("nothing to do (use --fit)")
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    corr, rep = FactorisedCorrector.fit(nside=args.nside, n_maps=args.maps, samples_per_map=args.samples, log=log)
    with open(args.out, "w") as f:
        f.write(corr.to_json(trained_on=f"{rep.n_maps} synthetic nside-{args.nside} skies x {rep.samples_per_map} "
                                        f"random pointings, 30' beam scaled to the pixel size",
                             train_rms_ratio=rep.train_ratio) + "\n")
    log(f"training RMS ratio {rep.train_ratio:.4f} ({rep.seconds:.1f}s) -> {args.out}")


if __name__ == "__main__":
    main()


