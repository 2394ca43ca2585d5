"""Train and evaluate the learned interpolation corrector on one GPU.

Training (``train``): pairs of (bilinear features, bicubic target) at
uniformly random directions on seeded synthetic nside-512 skies (spectrum
C_l^TT = 1e3/(l(l+1)+1), lmax = 3 nside - 1), a held-out set on a new seed;
BF16-MFMA training kernels with FP32 master weights and Adam.

Evaluation (``evaluate``): beam TODs of several detectors on a synthetic
nside-2048 sky with the plain bilinear kernel, the fused corrector and the
bicubic judge; RMS and max residuals against the judge, kernel times, and
two rotation-stability spreads per method, both zero for a perfect
interpolator: (1) the same sky under global rotations about the polar axis
by random angles alpha (a_lm -> a_lm exp(-i m alpha), pointings phi -> phi +
alpha), which moves every pointing's sub-pixel position; (2) the TOD of one
pointing under random rolls of a circular beam (the metric of the
reference's docs/beam_interpolation_accuracy.rst:70-76).

    python -m tod_generation_mapbased_beam_amd.learned_eval --out results.json [--quick]
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import corrector as C
from . import inputs, kernels
from . import tod_rotations as R
from .device import DeviceMap, device


def log(msg):
    print(f"[learned] {msg}", file=sys.stderr, flush=True)


def train(nside=512, n_maps=8, n_pairs=10_000_000, n_held=1_000_000, epochs=8, batch=4096,
          lr=1e-3, seed=0, schedule="cosine"):
    t0 = time.time()
    X, W, T = C.make_pairs(nside, seeds=range(1, n_maps + 1), n_per_map=n_pairs // n_maps,
                           dir_seed=seed)
    Xh, Wh, Th = C.make_pairs(nside, seeds=[1000 + seed], n_per_map=n_held, dir_seed=seed + 1)
    t_data = time.time() - t0
    log(f"pairs: train {X.shape[0]}, held-out {Xh.shape[0]} ({t_data:.1f}s)")
    c = C.Corrector(seed=seed)
    torch.cuda.synchronize()
    t1 = time.time()
    hist = c.train(X, W, T, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log, schedule=schedule)
    torch.cuda.synchronize()
    t_train = time.time() - t1
    bil = C.bilinear_loss(Xh, Wh, Th)
    got = c.eval_loss(Xh, Wh, Th)
    res = {
        "nside": nside, "maps": n_maps, "train_pairs": int(X.shape[0]), "held_out_pairs": int(Xh.shape[0]),
        "epochs": epochs, "batch": batch, "lr": lr, "schedule": schedule, "train_seconds": t_train,
        "data_seconds": t_data,
        "loss_per_epoch": hist,
        "held_out_ms_residual_bilinear": bil, "held_out_ms_residual_corrector": got,
        "held_out_rms_ratio": math.sqrt(got / bil),
    }
    log(f"held-out RMS residual vs bicubic: corrector / bilinear = {res['held_out_rms_ratio']:.4f}")
    return c, res


def _sky(nside, seed, lmax):
    tt, ee, bb = inputs.cl_spectrum(lmax)
    rng = np.random.default_rng(seed)
    alms = [inputs._alm_realisation(cl, lmax, rng) for cl in (tt, ee, bb)]
    return inputs.synthesize_map(nside, alms, lmax, device=device()).astype(np.float32)


def _time(fn, reps=2):
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return min(ts)


def evaluate(c: C.Corrector, nside=2048, n_det=8, n_samples=1_728_000, fwhm=30.0, bound=4.0,
             lmax=None, seed=7, n_rot=32, n_rot_points=64, powers=(0.99, 0.95, 0.9)):
    dev = device()
    lmax = 3 * nside - 1 if lmax is None else lmax
    t0 = time.time()
    mp = _sky(nside, seed, lmax)
    log(f"sky nside {nside} lmax {lmax} ({time.time() - t0:.1f}s)")
    dmap = DeviceMap(mp, 1, 2)
    vec, bv, _ = inputs.gaussian_beam(2 * nside, fwhm, 0.99)
    S = vec.shape[0]
    out = {"nside": nside, "detectors": n_det, "samples_per_detector": n_samples, "beam_pixels": S,
           "bound": bound}
    res = {"bilinear": [], "learned": [], "bicubic": []}
    times = {"bilinear": 0.0, "learned": 0.0, "bicubic": 0.0}
    pres = {pw: [] for pw in powers}
    ptimes = {pw: 0.0 for pw in powers}
    stats = torch.zeros(2, dtype=torch.int64, device=dev)
    for d in range(n_det):
        th, ph, ps = inputs.ring_scan(n_samples, 20.0, t0=3600.0 * d)
        ps = ps + 0.37 * d
        params = [torch.as_tensor(p, device=dev) for p in R.rotation_params_for_pointing(th, ph, ps)]
        v = torch.as_tensor(vec, device=dev)
        b = torch.as_tensor(bv, device=dev)
        tods = {k: torch.zeros((3, n_samples), dtype=torch.float64, device=dev) for k in res}

        def run(kind):
            t = tods[kind]
            t.zero_()
            if kind == "bilinear":
                kernels._gather_accum_fused_jit(v, *params, nside, dmap, b, n_samples, S, t, 1, 2)
            elif kind == "bicubic":
                kernels._gather_accum_bicubic_jit(v, *params, nside, dmap, b, n_samples, S, t, 1, 2)
            else:
                C.gather_accum_learned(v, *params, nside, dmap, b, n_samples, S, t, 1, 2,
                                       params=c.params, bound=bound)

        for k in res:
            times[k] += _time(lambda: run(k), reps=1)
        for pw in powers:  # the corrector on the strongest beam pixels only
            tp = torch.zeros((3, n_samples), dtype=torch.float64, device=dev)

            def run_p():
                tp.zero_()
                C.gather_accum_learned(v, *params, nside, dmap, b, n_samples, S, tp, 1, 2,
                                       params=c.params, bound=bound, power=pw)

            ptimes[pw] += _time(run_p, reps=1)
            pres[pw].append(tp)
        C.gather_accum_learned(v, *params, nside, dmap, b, n_samples, S,
                               torch.zeros((3, n_samples), dtype=torch.float64, device=dev), 1, 2,
                               params=c.params, bound=bound, stats=stats)
        for k in res:
            res[k].append(tods[k])
        log(f"detector {d}: bilinear {times['bilinear']:.3f}s learned {times['learned']:.3f}s "
            f"bicubic {times['bicubic']:.3f}s (cumulative)")
    judge = torch.cat(res["bicubic"], dim=1)
    for k in ("bilinear", "learned"):
        t = torch.cat(res[k], dim=1)
        diff = t - judge
        rms_sig = torch.sqrt(torch.mean(judge ** 2, dim=1))
        out[k] = {
            "rms_residual_vs_bicubic": torch.sqrt(torch.mean(diff ** 2, dim=1)).tolist(),
            "max_residual_vs_bicubic": torch.max(torch.abs(diff), dim=1).values.tolist(),
            "signal_rms": rms_sig.tolist(),
            "kernel_seconds": times[k],
        }
    out["bicubic"] = {"kernel_seconds": times["bicubic"]}
    rb0 = torch.sqrt(torch.mean((torch.cat(res["bilinear"], dim=1) - judge) ** 2, dim=1))
    out["learned_power_sweep"] = []
    for pw in powers:
        rp = torch.sqrt(torch.mean((torch.cat(pres[pw], dim=1) - judge) ** 2, dim=1))
        out["learned_power_sweep"].append({
            "power": pw, "kernel_seconds": ptimes[pw],
            "time_ratio_over_bilinear": ptimes[pw] / times["bilinear"],
            "rms_ratio_over_bilinear": (rp / rb0).tolist()})
    rb = np.array(out["bilinear"]["rms_residual_vs_bicubic"])
    rl = np.array(out["learned"]["rms_residual_vs_bicubic"])
    out["rms_ratio_learned_over_bilinear"] = (rl / rb).tolist()
    out["time_ratio_learned_over_bilinear"] = times["learned"] / times["bilinear"]
    st = stats.cpu().numpy()
    out["learned_corrected_fraction"] = float(st[0] / max(1, st[0] + st[1]))
    out["rotation_stability_roll"] = rotation_stability(c, dmap, vec, bv, nside, n_rot, n_rot_points, bound)
    del dmap, mp
    out["rotation_stability_global"] = global_rotation_stability(c, nside, lmax, vec, bv, bound=bound,
                                                                 seed=seed)
    return out


def rotation_stability(c, dmap, vec, bv, nside, n_rot=32, n_points=64, bound=4.0, seed=5):
    """RMS spread over random rolls psi of the T-TOD at fixed pointings
    (circular beam), averaged over pointings, per method."""
    dev = device()
    rng = np.random.default_rng(seed)
    th = np.repeat(np.arccos(rng.uniform(-0.7, 0.7, n_points)), n_rot)
    ph = np.repeat(rng.uniform(0, 2 * math.pi, n_points), n_rot)
    ps = rng.uniform(0, 2 * math.pi, n_points * n_rot)
    params = [torch.as_tensor(p, device=dev) for p in R.rotation_params_for_pointing(th, ph, ps)]
    B = th.shape[0]
    S = vec.shape[0]
    v = torch.as_tensor(vec, device=dev)
    b = torch.as_tensor(bv, device=dev)
    out = {}
    for k in ("bilinear", "learned", "bicubic"):
        t = torch.zeros((3, B), dtype=torch.float64, device=dev)
        if k == "bilinear":
            kernels._gather_accum_fused_jit(v, *params, nside, dmap, b, B, S, t, 1, 2)
        elif k == "bicubic":
            kernels._gather_accum_bicubic_jit(v, *params, nside, dmap, b, B, S, t, 1, 2)
        else:
            C.gather_accum_learned(v, *params, nside, dmap, b, B, S, t, 1, 2, params=c.params, bound=bound)
        T = t[0].reshape(n_points, n_rot)  # the intensity channel
        out[k] = float(torch.sqrt(torch.mean(torch.var(T, dim=1, unbiased=False))))
    return out


def global_rotation_stability(c, nside, lmax, vec, bv, n_rot=3, n_samples=200_000, bound=4.0,
                              seed=7):
    """The intensity sky rotated about the polar axis by alpha = 0 and n_rot
    random angles (a_lm -> a_lm exp(-i m alpha)), scanned with pointings
    phi -> phi + alpha: every method's T-TOD should not change.  Returns per
    method the RMS over samples of the deviation from the mean over alpha,
    and that RMS relative to the T signal RMS."""
    dev = device()
    tt, _, _ = inputs.cl_spectrum(lmax)
    rng = np.random.default_rng(seed)
    alm = inputs._alm_realisation(tt, lmax, rng)
    alphas = [0.0] + list(np.random.default_rng(seed + 1).uniform(0, 2 * math.pi, n_rot))
    th, ph, ps = inputs.ring_scan(n_samples, 20.0)
    S = vec.shape[0]
    v = torch.as_tensor(vec, device=dev)
    b = torch.as_tensor(bv, device=dev)
    mm = np.arange(lmax + 1)
    tods = {"bilinear": [], "learned": [], "bicubic": []}
    for a in alphas:
        t0 = time.time()
        rot = alm * np.exp(-1j * mm * a)[None, :]
        mp = inputs.synthesize_map(nside, [rot], lmax, device=dev).astype(np.float32)
        dmap = DeviceMap(mp, -1, -1)
        params = [torch.as_tensor(p, device=dev)
                  for p in R.rotation_params_for_pointing(th, np.mod(ph + a, 2 * math.pi), ps)]
        for k in tods:
            t = torch.zeros((1, n_samples), dtype=torch.float64, device=dev)
            if k == "bilinear":
                kernels._gather_accum_fused_jit(v, *params, nside, dmap, b, n_samples, S, t)
            elif k == "bicubic":
                kernels._gather_accum_bicubic_jit(v, *params, nside, dmap, b, n_samples, S, t)
            else:
                C.gather_accum_learned(v, *params, nside, dmap, b, n_samples, S, t, params=c.params,
                                       bound=bound)
            tods[k].append(t[0])
        del dmap, mp
        log(f"global rotation alpha={a:.4f} ({time.time() - t0:.1f}s)")
    out = {"alphas": [float(a) for a in alphas], "samples": n_samples}
    for k, ts in tods.items():
        T = torch.stack(ts)
        dev_rms = float(torch.sqrt(torch.mean((T - T.mean(dim=0, keepdim=True)) ** 2)))
        sig = float(torch.sqrt(torch.mean(T[0] ** 2)))
        out[k] = {"rms_spread": dev_rms, "rms_spread_over_signal": dev_rms / sig}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--quick", action="store_true", help="reduced sizes (minutes -> seconds)")
    ap.add_argument("--params-out", default=None, help="save the trained parameters (.npy)")
    args = ap.parse_args()
    if args.quick:
        c, tr = train(nside=128, n_maps=2, n_pairs=400_000, n_held=100_000, epochs=3)
        ev = evaluate(c, nside=256, n_det=2, n_samples=20_000, fwhm=120.0, n_rot=16, n_rot_points=16)
    else:
        c, tr = train()
        ev = evaluate(c)
    res = {"train": tr, "evaluate": ev}
    if args.params_out:
        np.save(args.params_out, c.params.cpu().numpy())
    txt = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt)
    print(txt)


if __name__ == "__main__":
    main()
