"""Production with the learned corrector in its fast (factorised) form:
synthetic nside-2048 I/Q/U sky (spectrum of inputs.cl_spectrum, lmax 3
nside - 1), 30' circular Gaussian beam at nside_beam 4096 with a 0.99 cut, 8
detectors x 1,728,000 samples (20 Hz ring scans an hour apart, rolled psi),
against the plain bilinear kernel and the bicubic judge on the same GPU.

Checks: the coefficients fitted here on 8 nside-512 skies (and the ones the
package ships) bring the RMS TOD residual against the judge to <= 0.20 of
bilinear's on every field, on a held-out nside-512 sky and on the nside-2048
production sky, and so do the channels of the shipped MLP stencil corrector
(the network evaluated per map pixel on BF16 MFMA); the corrected kernel
takes <= 1.5x the plain kernel's time (and the MLP's channel build for the
sky, counted once against the 8 detectors, adds a few per cent);
and the corrected TOD is steadier than bilinear's under rolls of the
circular beam and under rotations of the sky about the polar axis (both
leave a perfect interpolator's TOD unchanged).
"""

import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSIDE = 2048
N_DET = 8
B = 1_728_000


@pytest.fixture(scope="module")
def torch():
    import torch

    return torch


@pytest.fixture(scope="module")
def fitted():
    from tod_generation_mapbased_beam_amd import fcorr as F

    corr, rep = F.FactorisedCorrector.fit(nside=512, n_maps=8, samples_per_map=40_000,
                                          log=lambda m: print(m, flush=True))
    print(f"fit: training RMS ratio {rep.train_ratio:.4f} per map {np.round(rep.per_map_ratio, 4).tolist()} "
          f"({rep.seconds:.1f}s)")
    return corr, rep


@pytest.fixture(scope="module")
def sky(torch):
    from tod_generation_mapbased_beam_amd import inputs

    t0 = time.time()
    mp = torch.as_tensor(inputs.synfast_iqu(NSIDE, seed=7, device="cuda"), device="cuda",
                         dtype=torch.float32).contiguous()
    print(f"sky nside {NSIDE} ({time.time() - t0:.1f}s)", flush=True)
    return mp


def _beam(torch, nside_beam, fwhm):
    from tod_generation_mapbased_beam_amd import inputs

    vec, bv, _ = inputs.gaussian_beam(nside_beam, fwhm, 0.99)
    return (torch.as_tensor(vec, device="cuda"), torch.as_tensor(bv, device="cuda"), int(bv.shape[0]))


def _tods(torch, nside, mp, maps, v, b, S, params, n, variant=0):
    """bilinear, corrected and bicubic TODs (C, n) of one pointing set."""
    from tod_generation_mapbased_beam_amd import fcorr as F, kernels as K

    C = mp.C if hasattr(mp, "C") else int(mp.shape[0])
    qu = (1, 2) if C == 3 else (-1, -1)
    out = [torch.zeros((C, n), dtype=torch.float64, device="cuda") for _ in range(3)]
    K._gather_accum_fused_jit(v, *params, nside, mp, b, n, S, out[0], *qu, variant=variant)
    F.gather_accum_fcorr(v, *params, nside, maps, b, n, S, out[1])
    K._gather_accum_bicubic_jit(v, *params, nside, mp, b, n, S, out[2], *qu)
    return out


def test_fit_and_held_out_sky(torch, fitted):
    from tod_generation_mapbased_beam_amd import fcorr as F, inputs

    corr, rep = fitted
    assert rep.train_ratio <= 0.20
    nside = 512
    v, b, S = _beam(torch, 2 * nside, 120.0)
    params = [torch.as_tensor(p, device="cuda") for p in F.random_pointings(100_000, 4242)]
    mp = torch.as_tensor(inputs.synfast_iqu(nside, seed=4243, device="cuda")[:1], device="cuda",
                         dtype=torch.float32).contiguous()
    for name, c in (("fitted here", corr), ("shipped", F.FactorisedCorrector.load_default())):
        tb, tf, tc = _tods(torch, nside, mp, c.maps(mp), v, b, S, params, 100_000)
        j = F.judge(tf, tb, tc)
        print(f"held-out nside-512 sky, coefficients {name}: RMS ratio {j['rms_ratio']}, max {j['max_ratio']}")
        assert j["rms_ratio"][0] <= 0.20


def test_production_residual_and_time(torch, fitted, sky):
    """8 detectors x 1.728M samples: RMS ratio per field and kernel times."""
    from tod_generation_mapbased_beam_amd import fcorr as F, inputs, kernels as K
    from tod_generation_mapbased_beam_amd import tod_rotations as R
    from tod_generation_mapbased_beam_amd.device import DeviceMap

    from tod_generation_mapbased_beam_amd import smlp

    corr, _ = fitted
    v, b, S = _beam(torch, 2 * NSIDE, 30.0)
    dmap = DeviceMap(sky, 1, 2)
    maps = corr.maps(sky, 1, 2)
    net = smlp.StencilMLPCorrector.load_default()
    net.maps(sky, 1, 2)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    nmaps = net.maps(sky, 1, 2)
    e1.record()
    torch.cuda.synchronize()
    build_ms = e0.elapsed_time(e1)
    ss = {k: torch.zeros(3, dtype=torch.float64, device="cuda") for k in ("bil", "fc", "ml")}
    mx = {k: torch.zeros(3, dtype=torch.float64, device="cuda") for k in ("bil", "fc", "ml")}
    t_ms = {"bil": 0.0, "fc": 0.0, "ml": 0.0}
    for d in range(N_DET):
        th, ph, ps = inputs.ring_scan(B, 20.0, t0=3600.0 * d)
        params = [torch.as_tensor(p, device="cuda")
                  for p in R.rotation_params_for_pointing(th, ph, ps + 0.37 * d)]
        tb, tf, tc = _tods(torch, NSIDE, dmap, maps, v, b, S, params, B)
        tm = torch.zeros_like(tf)
        F.gather_accum_fcorr(v, *params, NSIDE, nmaps, b, B, S, tm)
        for k, t in (("bil", tb), ("fc", tf), ("ml", tm)):
            dd = t - tc
            ss[k] += torch.sum(dd * dd, dim=1)
            mx[k] = torch.maximum(mx[k], torch.max(torch.abs(dd), dim=1).values)
        # kernel times: best of two launches each
        for k in ("bil", "fc", "ml"):
            best = math.inf
            for _ in range(2):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                if k == "bil":
                    K._gather_accum_fused_jit(v, *params, NSIDE, dmap, b, B, S, tb, 1, 2, variant=0)
                else:
                    F.gather_accum_fcorr(v, *params, NSIDE, maps if k == "fc" else nmaps, b, B, S, tf)
                e1.record()
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1))
            t_ms[k] += best
        print(f"detector {d}: bilinear {t_ms['bil']:.1f} ms, corrected {t_ms['fc']:.1f} ms, MLP stencil "
              f"{t_ms['ml']:.1f} ms (cumulative)", flush=True)
    rms = (torch.sqrt(ss["fc"] / ss["bil"])).tolist()
    mxr = (mx["fc"] / mx["bil"]).tolist()
    ratio = t_ms["fc"] / t_ms["bil"]
    print(f"8 x 1.728M: RMS residual ratio (rows I, Q, U) {np.round(rms, 4).tolist()}, max ratio "
          f"{np.round(mxr, 4).tolist()}; time {t_ms['fc']:.1f} / {t_ms['bil']:.1f} ms = {ratio:.3f}x")
    rms_ml = (torch.sqrt(ss["ml"] / ss["bil"])).tolist()
    mxr_ml = (mx["ml"] / mx["bil"]).tolist()
    ratio_ml = t_ms["ml"] / t_ms["bil"]
    ratio_build = (t_ms["ml"] + build_ms) / t_ms["bil"]
    print(f"MLP stencil corrector: RMS residual ratio {np.round(rms_ml, 4).tolist()}, max ratio "
          f"{np.round(mxr_ml, 4).tolist()}; time {ratio_ml:.3f}x, with the sky's channel build "
          f"({build_ms:.1f} ms) {ratio_build:.3f}x")
    assert max(rms) <= 0.20
    assert ratio <= 1.5
    assert max(rms_ml) <= 0.20
    # the kernel is the target's measure; the sky's channel build, once for
    # the 8 detectors, is reported beside it (measured 1.47x with it)
    assert ratio_ml <= 1.5 and ratio_build <= 1.6


def test_roll_stability(torch, fitted, sky):
    """64 pointings x 32 rolls of the circular beam: RMS spread of the T TOD
    over the rolls (zero for a perfect interpolator)."""
    from tod_generation_mapbased_beam_amd import tod_rotations as R
    from tod_generation_mapbased_beam_amd.device import DeviceMap

    corr, _ = fitted
    v, b, S = _beam(torch, 2 * NSIDE, 30.0)
    rng = np.random.default_rng(5)
    n_pt, n_rot = 64, 32
    th = np.repeat(np.arccos(rng.uniform(-0.7, 0.7, n_pt)), n_rot)
    ph = np.repeat(rng.uniform(0, 2 * math.pi, n_pt), n_rot)
    ps = rng.uniform(0, 2 * math.pi, n_pt * n_rot)
    params = [torch.as_tensor(p, device="cuda") for p in R.rotation_params_for_pointing(th, ph, ps)]
    tods = _tods(torch, NSIDE, DeviceMap(sky, 1, 2), corr.maps(sky, 1, 2), v, b, S, params, n_pt * n_rot)
    spread = [float(torch.sqrt(torch.mean(torch.var(t[0].reshape(n_pt, n_rot), dim=1, unbiased=False))))
              for t in tods]
    print(f"roll spread: bilinear {spread[0]:.3e}, corrected {spread[1]:.3e}, bicubic {spread[2]:.3e}")
    assert spread[1] < spread[0]


def test_polar_rotation_stability(torch, fitted):
    """The T sky rotated about the polar axis (a_lm -> a_lm exp(-i m alpha))
    and scanned with phi -> phi + alpha, at nside 512 (the beam scaled to
    120'): RMS deviation of the TOD from its mean over 4 angles."""
    from tod_generation_mapbased_beam_amd import inputs
    from tod_generation_mapbased_beam_amd import tod_rotations as R

    corr, _ = fitted
    nside, lmax, n = 512, 3 * 512 - 1, 100_000
    v, b, S = _beam(torch, 2 * nside, 120.0)
    tt, _, _ = inputs.cl_spectrum(lmax)
    alm = inputs._alm_realisation(tt, lmax, np.random.default_rng(11))
    th, ph, ps = inputs.ring_scan(n, 20.0 * nside / 2048)
    mm = np.arange(lmax + 1)
    res = [[], [], []]
    for a in [0.0, 0.31, 1.7, 4.4]:
        mp = inputs.synthesize_map(nside, [alm * np.exp(-1j * mm * a)[None, :]], lmax, device="cuda")
        mp = torch.as_tensor(mp, device="cuda", dtype=torch.float32).contiguous()
        params = [torch.as_tensor(p, device="cuda")
                  for p in R.rotation_params_for_pointing(th, np.mod(ph + a, 2 * math.pi), ps)]
        for k, t in enumerate(_tods(torch, nside, mp, corr.maps(mp), v, b, S, params, n)):
            res[k].append(t[0])
    spread = []
    for ts in res:
        T = torch.stack(ts)
        spread.append(float(torch.sqrt(torch.mean((T - T.mean(dim=0, keepdim=True)) ** 2))))
    print(f"polar rotation spread: bilinear {spread[0]:.3e}, corrected {spread[1]:.3e}, bicubic {spread[2]:.3e}")
    assert spread[1] < spread[0]


def _rod(k, c, s):
    """Rodrigues rotation matrices about unit axes k (n, 3) by angles with
    cosines c and sines s (n,): v cos + (k x v) sin + k (k . v)(1 - cos), as
    the kernels apply them (healpix_ring.hpp, rodrigues)."""
    n = k.shape[0]
    K = np.zeros((n, 3, 3))
    K[:, 0, 1], K[:, 0, 2] = -k[:, 2], k[:, 1]
    K[:, 1, 0], K[:, 1, 2] = k[:, 2], -k[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -k[:, 1], k[:, 0]
    kk = k[:, :, None] * k[:, None, :]
    return (c[:, None, None] * np.eye(3)[None] + s[:, None, None] * K + (1 - c)[:, None, None] * kk)


def _rotated_operands(params, Rm, centre):
    """Kernel operands of R . R_b for every sample's rotation R_b: boresight
    R b, the minimal rotation taking the beam centre there, and the roll
    about the new boresight that completes R . R_b."""
    axes, ca, sa, bp, cp, sp = (np.asarray(p, np.float64) for p in params)
    Rb = _rod(bp, cp, sp) @ _rod(axes, ca, sa)
    Q = Rm[None] @ Rb
    b2 = Q @ centre
    k = np.cross(np.broadcast_to(centre, b2.shape), b2)
    kn = np.linalg.norm(k, axis=1, keepdims=True)
    k = np.where(kn > 1e-12, k / np.maximum(kn, 1e-300), np.array([1.0, 0.0, 0.0]))
    ang = np.arccos(np.clip(b2 @ centre, -1.0, 1.0))
    A = _rod(k, np.cos(ang), np.sin(ang))
    P = Q @ np.transpose(A, (0, 2, 1))  # a rotation about b2
    anti = 0.5 * np.stack([P[:, 2, 1] - P[:, 1, 2], P[:, 0, 2] - P[:, 2, 0], P[:, 1, 0] - P[:, 0, 1]], 1)
    p = np.arctan2(np.sum(anti * b2, axis=1), 0.5 * (np.trace(P, axis1=1, axis2=2) - 1.0))
    f32 = np.float32
    return (k.astype(f32), np.cos(ang).astype(f32), np.sin(ang).astype(f32), b2.astype(f32),
            np.cos(p).astype(f32), np.sin(p).astype(f32))


def _sky_at(alm, lmax, z, phi, torch, chunk=32768):
    """sum_lm a_lm Y_lm at directions (z, phi): the spin-0 synthesis of
    inputs.synthesize_map evaluated point by point."""
    from tod_generation_mapbased_beam_amd import inputs

    out = np.empty(z.shape[0])
    w = np.full(lmax + 1, 2.0)
    w[0] = 1.0
    mm = np.arange(lmax + 1)
    for i in range(0, z.shape[0], chunk):
        th = np.arccos(np.clip(z[i:i + chunk], -1.0, 1.0))
        F = inputs._ring_sums([alm], lmax, th, "cuda", None)[0]  # (n, M)
        e = torch.as_tensor(np.exp(1j * np.outer(phi[i:i + chunk], mm)) * w[None, :], device="cuda")
        out[i:i + chunk] = torch.real(torch.sum(F * e, dim=1)).cpu().numpy()
    return out


def test_general_rotation_stability(torch, fitted):
    """The T sky under 3 random rotations R of SO(3) (sky f o R^-1, evaluated
    exactly from its a_lm at every pixel centre) scanned with the rotated
    pointings R . R_b: a perfect interpolator's TOD does not change.  nside
    128 (lmax 383), the beam scaled to the pixel size (480' at nside_beam
    256), 50,000 pointings over the sphere; RMS deviation of each method's TOD
    from its mean over the identity and the 3 rotations."""
    import oracle as O

    from tod_generation_mapbased_beam_amd import fcorr as F, inputs, healpix_np as H

    corr, _ = fitted
    nside, lmax, n = 128, 383, 50_000
    v, b, S = _beam(torch, 2 * nside, 480.0)
    vec = v.cpu().numpy()
    tt, _, _ = inputs.cl_spectrum(lmax)
    alm = inputs._alm_realisation(tt, lmax, np.random.default_rng(21))
    pts = F.random_pointings(n, 77)
    centre = np.array([0.0, 0.0, 1.0])  # beam built around the north pole
    pix = np.arange(H.npix(nside))
    zp, php = H.pix2zphi(nside, pix)
    stp = np.sqrt(np.maximum(0.0, 1 - zp * zp))
    npv = np.stack([stp * np.cos(php), stp * np.sin(php), zp], 1)
    rng = np.random.default_rng(8)
    res = [[], [], []]
    for r in range(4):
        if r == 0:
            Rm = np.eye(3)
        else:
            qm, rr = np.linalg.qr(rng.standard_normal((3, 3)))
            Rm = qm * np.sign(np.diag(rr))[None, :]
            if np.linalg.det(Rm) < 0:
                Rm[:, 0] = -Rm[:, 0]
        q = npv @ Rm  # rows: R^-1 n_p (R orthogonal)
        mp = _sky_at(alm, lmax, q[:, 2], np.mod(np.arctan2(q[:, 1], q[:, 0]), 2 * math.pi), torch)
        if r == 0:  # the point synthesis is the ring synthesis at the pixel centres
            ref = inputs.synthesize_map(nside, [alm], lmax, device="cuda")[0]
            assert np.max(np.abs(mp - ref)) < 1e-10 * np.std(ref)
        params = pts if r == 0 else _rotated_operands(pts, Rm, centre)
        if r:
            # the composed operands rotate every beam vector as R . R_b
            d0 = O.rotated_dirs(vec, *pts, 0, 64)
            d1 = O.rotated_dirs(vec, *params, 0, 64)
            assert np.max(np.abs(d1 - d0 @ Rm.T)) < 2e-6
        m = torch.as_tensor(mp[None], device="cuda", dtype=torch.float32).contiguous()
        dp = [torch.as_tensor(p, device="cuda") for p in params]
        for k, t in enumerate(_tods(torch, nside, m, corr.maps(m), v, b, S, dp, n)):
            res[k].append(t[0])
    spread = []
    for ts in res:
        T = torch.stack(ts)
        spread.append(float(torch.sqrt(torch.mean((T - T.mean(dim=0, keepdim=True)) ** 2))))
    sig = float(torch.sqrt(torch.mean(res[2][0] ** 2)))
    print(f"general rotation spread / T RMS: bilinear {spread[0] / sig:.3e}, corrected {spread[1] / sig:.3e}, "
          f"bicubic {spread[2] / sig:.3e}")
    assert spread[1] < spread[0]
