"""GPU parity of the factorised corrector (footprint mode 4 and its stencil
kernel) against the CPU oracle, on identical inputs.

Bar: channels and TOD within 1e-6 of the oracle (max |gpu - oracle| / RMS);
identity coefficients give the plain bilinear kernel's TOD to rounding; the
output repeats bit for bit; non-finite pointings are skipped with the status
bit; footprints beyond the split weight grid (a 60' beam at nside 2048) take
the direct path and still match.
"""

import numpy as np
import pytest

import oracle as O
from cases import rel_err, small_config

pytestmark = pytest.mark.gpu

TOL = 1e-6


@pytest.fixture(scope="module")
def torch():
    import torch

    return torch


def _field_major(maps):
    """The packed channel maps back in (C, npix) field order (host)."""
    cs = maps.cs
    p0 = maps.packed0.view(-1, cs).cpu().numpy()
    p1 = maps.packed1.view(-1, cs).cpu().numpy()
    A0 = np.zeros((maps.C, maps.npix), np.float32)
    A1 = np.zeros_like(A0)
    for k, c in enumerate(maps.perm):
        if c >= 0:
            A0[c] = p0[:, k]
            A1[c] = p1[:, k]
    return A0, A1


def _coef(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return np.concatenate([[1.0], scale * rng.standard_normal(31)])


@pytest.mark.parametrize("nside", [64, 256])
def test_channels_match_oracle(torch, nside):
    from tod_generation_mapbased_beam_amd import fcorr as F

    rng = np.random.default_rng(nside)
    m = rng.standard_normal(12 * nside * nside).astype(np.float32)
    coef = rng.standard_normal(32)
    a0 = torch.empty(m.size, dtype=torch.float32, device="cuda")
    a1 = torch.empty_like(a0)
    F.channels(nside, torch.as_tensor(m, device="cuda"), coef, a0, a1, 1)
    o0, o1 = O.fcorr_channels(nside, m, coef)
    e0 = np.max(np.abs(a0.cpu().numpy() - o0)) / np.sqrt(np.mean(o0 ** 2))
    e1 = np.max(np.abs(a1.cpu().numpy() - o1)) / np.sqrt(np.mean(o1 ** 2))
    print(f"nside {nside}: channel errors {e0:.2e} {e1:.2e}")
    assert e0 <= TOL and e1 <= TOL


@pytest.mark.parametrize("qu", [(1, 2), (-1, -1)])
def test_small_config_tod_matches_oracle(torch, qu):
    from tod_generation_mapbased_beam_amd import fcorr as F

    c = small_config(n_samples=4000)
    mp = c["map"] if qu[0] >= 0 else c["map"][:1]
    maps = F.FcorrMaps(mp, _coef(1), *qu)
    tod = torch.zeros((mp.shape[0], c["B"]), dtype=torch.float64, device="cuda")
    F.gather_accum_fcorr(c["vec"], *c["params"], c["nside"], maps, c["bv"], c["B"], c["S"], tod)
    A0, A1 = _field_major(maps)
    ref = np.zeros((mp.shape[0], c["B"]))
    O.gather_accum_fcorr(c["vec"], *c["params"], c["nside"], A0, A1, c["bv"], c["B"], c["S"], ref, *qu)
    err = rel_err(tod.cpu().numpy(), ref)
    print(f"qu={qu}: max |gpu - oracle| / RMS = {err:.2e}")
    assert err <= TOL


def test_identity_coefficients_give_the_bilinear_tod(torch):
    from tod_generation_mapbased_beam_amd import fcorr as F, kernels as K

    c = small_config(n_samples=4000)
    maps = F.FcorrMaps(c["map"], F.identity_coef(), 1, 2)
    t = torch.zeros((3, c["B"]), dtype=torch.float64, device="cuda")
    F.gather_accum_fcorr(c["vec"], *c["params"], c["nside"], maps, c["bv"], c["B"], c["S"], t)
    p = torch.zeros_like(t)
    K._gather_accum_fused_jit(c["vec"], *c["params"], c["nside"], c["map"], c["bv"], c["B"], c["S"], p, 1, 2)
    assert rel_err(t.cpu().numpy(), p.cpu().numpy()) < 1e-12


def test_repeats_bit_for_bit(torch):
    from tod_generation_mapbased_beam_amd import fcorr as F

    c = small_config(n_samples=4000)
    maps = F.FcorrMaps(c["map"], _coef(2), 1, 2)
    outs = []
    for _ in range(3):
        t = torch.zeros((3, c["B"]), dtype=torch.float64, device="cuda")
        F.gather_accum_fcorr(c["vec"], *c["params"], c["nside"], maps, c["bv"], c["B"], c["S"], t)
        outs.append(t.cpu().numpy())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_non_finite_pointing_is_skipped_with_status(torch):
    from tod_generation_mapbased_beam_amd import fcorr as F

    c = small_config(n_samples=2000)
    params = [np.array(p, copy=True) for p in c["params"]]
    bad = [5, 777]
    params[1][bad[0]] = np.nan
    params[3].reshape(-1)[3 * bad[1] + 2] = np.inf  # ax_pts (B, 3)
    maps = F.FcorrMaps(c["map"], _coef(3), 1, 2)
    t = torch.full((3, c["B"]), 7.0, dtype=torch.float64, device="cuda")
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    F.gather_accum_fcorr(c["vec"], *params, c["nside"], maps, c["bv"], c["B"], c["S"], t, status=st)
    got = t.cpu().numpy()
    assert int(st.item()) & 1
    assert np.all(got[:, bad] == 7.0)
    ok = np.setdiff1d(np.arange(c["B"]), bad)
    A0, A1 = _field_major(maps)
    sub = tuple(np.ascontiguousarray(p[ok[:300]]) for p in params)
    ref = np.full((3, 300), 7.0)
    O.gather_accum_fcorr(c["vec"], *sub, c["nside"], A0, A1, c["bv"], 300, c["S"], ref, 1, 2)
    assert rel_err(got[:, ok[:300]], ref) <= TOL


@pytest.mark.parametrize("fwhm,label", [(30.0, "production beam, laid out"),
                                        (60.0, "footprint beyond the split grid: direct path")])
def test_nside2048_day_samples_match_oracle(torch, fwhm, label):
    """1.728M-sample day scan at nside 2048 (20 Hz), one launch; 48 random
    samples checked against the oracle."""
    from tod_generation_mapbased_beam_amd import fcorr as F, inputs
    from tod_generation_mapbased_beam_amd import tod_rotations as R

    nside, B = 2048, 1_728_000
    vec, bv, _ = inputs.gaussian_beam(4096, fwhm, 0.99)
    S = vec.shape[0]
    th, ph, ps = inputs.ring_scan(B, 20.0)
    params = R.rotation_params_for_pointing(th, ph, ps)
    rng = np.random.default_rng(9)
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    maps = F.FcorrMaps(mp, _coef(4), 1, 2)
    dev = [torch.as_tensor(p, device="cuda") for p in params]
    t = torch.zeros((3, B), dtype=torch.float64, device="cuda")
    F.gather_accum_fcorr(torch.as_tensor(vec, device="cuda"), *dev, nside, maps,
                         torch.as_tensor(bv, device="cuda"), B, S, t)
    idx = np.sort(rng.choice(B, 48, replace=False))
    A0, A1 = _field_major(maps)
    sub = tuple(np.ascontiguousarray(p[idx]) for p in params)
    ref = np.zeros((3, idx.size))
    O.gather_accum_fcorr(vec, *sub, nside, A0, A1, bv, idx.size, S, ref, 1, 2)
    err = rel_err(t.cpu().numpy()[:, idx], ref)
    print(f"{label}: S {S}, max |gpu - oracle| / RMS = {err:.2e}")
    assert err <= TOL
