"""GPU parity: HIP kernels vs the CPU oracle on identical inputs.

Bar: neighbour / nearest pixel indices bit-exact; TOD within 1e-6 of the
oracle, measured as max |gpu - oracle| / RMS(oracle component).
"""

import math

import numpy as np
import pytest

import oracle as O
from cases import random_rot_params, recenter_params, rel_err, small_config

pytestmark = pytest.mark.gpu

TOL = 1e-6


@pytest.fixture(scope="module")
def K():
    import tod_generation_mapbased_beam_amd.kernels as K

    return K


def _gpu_tod(K, mode, vec, params, nside, mp, bv, B, S, c_q, c_u, variant=0):
    tod = np.zeros((mp.shape[0], B))
    fn = K._gather_accum_fused_jit if mode == "bilinear" else K._gather_accum_nearest_jit
    fn(vec, *params, nside, mp, bv, B, S, tod, c_q, c_u, variant=variant)
    return tod


def _cpu_tod(mode, vec, params, nside, mp, bv, B, S, c_q, c_u):
    tod = np.zeros((mp.shape[0], B))
    fn = O.gather_accum_fused if mode == "bilinear" else O.gather_accum_nearest
    fn(vec, *params, nside, mp, bv, B, S, tod, c_q, c_u)
    return tod


@pytest.mark.parametrize("variant", [0, 8, 2])
@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
@pytest.mark.parametrize("qu", [(1, 2), (-1, -1)])
def test_small_config_tod(K, mode, qu, variant):
    c = small_config()
    args = (c["vec"], c["params"], c["nside"], c["map"], c["bv"], c["B"], c["S"], *qu)
    g = _gpu_tod(K, mode, *args, variant=variant)
    r = _cpu_tod(mode, *args)
    err = rel_err(g, r)
    print(f"{mode} qu={qu} rel err {err:.3e}")
    assert err <= TOL


def test_small_config_bilinear_indices_bit_exact(K):
    c = small_config()
    nb = 2000
    pg, wg = K.bilinear_indices(c["vec"], *c["params"], c["nside"], c["B"], c["S"], 0, nb)
    pc, wc = O.fused_indices(c["vec"], *c["params"], c["nside"], 0, nb)
    assert np.array_equal(pg, pc), f"{np.count_nonzero(pg != pc)} index mismatches"
    assert np.max(np.abs(wg - wc)) < 1e-10  # weights are a tolerance path


def test_small_config_nearest_indices_bit_exact(K):
    c = small_config()
    nb = 2000
    pg = K.nearest_indices(c["vec"], *c["params"], c["nside"], c["B"], c["S"], 0, nb)
    pc = O.nearest_indices(c["vec"], *c["params"], c["nside"], 0, nb)
    assert np.array_equal(pg, pc), f"{np.count_nonzero(pg != pc)} index mismatches"


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64, 256])
def test_whole_sphere_indices_bit_exact(K, nside):
    """Random rotations covering the whole sphere, both poles included."""
    rng = np.random.default_rng(100 + nside)
    B, S = 64, 512
    vec, params = random_rot_params(B, S, rng, pole_safe=False)
    pg, wg = K.bilinear_indices(vec, *params, nside, B, S, 0, B)
    pc, wc = O.fused_indices(vec, *params, nside, 0, B)
    assert np.array_equal(pg, pc)
    assert np.max(np.abs(wg - wc)) < 1e-10  # weights are a tolerance path
    ng = K.nearest_indices(vec, *params, nside, B, S, 0, B)
    nc = O.nearest_indices(vec, *params, nside, 0, B)
    assert np.array_equal(ng, nc)


@pytest.mark.parametrize("variant", [0, 8, 1])
@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
@pytest.mark.parametrize("nside", [4, 16, 64])
@pytest.mark.parametrize("C,qu", [(3, (1, 2)), (3, (-1, -1)), (2, (0, 1)), (1, (-1, -1)), (2, (0, 0))])
def test_random_rotations_tod(K, mode, nside, C, qu, variant):
    rng = np.random.default_rng(7 * nside + C)
    B, S = 40, 300
    vec, params = random_rot_params(B, S, rng, pole_safe=False)
    mp = rng.uniform(-1.0, 1.0, (C, 12 * nside * nside)).astype(np.float32)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    g = _gpu_tod(K, mode, vec, params, nside, mp, bv, B, S, *qu, variant=variant)
    r = _cpu_tod(mode, vec, params, nside, mp, bv, B, S, *qu)
    assert rel_err(g, r) <= TOL


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_direct_tile_shapes_identical(K, tile):
    """The direct kernel's per-lane sums do not depend on its sample tile."""
    c = small_config()
    B = 3001
    params = tuple(p[:B] for p in c["params"])
    ref = _gpu_tod(K, "bilinear", c["vec"], params, c["nside"], c["map"], c["bv"], B, c["S"], 1, 2, variant=1)
    got = _gpu_tod(K, "bilinear", c["vec"], params, c["nside"], c["map"], c["bv"], B, c["S"], 1, 2, variant=tile)
    assert np.array_equal(ref, got)


@pytest.mark.parametrize("variant", [0, 8, 16, 24])
@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_footprint_deterministic(K, mode, variant):
    """Same inputs, same footprint-kernel result bit for bit: weight-grid
    deposits are rounded onto a grid on which their sums are exact, so the
    order of the LDS atomics does not matter (the reference's result is
    thread-count independent, tests/test_tod_bilinear.py:830-870)."""
    c = small_config()
    B = 3001
    params = tuple(p[:B] for p in c["params"])
    a = _gpu_tod(K, mode, c["vec"], params, c["nside"], c["map"], c["bv"], B, c["S"], 1, 2, variant)
    b = _gpu_tod(K, mode, c["vec"], params, c["nside"], c["map"], c["bv"], B, c["S"], 1, 2, variant)
    assert np.array_equal(a, b)
    r = _cpu_tod(mode, c["vec"], params, c["nside"], c["map"], c["bv"], B, c["S"], 1, 2)
    assert rel_err(a, r) <= TOL


@pytest.mark.parametrize("variant", list(range(16, 25)))
@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
@pytest.mark.parametrize("qu", [(1, 2), (-1, -1)])
def test_tile_variants_tod(K, mode, variant, qu):
    """Every tile variant (claims of 128/256/512 beam pixels x 3/4/5 sweep
    chunks) against the oracle, for both channel layouts they are built for
    and for one they fall back from (C = 3 without Q/U)."""
    c = small_config()
    B = 1500
    params = tuple(p[:B] for p in c["params"])
    mp = c["map"] if qu != (-1, -1) else c["map"][:1]
    for m in (mp, c["map"]) if qu == (-1, -1) else (mp,):
        g = _gpu_tod(K, mode, c["vec"], params, c["nside"], m, c["bv"], B, c["S"], *qu, variant)
        r = _cpu_tod(mode, c["vec"], params, c["nside"], m, c["bv"], B, c["S"], *qu)
        assert rel_err(g, r) <= TOL


def test_unknown_variant_rejected(K):
    c = small_config()
    params = tuple(p[:10] for p in c["params"])
    with pytest.raises(ValueError):
        _gpu_tod(K, "bilinear", c["vec"], params, c["nside"], c["map"], c["bv"], 10, c["S"], 1, 2, 3)


@pytest.mark.parametrize("mode", ["bilinear", "nearest", "bicubic"])
@pytest.mark.parametrize("variant", [0, 1])
def test_non_finite_pointing_skipped_and_flagged(K, mode, variant):
    """A sample with a non-finite rotation operand is skipped (its tod
    entries keep their values), the others are computed as without it, and
    todg_gather_accum's status word says so."""
    import torch

    if mode == "bicubic" and variant == 1:
        pytest.skip("the bicubic judge runs on the footprint kernel only")
    c = small_config()
    B = 600
    params = [p[:B].copy() for p in c["params"]]
    bad = [0, 1, 77, 300, 599]
    params[0][bad[0], 1] = np.nan
    params[1][bad[1]] = np.inf
    params[3][bad[2], 2] = np.nan
    params[4][bad[3]] = -np.inf
    params[5][bad[4]] = np.nan
    good = np.setdiff1d(np.arange(B), bad)
    tod = np.full((3, B), 5.0)
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    K.gather_accum(mode, c["vec"], *params, c["nside"], c["map"], c["bv"], B, c["S"], tod, 1, 2,
                   variant=variant, status=status)
    assert int(status.item()) & 1
    assert np.all(tod[:, bad] == 5.0)
    ref = np.zeros((3, good.size))
    sub = [p[good] for p in params]
    fn = {"bilinear": O.gather_accum_fused, "nearest": O.gather_accum_nearest,
          "bicubic": O.gather_accum_bicubic}[mode]
    fn(c["vec"], *sub, c["nside"], c["map"], c["bv"], good.size, c["S"], ref, 1, 2)
    assert rel_err(tod[:, good] - 5.0, ref) <= TOL
    status.zero_()
    ok = [p[good] for p in params]
    t2 = np.zeros((3, good.size))
    K.gather_accum(mode, c["vec"], *ok, c["nside"], c["map"], c["bv"], good.size, c["S"], t2, 1, 2,
                   variant=variant, status=status)
    assert int(status.item()) == 0


def test_fallback_without_mp_stacked():
    """beam_tod_batch without mp_stacked (the reference's healpy split
    pipeline, tod_core.py:169-180: batch Rodrigues with float32 rotated
    vectors, vec2ang, get_interp_weights, no spin-2) against the oracle's
    restatement of that pipeline, to the reference's own tolerance for the
    split-vs-fused rounding gap (1e-5 absolute,
    tests/test_tod_bilinear.py:874-914).  vec2ang is healpy's (absent here):
    theta = arccos(z / |v|), phi = atan2(y, x) wrapped to [0, 2 pi)."""
    from tod_generation_mapbased_beam_amd import beam_tod_batch
    from tod_generation_mapbased_beam_amd import tod_rotations as R

    rng = np.random.default_rng(45)
    for nside in (4, 16, 64):
        B, S, C = 8, 20, 3
        vec, _ = random_rot_params(B, S, rng)
        th = rng.uniform(0.2, np.pi - 0.2, B).astype(np.float32)
        ph = rng.uniform(0.0, 2 * np.pi, B).astype(np.float32)
        ps = rng.uniform(0.0, 2 * np.pi, B).astype(np.float32)
        rot, beta = R.precompute_rotation_vector_batch(np.zeros((1, 1)), np.full((1, 1), np.pi / 2),
                                                       ph, th, center_idx=(0, 0))
        psis = -beta + ps
        mp = rng.uniform(0.5, 1.5, (C, 12 * nside * nside)).astype(np.float32)
        bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
        bv /= bv.sum()
        data = dict(vec_orig=vec, beam_vals=bv, comp_indices=[0, 1, 2])
        out = beam_tod_batch(nside, mp, data, rot, ph, th, psis)
        got = np.stack([out[k] for k in (0, 1, 2)]).astype(np.float64)
        params = R._rotation_params(rot, ph, th, psis)
        vr = O.rodrigues_batch(vec, *params).reshape(-1, 3).astype(np.float64)
        r = np.linalg.norm(vr, axis=1)
        theta = np.arccos(vr[:, 2] / r)
        phi = np.arctan2(vr[:, 1], vr[:, 0])
        phi[phi < 0] += 2 * np.pi
        pix, wgt = O.get_interp_weights(nside, theta, phi)
        ref = np.zeros((C, B))
        O.gather_accum(pix, wgt, bv, mp, B, S, ref)
        np.testing.assert_allclose(got, ref, atol=1e-5)


def test_accumulates_in_place(K):
    c = small_config()
    B = 500
    params = tuple(p[:B] for p in c["params"])
    tod = np.zeros((3, B))
    K._gather_accum_fused_jit(c["vec"], *params, c["nside"], c["map"], c["bv"], B, c["S"], tod, 1, 2)
    first = tod.copy()
    K._gather_accum_fused_jit(c["vec"], *params, c["nside"], c["map"], c["bv"], B, c["S"], tod, 1, 2)
    np.testing.assert_allclose(tod, 2 * first, rtol=0, atol=1e-10)


@pytest.mark.parametrize("variant", [0, 8, 2])
@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_empty_and_single_inputs(K, mode, variant):
    """No samples or no beam pixels leave tod as it is (the reference's
    prange over an empty range, tod_bilinear.py:233); one sample, one beam
    pixel, or both, give the oracle's TOD.  With one sample a component's
    RMS over samples is that one value, which the spin-2 mixing of Q and U
    can make arbitrarily small, so those cases are normalised by the RMS of
    the map field instead (the scale of a normalised beam's TOD)."""
    c = small_config()
    nside, mp, vec, bv, S = c["nside"], c["map"], c["vec"], c["bv"], c["S"]
    fn = K._gather_accum_fused_jit if mode == "bilinear" else K._gather_accum_nearest_jit
    p0 = tuple(p[:0] for p in c["params"])
    tod = np.zeros((3, 0))
    fn(vec, *p0, nside, mp, bv, 0, S, tod, 1, 2, variant=variant)
    B = 64
    pB = tuple(p[:B] for p in c["params"])
    tod = np.full((3, B), 0.25)
    fn(vec[:0], *pB, nside, mp, bv[:0], B, 0, tod, 1, 2, variant=variant)
    assert np.array_equal(tod, np.full((3, B), 0.25))
    for b, s in ((1, S), (B, 1), (1, 1)):
        args = (vec[:s], tuple(p[:b] for p in c["params"]), nside, mp, bv[:s], b, s, 1, 2)
        g = _gpu_tod(K, mode, *args, variant=variant)
        r = _cpu_tod(mode, *args)
        if b == 1:
            scale = np.sqrt(np.mean(mp.astype(np.float64) ** 2, axis=1, keepdims=True))
            err = float(np.max(np.abs(g - r) / scale))
        else:
            err = rel_err(g, r)
        print(f"{mode} variant {variant} B={b} S={s}: rel err {err:.3e}")
        assert err <= TOL


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_non_finite_beam_value_propagates(K, mode):
    """A NaN beam value makes every sample's TOD NaN, as the reference's sums
    do (tod_bilinear.py:233-477 adds beam_val * value for every pixel); such a
    beam has no weight grid and runs on the direct path."""
    c = small_config()
    B = 200
    params = tuple(p[:B] for p in c["params"])
    bv = c["bv"].copy()
    bv[7] = np.nan
    g = _gpu_tod(K, mode, c["vec"], params, c["nside"], c["map"], bv, B, c["S"], 1, 2)
    assert np.all(np.isnan(g))


def test_zero_beam_and_constant_map(K):
    rng = np.random.default_rng(41)
    nside, B, S = 32, 16, 40
    vec, params = random_rot_params(B, S, rng)
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    z = _gpu_tod(K, "bilinear", vec, params, nside, mp, np.zeros(S, np.float32), B, S, 1, 2)
    assert np.array_equal(z, np.zeros_like(z))
    const = np.full((2, 12 * nside * nside), 7.5, np.float32)
    bv = np.full(S, 1.0 / S, np.float32)
    t = _gpu_tod(K, "bilinear", vec, params, nside, const, bv, B, S, -1, -1)
    np.testing.assert_allclose(t, 7.5, atol=1e-4)


def test_q_one_at_boresight(K):
    """Q=1 map, single beam pixel on the boresight -> (T,Q,U) = (0,1,0) within
    the spin-2 residual of one pixel (tests/test_tod_bilinear.py:1027-1076)."""
    nside = 64
    rng = np.random.default_rng(50)
    B = 4
    ax = rng.standard_normal((B, 3))
    ax /= np.linalg.norm(ax, axis=-1, keepdims=True)
    ax[ax[:, 2] < -0.95, 2] = -0.9
    ax /= np.linalg.norm(ax, axis=-1, keepdims=True)
    params = recenter_params(ax.astype(np.float32))
    vec = np.array([[0.0, 0.0, 1.0]], np.float32)
    mp = np.zeros((3, 12 * nside * nside), np.float32)
    mp[1] = 1.0
    t = _gpu_tod(K, "bilinear", vec, params, nside, mp, np.ones(1, np.float32), B, 1, 1, 2)
    np.testing.assert_allclose(t[0], 0.0, atol=1e-3)
    np.testing.assert_allclose(t[1], 1.0, atol=1e-3)
    np.testing.assert_allclose(t[2], 0.0, atol=1e-3)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_spin2_guard_at_coincident_pixel(K, mode, variant):
    """Boresight exactly on a map pixel centre: the reference returns
    (cos 2d, sin 2d) = (1, 0) for that pixel (tod_bilinear.py:108-109).  At
    nside 1 the equator ring's pixel 5 sits at phi = pi/2, z = 0, which the
    float32 boresight (0, 1, 0) hits bit for bit; the beam spreads around it
    so that the coincident pixel and its neighbours all carry weight."""
    nside = 1
    rng = np.random.default_rng(61)
    ax = np.array([[0.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.6, 0.8]], np.float32)
    params = recenter_params(ax)
    S = 40
    v = rng.standard_normal((S, 3)) * np.array([0.3, 0.3, 0.0]) + np.array([0.0, 0.0, 1.0])
    v[0] = [0.0, 0.0, 1.0]
    vec = (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)
    mp = rng.uniform(-1.0, 1.0, (3, 12)).astype(np.float32)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    B = ax.shape[0]
    g = _gpu_tod(K, mode, vec, params, nside, mp, bv, B, S, 1, 2, variant=variant)
    r = _cpu_tod(mode, vec, params, nside, mp, bv, B, S, 1, 2)
    assert np.isfinite(g).all()  # no 0/0 from the coincident pixel's spin-2 terms
    assert rel_err(g, r) <= TOL


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_footprint_centre_just_below_two_pi(K, mode):
    """Boresights within half a step of the footprint's longitude grid (512
    steps) below 2 pi, where the grid index wraps to 0: the footprint's
    windows are then laid out from longitude origin 0 as well."""
    nside = 256
    rng = np.random.default_rng(71)
    B, S = 48, 200
    th = rng.uniform(0.3, 2.8, B)
    ph = 2.0 * np.pi - rng.uniform(0.0, 0.95 * np.pi / 512, B)
    ax = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1).astype(np.float32)
    params = recenter_params(ax)
    v = rng.standard_normal((S, 3)) * np.array([0.01, 0.01, 0.0]) + np.array([0.0, 0.0, 1.0])
    vec = (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)
    mp = rng.uniform(-1.0, 1.0, (3, 12 * nside * nside)).astype(np.float32)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    g = _gpu_tod(K, mode, vec, params, nside, mp, bv, B, S, 1, 2)
    r = _cpu_tod(mode, vec, params, nside, mp, bv, B, S, 1, 2)
    assert rel_err(g, r) <= TOL


def test_beam_tod_batch_dropin():
    """The reference's dispatcher contract end to end (numpy in, dict of
    float32 out) against the oracle."""
    from tod_generation_mapbased_beam_amd import beam_tod_batch
    from tod_generation_mapbased_beam_amd import tod_rotations as R

    c = small_config()
    B = 2000
    th = c["theta"][:B].astype(np.float32)
    ph = c["phi"][:B].astype(np.float32)
    ps = c["psi"][:B].astype(np.float32)
    rot, beta = R.precompute_rotation_vector_batch(np.zeros((1, 1)), np.full((1, 1), math.pi / 2), ph, th, center_idx=(0, 0))
    psis = -beta + ps
    data = dict(vec_orig=c["vec"], beam_vals=c["bv"], comp_indices=[0, 1, 2], mp_stacked=c["map"])
    for mode in ("bilinear", "nearest"):
        out = beam_tod_batch(c["nside"], None, data, rot, ph, th, psis, interp_mode=mode)
        assert sorted(out) == [0, 1, 2]
        assert all(v.dtype == np.float32 and v.shape == (B,) for v in out.values())
        params = R._rotation_params(rot, ph, th, psis)
        ref = _cpu_tod(mode, c["vec"], params, c["nside"], c["map"], c["bv"], B, c["S"], 1, 2)
        got = np.stack([out[k] for k in (0, 1, 2)]).astype(np.float64)
        assert rel_err(got, ref.astype(np.float32)) <= 2e-6
    out = beam_tod_batch(c["nside"], None, data, rot, ph, th, psis, interp_mode="bicubic")
    params = R._rotation_params(rot, ph, th, psis)
    ref = np.zeros((3, B))
    O.gather_accum_bicubic(c["vec"], *params, c["nside"], c["map"], c["bv"], B, c["S"], ref, 1, 2)
    got = np.stack([out[k] for k in (0, 1, 2)]).astype(np.float64)
    assert rel_err(got, ref.astype(np.float32)) <= 2e-6
    with pytest.raises(ValueError):
        beam_tod_batch(c["nside"], None, data, rot, ph, th, psis, interp_mode="gaussian")
