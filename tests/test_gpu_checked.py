"""The checked build on the production paths.

libtodgpu_checked.so is the library compiled from the same sources with
-DTODG_CHECKED: every map gather, weight-grid deposit and ring-slot lookup
also tests its index and counts a violation (the clamps stay, so nothing is
accessed out of bounds).  Each case runs once through the release library
and once through the checked one, and asserts that the checked run counted
no violation and returned the release library's output bit for bit (the
checks only count; the arithmetic is the same).  The cases cover the fast
and exact footprint paths, direct accumulation (footprints wider than the
LDS grid), poles, the phi = 0 seam, the equatorial-belt edges, tiny nsides
and non-finite pointings.  One positive control (the reference's float-mod
corner at the last ring, where a neighbour index reaches npix) shows that
the counters fire.
"""

import math

import numpy as np
import pytest

from cases import random_rot_params, recenter_params, small_config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    import torch  # noqa: F401

    from tod_generation_mapbased_beam_amd import _lib

    _lib.lib()
    _lib.check_counts(reset=True)
    return _lib


def _host(x):
    if isinstance(x, (tuple, list)):
        return [_host(v) for v in x]
    if hasattr(x, "cpu"):
        return x.cpu().numpy()
    return np.asarray(x)


def _release_and_checked(L, fn):
    """fn() under the release library, then under the checked one; returns
    both outputs (host) and the checked run's violation counts."""
    ref = _host(fn())
    L.check_counts(reset=True)
    with L.checked():
        got = _host(fn())
    return ref, got, L.check_counts(reset=True)


def _assert_same(ref, got, counts, label):
    print(f"{label}: counts {counts}", flush=True)
    assert counts == {k: 0 for k in counts}, f"{label}: index violations {counts}"
    for a, b in zip(ref if isinstance(ref, list) else [ref], got if isinstance(got, list) else [got]):
        assert a.shape == b.shape
        assert np.array_equal(a, b, equal_nan=True), f"{label}: checked build differs from release"


def _tod_fn(mode, vec, params, nside, mp, bv, c_q, c_u, variant=0):
    from tod_generation_mapbased_beam_amd import kernels as K

    fn = {"bilinear": K._gather_accum_fused_jit, "nearest": K._gather_accum_nearest_jit,
          "bicubic": K._gather_accum_bicubic_jit}[mode]
    B, S = params[0].shape[0], vec.shape[0]

    def run():
        tod = np.zeros((mp.shape[0], B))
        fn(vec, *params, nside, mp, bv, B, S, tod, c_q, c_u, variant=variant)
        return tod

    return run


def test_release_library_rejects_check_counts(L):
    import ctypes

    buf = (ctypes.c_ulonglong * 3)()
    assert L.lib().todg_check_counts(buf, 3, 1) == L.TODG_ERR_ARG
    assert "TODG_CHECKED" in L.last_error()


def test_counts_the_last_ring_float_mod_corner(L):
    """Positive control of the counters: in the south polar cap, a longitude
    one ulp below the last ring's phi0 makes the slot coordinate round to
    exactly n (the reference's float-mod corner, numba_healpy.py:454-456),
    so one neighbour index is npix, one past the map, which the reference
    (no bounds checks) would read.  Both libraries clamp it to the last
    pixel, as the oracle does; the checked one counts it once per such
    direction, and nothing else."""
    import oracle as O
    from tod_generation_mapbased_beam_amd import kernels as K

    nside = 4  # last ring: 4 pixels, phi0 = pi/4, dphi = pi/2
    corner = float(np.nextafter(math.pi / 4, 0.0))
    assert ((corner - math.pi / 4) / (math.pi / 2)) % 4.0 == 4.0
    z = np.array([-0.99, -0.995, 0.3, -0.99, 0.99])
    phi = np.array([corner, corner, corner, 1.0, corner])
    m = np.random.default_rng(9).standard_normal(12 * nside * nside).astype(np.float32)
    ref, got, counts = _release_and_checked(L, lambda: K.interp_points(nside, m, z, phi, "bilinear"))
    print(f"corner: counts {counts}", flush=True)
    assert counts == {"map_pixel": 2, "grid_cell": 0, "ring_slot": 0}
    assert np.array_equal(ref, got)
    want = O.interp_values(nside, m, z, phi, "bilinear")
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("mode", ["bilinear", "nearest", "bicubic"])
@pytest.mark.parametrize("variant", [0, 8])
def test_compact_beam_ring_scan(L, mode, variant):
    """The production shape: a compact beam on a ring scan (fast path)."""
    c = small_config(n_samples=3000)
    ref, got, counts = _release_and_checked(
        L, _tod_fn(mode, c["vec"], c["params"], c["nside"], c["map"], c["bv"], 1, 2, variant))
    _assert_same(ref, got, counts, f"{mode} v{variant}")


@pytest.mark.parametrize("mode", ["bilinear", "nearest", "bicubic"])
@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64])
def test_beams_spread_over_the_sphere(L, mode, nside):
    """Beam pixels anywhere on the sphere, pointings anywhere (poles
    included): wide footprints, the exact path and direct accumulation."""
    rng = np.random.default_rng(100 + nside)
    B, S = 12, 48
    vec, params = random_rot_params(B, S, rng, pole_safe=False)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    ref, got, counts = _release_and_checked(L, _tod_fn(mode, vec, params, nside, mp, bv, 1, 2))
    _assert_same(ref, got, counts, f"{mode} nside {nside}")


@pytest.mark.parametrize("mode", ["bilinear", "nearest", "bicubic"])
def test_poles_seam_and_belt_edges(L, mode):
    """A compact beam centred on the poles, on the phi = 0 seam, on the
    belt/cap boundaries |z| = 2/3 and on the first and last rings."""
    from tod_generation_mapbased_beam_amd import inputs

    nside = 64
    vec, bv, _ = inputs.gaussian_beam(256, 120.0, 0.99)
    pts = []
    for z in (1.0, -1.0, 2.0 / 3.0, -2.0 / 3.0, 0.0, 1.0 - 1.0 / (3 * nside * nside),
              -1.0 + 1.0 / (3 * nside * nside)):
        for ph in (0.0, 1e-12, 2 * math.pi - 1e-12, math.pi / 4):
            s = math.sqrt(max(0.0, 1.0 - z * z))
            pts.append([s * math.cos(ph), s * math.sin(ph), z])
    params = recenter_params(np.asarray(pts, np.float32))
    mp = np.random.default_rng(7).standard_normal((3, 12 * nside * nside)).astype(np.float32)
    ref, got, counts = _release_and_checked(L, _tod_fn(mode, vec, params, nside, mp, bv, 1, 2))
    _assert_same(ref, got, counts, f"{mode} poles/seam")


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_non_finite_pointings(L, mode):
    """Samples with non-finite rotation operands are skipped; nothing they
    would have indexed is touched."""
    c = small_config(n_samples=200, nside=64)
    params = [p.copy() for p in c["params"]]
    params[1][3] = np.nan
    params[3][7] = np.inf
    params[5][11] = np.nan
    ref, got, counts = _release_and_checked(
        L, _tod_fn(mode, c["vec"], params, c["nside"], c["map"], c["bv"], 1, 2))
    _assert_same(ref, got, counts, f"{mode} non-finite")


def test_learned_corrector_kernels(L):
    """The factorised corrector's channel kernel and fused TOD, and the MLP
    stencil corrector's channel kernel."""
    import torch

    from tod_generation_mapbased_beam_amd import fcorr as F
    from tod_generation_mapbased_beam_amd import smlp

    c = small_config(n_samples=2000)
    coef = np.concatenate([[1.0], 0.3 * np.random.default_rng(1).standard_normal(31)])

    def run_fcorr():
        maps = F.FcorrMaps(c["map"], coef, 1, 2)
        tod = torch.zeros((3, c["B"]), dtype=torch.float64, device="cuda")
        F.gather_accum_fcorr(c["vec"], *c["params"], c["nside"], maps, c["bv"], c["B"], c["S"], tod)
        return [maps.packed0, maps.packed1, tod]

    ref, got, counts = _release_and_checked(L, run_fcorr)
    _assert_same(ref, got, counts, "factorised corrector")

    nside = 32
    m = torch.as_tensor(np.random.default_rng(3).standard_normal(12 * nside * nside).astype(np.float32),
                        device="cuda")
    flat = smlp.init_params(1, F.FactorisedCorrector.load_default().coef, 1.0)

    def run_smlp():
        A0 = torch.empty(m.numel(), dtype=torch.float32, device="cuda")
        A1 = torch.empty_like(A0)
        smlp.channels(flat, nside, m, A0, A1, 1)
        return [A0, A1]

    ref, got, counts = _release_and_checked(L, run_smlp)
    _assert_same(ref, got, counts, "MLP stencil channels")


def test_per_beam_pixel_mlp_corrector(L):
    from tod_generation_mapbased_beam_amd import corrector as M
    from tod_generation_mapbased_beam_amd.corrector import gather_accum_learned

    rng = np.random.default_rng(22)
    nside, B, S = 16, 6, 40
    vec, params = random_rot_params(B, S, rng)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    flat = M.init_params(3) * 4.0

    def run():
        got = np.zeros((3, B))
        gather_accum_learned(vec, *params, nside, mp, bv, B, S, got, 1, 2, params=flat)
        return got

    ref, got, counts = _release_and_checked(L, run)
    _assert_same(ref, got, counts, "per-beam-pixel MLP corrector")


def test_batch_and_points(L):
    """The persistent batched kernel over uneven detectors, and the point
    interpolation kernels (bilinear and bicubic) at poles and the seam."""
    import torch

    from tod_generation_mapbased_beam_amd import inputs, kernels as K
    from tod_generation_mapbased_beam_amd import tod_rotations as R
    from tod_generation_mapbased_beam_amd.batch import BatchPlan, Detector
    from tod_generation_mapbased_beam_amd.device import DeviceMap

    nside = 64
    rng = np.random.default_rng(5)
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    dets = []
    for i, (fwhm, nb) in enumerate([(90.0, 300), (60.0, 57), (120.0, 500)]):
        vec, bv, _ = inputs.gaussian_beam(256, fwhm, 0.99, ellipticity=0.1, angle=0.3 * i)
        th, ph, ps = inputs.ring_scan(nb, 10.0, t0=100.0 * i)
        dets.append(Detector(vec, bv, R.rotation_params_for_pointing(th, ph, ps)))

    def run_batch():
        plan = BatchPlan(dets, DeviceMap(mp, 1, 2), chunk=7)
        plan.launch()
        torch.cuda.synchronize()
        assert plan.failed_items() == []
        return [t.clone() for t in plan.tods]

    ref, got, counts = _release_and_checked(L, run_batch)
    _assert_same(ref, got, counts, "batch")

    z = np.concatenate([rng.uniform(-1, 1, 2000), [1.0, -1.0, 2 / 3, -2 / 3, 0.0]])
    phi = np.concatenate([rng.uniform(0, 2 * math.pi, 2000), [0.0, 2 * math.pi - 1e-12, 0.0, 1e-12, 0.0]])
    for mode in ("bilinear", "bicubic"):
        ref, got, counts = _release_and_checked(L, lambda: K.interp_points(nside, mp[0], z, phi, mode))
        _assert_same(ref, got, counts, f"points {mode}")
