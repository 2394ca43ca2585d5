"""Host-side checks that need no GPU: the C-ABI library loads and exports
every symbol include/todgpu.h declares, its pure-host entry points agree
with the oracle, the Python boundary raises like the reference, and the
input generators keep their contracts."""

import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(REPO, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(REPO, "include", name)).read()
            syms |= set(re.findall(r"^int\s+\**(todg_\w+)\s*\(", text, flags=re.M))
    return syms


def test_library_exports_every_declared_symbol():
    from tod_generation_mapbased_beam_amd import _lib

    L = _lib.lib()
    syms = _declared_symbols()
    assert len(syms) >= 9
    for s in syms:
        assert hasattr(L, s), s
    assert set(_lib.SIGNATURES) >= syms
    assert L.todg_version() == _lib.ABI_VERSION


def test_checked_library_exports_the_same_symbols():
    """libtodgpu_checked.so (the -DTODG_CHECKED build) binds the same C-ABI;
    the release library refuses its counter query without touching a GPU."""
    from tod_generation_mapbased_beam_amd import _lib

    C = _lib._load(_lib.CHECKED_LIB_PATH)
    for s in _declared_symbols():
        assert hasattr(C, s), s
    assert C.todg_version() == _lib.ABI_VERSION
    buf = (ctypes.c_ulonglong * 3)()
    assert _lib.lib().todg_check_counts(buf, 3, 1) == _lib.TODG_ERR_ARG
    assert "TODG_CHECKED" in _lib.last_error()
    assert _lib.lib().todg_check_counts(None, 3, 1) == _lib.TODG_ERR_ARG


def test_checked_context_swaps_and_restores_the_library():
    """Inside _lib.checked() every call goes to the checked build (here a
    host-only entry point, which must agree with the release one); the
    release library is back afterwards, also after an exception."""
    from tod_generation_mapbased_beam_amd import _lib

    release = _lib.lib()

    def table(h, nside=16):
        n = 4 * nside + 1
        arrs = [np.zeros(n) for _ in range(4)] + [np.zeros(n, np.int64), np.zeros(n, np.int64)]
        _lib.check(h.todg_ring_table(nside, *(a.ctypes.data for a in arrs)))
        return arrs

    with _lib.checked() as C:
        assert _lib.lib() is C and C is not release
        got = table(_lib.lib())
    assert _lib.lib() is release
    for a, b in zip(got, table(release)):
        assert np.array_equal(a, b)
    with pytest.raises(RuntimeError):
        with _lib.checked():
            raise RuntimeError("inside")
    assert _lib.lib() is release


@pytest.mark.parametrize("nside", [1, 4, 128, 2048])
def test_ring_table_bit_identical_to_oracle(nside):
    from tod_generation_mapbased_beam_amd import _lib

    n = 4 * nside + 1
    z, th, p0, dp = (np.zeros(n) for _ in range(4))
    fp, nn = np.zeros(n, np.int64), np.zeros(n, np.int64)
    _lib.check(_lib.lib().todg_ring_table(nside, *(a.ctypes.data for a in (z, th, p0, dp, fp, nn))))
    for ir in range(1, 4 * nside, max(1, nside // 7)):
        on, ofp, ophi0, odphi = O.ring_info(nside, ir)
        assert (nn[ir], fp[ir]) == (on, ofp)
        assert p0[ir] == ophi0 and dp[ir] == odphi
        assert z[ir] == O.ring_z(nside, ir)
        assert th[ir] == math.acos(O.ring_z(nside, ir))


def test_map_slots_and_argument_errors():
    from tod_generation_mapbased_beam_amd import _lib
    from tod_generation_mapbased_beam_amd.device import map_slots

    assert map_slots(3, 1, 2) == (4, [1, 2, 0, -1], [1, 2, 0, 0])
    assert map_slots(3, -1, -1) == (4, [0, 1, 2, -1], [0, 1, 2, 0])
    assert map_slots(1, -1, -1)[0] == 1
    assert map_slots(2, 0, 0) == (4, [0, 0, 1, -1], [0, 0, 1, 0])  # c_q == c_u
    with pytest.raises(ValueError):
        map_slots(5, -1, -1)
    with pytest.raises(ValueError):
        map_slots(3, 3, 1)
    L = _lib.lib()
    rc = L.todg_ring_table(3, None, None, None, None, None, None)
    assert rc == _lib.TODG_ERR_ARG and "power of two" in _lib.last_error()


def test_beam_tod_batch_rejects_unknown_mode():
    """tod_core.py:114-119: ValueError before any work."""
    from tod_generation_mapbased_beam_amd import beam_tod_batch

    data = dict(vec_orig=np.zeros((1, 3), np.float32), beam_vals=np.ones(1, np.float32),
                comp_indices=[0], mp_stacked=np.zeros((1, 12), np.float32))
    with pytest.raises(ValueError):
        beam_tod_batch(1, None, data, np.zeros((2, 3)), np.zeros(2, np.float32),
                       np.zeros(2, np.float32), np.zeros(2), interp_mode="gaussian")


def test_qu_slot_derivation():
    from tod_generation_mapbased_beam_amd.tod_core import _qu_slots

    assert _qu_slots([0, 1, 2]) == (1, 2)
    assert _qu_slots([1, 2]) == (0, 1)
    assert _qu_slots([0]) == (-1, -1)
    assert _qu_slots([1]) == (0, -1)  # Q without U: no spin-2 (tod_bilinear.py:276)


def test_db_threshold_and_selection():
    """tod_utils.py:254-297 and sample_based_tod_generation_gridint.py:127-129:
    the threshold pixel is the weakest whose cumulative (from the top) still
    reaches power_cut of the sum; the selection keeps the pixels strictly
    above it, so kept + threshold pixel reaches the target and kept alone
    does not."""
    from tod_generation_mapbased_beam_amd import inputs

    rng = np.random.default_rng(0)
    v = rng.exponential(size=5000)
    vecs = rng.standard_normal((5000, 3))
    for cut in (0.9, 0.99, 0.999):
        vec, bv, sel = inputs.select_beam(v, vecs, cut)
        kept = v[sel]
        thr = v[~sel].max()
        assert kept.sum() + thr >= cut * v.sum() * (1 - 1e-12)
        assert kept.sum() < cut * v.sum()
        assert kept.min() > thr
        assert abs(bv.sum() - 1.0) < 1e-5 and bv.dtype == np.float32 and vec.dtype == np.float32


def test_gaussian_beam_size_and_shape():
    from tod_generation_mapbased_beam_amd import inputs

    vec, bv, vals = inputs.gaussian_beam(512, 60.0, 0.99)
    # 99% of a Gaussian lies within 3.035 sigma: S ~ pi r^2 / pixel area
    sigma = 60.0 / math.sqrt(8 * math.log(2)) * math.pi / 10800
    r = sigma * math.sqrt(2 * math.log(100))
    expect = math.pi * r * r / (4 * math.pi / (12 * 512 ** 2))
    assert abs(vec.shape[0] - expect) / expect < 0.1
    np.testing.assert_allclose(np.linalg.norm(vec, axis=1), 1.0, atol=1e-6)
    assert np.all(np.arccos(vec[:, 2]) < 1.1 * r)


def test_ring_scan_contract():
    from tod_generation_mapbased_beam_amd import inputs

    th, ph, ps = inputs.ring_scan(2000, 10.0)
    assert np.all((th >= math.radians(45) - 1e-9) & (th <= math.radians(135) + 1e-9))
    assert np.all((ph >= 0) & (ph < 2 * math.pi)) and np.all((ps >= 0) & (ps < 2 * math.pi))
    # 1 rpm spin at 10 Hz: 600 samples per revolution
    np.testing.assert_allclose(ps[600], ps[0], atol=1e-9)


def test_synthesised_map_spectrum():
    """Spherical-harmonic synthesis: exact single harmonics (Y00, Y10, Y11)
    and Parseval for a full realisation, sum_p f_p^2 (4pi/npix) =
    sum_lm |a_lm|^2 over m in [-l, l], to pixel-quadrature accuracy."""
    from tod_generation_mapbased_beam_amd import healpix_np as hpx
    from tod_generation_mapbased_beam_amd import inputs

    nside, lmax = 32, 95
    npix = 12 * nside * nside
    z, phi = hpx.pix2zphi(nside, np.arange(npix))
    st = np.sqrt(1 - z * z)

    def single(l, m, val):
        a = np.zeros((lmax + 1, lmax + 1), complex)
        a[l, m] = val
        return inputs.synthesize_map(nside, [a], lmax)[0]

    np.testing.assert_allclose(single(0, 0, 2.0), 2.0 / math.sqrt(4 * math.pi), rtol=1e-12)
    np.testing.assert_allclose(single(1, 0, 1.0), math.sqrt(3 / (4 * math.pi)) * z, atol=1e-12)
    # real field from a_11 = 1: 2 Re(Y_11) = -2 sqrt(3/8pi) sin(theta) cos(phi)
    np.testing.assert_allclose(single(1, 1, 1.0), -2 * math.sqrt(3 / (8 * math.pi)) * st * np.cos(phi),
                               atol=1e-12)
    tt, _, _ = inputs.cl_spectrum(lmax)
    a = inputs._alm_realisation(tt, 60, np.random.default_rng(1))
    a_full = np.zeros((lmax + 1, lmax + 1), complex)
    a_full[:61, :61] = a
    f = inputs.synthesize_map(nside, [a_full], lmax)[0]
    power = np.sum(np.abs(a[:, 0]) ** 2) + 2 * np.sum(np.abs(a[:, 1:]) ** 2)
    assert abs(np.sum(f * f) * 4 * math.pi / npix / power - 1) < 0.01


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    from tod_generation_mapbased_beam_amd import autotune

    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "tune.json"))
    k = autotune.workload_key("single", 2048, 6364, "bilinear", 3, True)
    assert "S~8192" in k and "nside=2048" in k
    assert autotune.lookup(k) is None
    assert autotune.tuned_variant(2048, 6364, "bilinear", 3, True, default=0) == 0
    autotune._store(k, {"variant": 8})
    assert autotune.tuned_variant(2048, 6000, "bilinear", 3, True) == 8  # same bucket
    assert autotune.tuned_variant(2048, 9000, "bilinear", 3, True) == 0  # other bucket


def test_rotation_params_split_equals_serial():
    """beam_tod_batch's numpy path computes _rotation_params in pieces on a
    thread pool: bit for bit the serial result, ragged pieces included."""
    from tod_generation_mapbased_beam_amd import inputs
    from tod_generation_mapbased_beam_amd import tod_rotations as R

    n = 3 * R.SPLIT_MIN + 4321
    th, ph, ps = inputs.ring_scan(n, 20.0)
    phb, thb, psb = (np.asarray(x, np.float32) for x in (ph, th, ps))
    rv, beta = R.precompute_rotation_vector_batch(np.zeros((1, 1)), np.full((1, 1), np.pi / 2), phb, thb)
    want = R._rotation_params(rv, phb, thb, -beta + psb)
    for workers in (1, 3, 16):
        got = R._rotation_params_split(rv, phb, thb, -beta + psb, workers=workers)
        for a, b in zip(want, got):
            assert a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b)


def test_rotation_params_pool_survives_fork():
    """A process forked after the shared pool exists makes its own pool
    (the inherited one has no threads) and gets the same operands."""
    from tod_generation_mapbased_beam_amd import inputs
    from tod_generation_mapbased_beam_amd import tod_rotations as R

    n = 2 * R.SPLIT_MIN + 11
    th, ph, ps = inputs.ring_scan(n, 20.0)
    phb, thb, psb = (np.asarray(x, np.float32) for x in (ph, th, ps))
    rv, beta = R.precompute_rotation_vector_batch(np.zeros((1, 1)), np.full((1, 1), np.pi / 2), phb, thb)
    want = R._rotation_params_split(rv, phb, thb, -beta + psb)
    pid = os.fork()
    if pid == 0:  # child: exit status 0 iff the operands agree
        got = R._rotation_params_split(rv, phb, thb, -beta + psb)
        os._exit(0 if all(np.array_equal(a, b) for a, b in zip(want, got)) else 1)
    _, status = os.waitpid(pid, 0)
    assert os.WIFEXITED(status) and os.WEXITSTATUS(status) == 0
