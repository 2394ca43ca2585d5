"""Beam transfer function: the oracle (oracle/bell_oracle.py) pinned by closed
forms, the GPU recurrence (todg_beam_bell via beam_tools.compute_bell)
against the oracle, and the clustering calibration sweep's contract
(tod_calibrate.py:392-530)."""

import math

import numpy as np
import pytest
from scipy.special import eval_legendre

import bell_oracle as BO


def test_oracle_delta_beam_is_flat():
    _, b = BO.compute_bell([0.0], [0.0], [1.0], lmax=500, power_cut=1.0)
    np.testing.assert_allclose(b, 1.0, rtol=0, atol=1e-12)


def test_oracle_two_point_beam_equals_legendre():
    ra = np.array([0.3, 1.1])
    dec = np.array([0.2, -0.4])
    w = np.array([0.25, 0.75])
    _, b = BO.compute_bell(ra, dec, w, lmax=300, power_cut=1.0, normalise=False)
    x = np.cos(dec) * np.cos(ra)
    ell = np.arange(301)
    ref = np.abs(w[0] * eval_legendre(ell, x[0]) + w[1] * eval_legendre(ell, x[1]))
    np.testing.assert_allclose(b, ref, rtol=1e-10, atol=1e-13)


def test_oracle_gaussian_beam_transfer():
    from tod_generation_mapbased_beam_amd import inputs

    fwhm = 60.0
    vec, bv, vals = inputs.gaussian_beam(1024, fwhm, 0.9999)
    th = np.arccos(np.clip(vec[:, 2].astype(np.float64), -1, 1))
    ph = np.arctan2(vec[:, 1], vec[:, 0]).astype(np.float64)
    # small (ra, dec) offsets from the beam centre, as for a CAR beam:
    # cos(dec) cos(ra) = cos(theta) up to O(theta^4)
    _, b = BO.compute_bell(th * np.cos(ph), th * np.sin(ph), bv, lmax=400, power_cut=1.0)
    sigma = fwhm * math.pi / 180 / 60 / math.sqrt(8 * math.log(2))
    ell = np.arange(401)
    ref = np.exp(-ell * (ell + 1) * sigma ** 2 / 2)
    assert np.max(np.abs(b[:200] - ref[:200])) < 5e-3


def test_oracle_db_threshold_keeps_power():
    rng = np.random.default_rng(2)
    v = rng.exponential(1.0, 5000)
    t = BO.db_threshold(v, 0.9)
    kept = v[10 * np.log10(v) > t].sum() / v.sum()
    assert 0.895 < kept <= 0.9 + 1e-3


@pytest.mark.gpu
def test_gpu_bell_matches_oracle():
    from tod_generation_mapbased_beam_amd import beam_tools, inputs

    vec, bv, vals = inputs.gaussian_beam(1024, 45.0, 0.999, ellipticity=0.1, angle=0.3)
    th = np.arccos(np.clip(vec[:, 2].astype(np.float64), -1, 1))
    ph = np.mod(np.arctan2(vec[:, 1], vec[:, 0]), 2 * math.pi).astype(np.float64)
    for pc in (1.0, 0.99):
        _, got = beam_tools.compute_bell(ph, math.pi / 2 - th, bv, lmax=2000, power_cut=pc, verbose=False)
        _, ref = BO.compute_bell(ph, math.pi / 2 - th, bv, lmax=2000, power_cut=pc)
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
def test_gpu_clustering_calibration_contract():
    from tod_generation_mapbased_beam_amd import beam_tools, inputs

    vec, bv, _ = inputs.gaussian_beam(512, 60.0, 0.999)
    data = {"beam": {"vec_orig": vec, "beam_vals": bv}}
    tf, K, table = beam_tools.calibrate_beam_clustering(
        data, error_threshold=1e-3, bell_lmax=200, tail_fractions=(0.01, 0.05),
        n_clusters=(10, 50, 200), verbose=False)
    assert (tf, K) in {(r[0], r[1]) for r in table}
    row = [r for r in table if (r[0], r[1]) == (tf, K)][0]
    passing = [r for r in table if r[4] <= 1e-3]
    if passing:
        assert row[3] == max(r[3] for r in passing)
    else:
        assert row[4] == min(r[4] for r in table)
