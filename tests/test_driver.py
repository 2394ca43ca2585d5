"""The per-day driver (tod_exact_gen_batched / run_days): scan files round
trip and per-day failure isolation on the CPU; on the GPU, saved days equal
the oracle on the same host-prepared rotation operands, for one and for
several batches per day."""

import math
import os

import numpy as np
import pytest

import oracle as O
from tod_generation_mapbased_beam_amd import driver, inputs
from tod_generation_mapbased_beam_amd import tod_rotations as R


def test_scan_files_round_trip(tmp_path):
    th, ph, ps = inputs.ring_scan(100, 10.0)
    driver.write_scan_day(str(tmp_path), 3, th, ph, ps)
    t2, p2, s2 = driver.open_scan_day(str(tmp_path), 3)
    assert isinstance(t2, np.memmap)
    np.testing.assert_array_equal(t2, th)
    np.testing.assert_array_equal(p2, ph)
    np.testing.assert_array_equal(s2, ps)


def test_missing_day_is_reported_and_others_go_on(tmp_path):
    """No scan files at all: every day fails before any GPU work, nothing is
    written, and the failures are returned (sample_based_tod_generation_gridint.py:319-333)."""
    nside = 8
    mp = np.zeros((3, 12 * nside * nside), np.float32)
    vec, bv, _ = inputs.gaussian_beam(16, 600.0, 0.99)
    bd = {"b": driver.beam_entry(vec, bv, [0, 1, 2], mp)}
    out = tmp_path / "out"
    res = driver.run_days(bd, mp, [0, 1], 100, str(tmp_path / "scan"), str(out), log=None)
    assert [r[0] for r in res] == [0, 1]
    assert all(not ok and err for _, ok, err in res)
    assert os.listdir(out) == []


def test_host_params_split_over_threads_changes_nothing(tmp_path):
    """The host path's rotation operands computed in pieces by a thread pool
    equal the serial numpy path's bit for bit, ragged pieces included."""
    from concurrent.futures import ThreadPoolExecutor

    n = 3 * driver.HOST_SPLIT_MIN + 12345
    th, ph, ps = inputs.ring_scan(n, 20.0)
    driver.write_scan_day(str(tmp_path), 0, th, ph, ps)
    mm = driver.open_scan_day(str(tmp_path), 0)
    ra0, dec0 = np.zeros((1, 1)), np.full((1, 1), math.pi / 2)
    serial = driver._host_params(*mm, 5, n, ra0, dec0)
    want = R._rotation_params(*driver._host_batch(*mm, 5, n, ra0, dec0))
    with ThreadPoolExecutor(max_workers=4) as pool:
        split = driver._host_params(*mm, 5, n, ra0, dec0, pool, 7)
    for a, b, c in zip(serial, split, want):
        assert a.dtype == b.dtype == c.dtype and a.shape == b.shape == c.shape
        assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.gpu
@pytest.mark.parametrize("prep", ["host", "device"])
@pytest.mark.parametrize("batch_size", [10_000, 333])
def test_days_match_oracle(tmp_path, batch_size, prep):
    """Saved days against the oracle on the operands the driver used: the
    numpy path's (prep="host"), or those todg_rotation_params computes on
    the GPU (prep="device", downloaded)."""
    rng = np.random.default_rng(5)
    nside = 32
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    vec, bv, _ = inputs.gaussian_beam(64, 240.0, 0.99)
    # I alone and a shared Q/U entry (spin-2 on), as two beam files would give
    bd = {"I": driver.beam_entry(vec, bv, [0], mp), "QU": driver.beam_entry(vec, bv, [1, 2], mp)}
    days = [0, 1]
    n = 1200
    for d in days:
        th, ph, ps = inputs.ring_scan(n, 10.0, t0=1000.0 * d)
        driver.write_scan_day(str(tmp_path / "scan"), d, th, ph, ps)
    res = driver.run_days(bd, mp, days + [7], batch_size, str(tmp_path / "scan"), str(tmp_path / "out"),
                          log=None, prep=prep)
    assert [ok for _, ok, _ in res] == [True, True, False]  # day 7 has no scan files
    S = vec.shape[0]
    for d in days:
        got = np.load(tmp_path / "out" / f"tod_day_{d}.npy")
        assert got.shape == (3, n) and got.dtype == np.float64
        th, ph, ps = driver.open_scan_day(str(tmp_path / "scan"), d)
        if prep == "host":
            phi_b, theta_b, psi_b = (np.array(x, dtype=np.float32) for x in (ph, th, ps))
            rv, betas = R.precompute_rotation_vector_batch(np.zeros((1, 1)), np.full((1, 1), math.pi / 2),
                                                          phi_b, theta_b)
            params = R._rotation_params(rv, phi_b, theta_b, -betas + psi_b)
        else:
            from tod_generation_mapbased_beam_amd import kernels

            ctr = R.beam_centre(np.zeros((1, 1)), np.full((1, 1), math.pi / 2))
            params = [p.cpu().numpy() for p in kernels.rotation_params(np.array(th), np.array(ph),
                                                                      np.array(ps), ctr)]
        ref = np.zeros((3, n))
        t1 = np.zeros((1, n))
        O.gather_accum_fused(vec, *params, nside, mp[[0]], bv, n, S, t1, -1, -1)
        t2 = np.zeros((2, n))
        O.gather_accum_fused(vec, *params, nside, mp[[1, 2]], bv, n, S, t2, 0, 1)
        ref[0] = t1[0].astype(np.float32)
        ref[1:] = t2.astype(np.float32)
        scale = np.sqrt(np.mean(ref ** 2, axis=1, keepdims=True))
        assert np.max(np.abs(got - ref) / scale) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("batch_size", [1200, 333])
def test_device_day_pipeline_chunks_change_nothing(tmp_path, monkeypatch, batch_size):
    """The device path's pipelined chunks (PIPE_CHUNK samples: two device
    buffers, uploads and downloads on a second stream) give the day bit for
    bit whatever the chunk size, ragged last chunks and several batches
    included, and equal the same day computed in one launch."""
    import torch

    from tod_generation_mapbased_beam_amd import kernels

    rng = np.random.default_rng(8)
    nside = 32
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    vec, bv, _ = inputs.gaussian_beam(64, 240.0, 0.99)
    bd = {"I": driver.beam_entry(vec, bv, [0], mp), "QU": driver.beam_entry(vec, bv, [1, 2], mp)}
    n = 1200
    th, ph, ps = inputs.ring_scan(n, 10.0)
    driver.write_scan_day(str(tmp_path), 0, th, ph, ps)
    days = {}
    for chunk in (1 << 19, 97, 256):
        monkeypatch.setattr(driver, "PIPE_CHUNK", chunk)
        days[chunk] = driver.tod_exact_gen_batched(bd, 0, mp, batch_size, folder_scan=str(tmp_path), log=None)
        assert days[chunk].shape == (3, n) and days[chunk].dtype == np.float64
    assert np.array_equal(days[97], days[1 << 19]) and np.array_equal(days[256], days[1 << 19])
    # one launch per entry over the whole day, as the unpipelined path ran it
    ctr = R.beam_centre(np.zeros((1, 1)), np.full((1, 1), math.pi / 2))
    params = kernels.rotation_params(np.array(th), np.array(ph), np.array(ps), ctr)
    ref = torch.zeros((3, n), dtype=torch.float64, device="cuda")
    for data in bd.values():
        for comp, vals in driver.beam_tod_batch(nside, mp, data, None, None, None, None,
                                                params=params).items():
            ref[comp] += vals
    assert np.array_equal(days[1 << 19], ref.cpu().numpy())
