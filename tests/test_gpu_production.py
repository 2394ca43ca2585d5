"""The production single-detector configuration at full size on the GPU.

nside 2048 I/Q/U map, 30' Gaussian beam at nside_beam 4096 with the 0.99
power cut (S = 6364), 1,728,000 samples of a 20 Hz ring scan -- the
throughput benchmark's workload.  The oracle cannot sweep a whole day in a
test, so the full-size launch is checked through a spot check of random
samples against the oracle and through properties that do not depend on the
size: linearity in the map, bit-for-bit determinism, and a constant map; in
both modes (bilinear and nearest).  Neighbour indices at this size are
checked in test_gpu_fastpath_indices.py.
"""

import numpy as np
import pytest

import oracle as O
from cases import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-6
NSIDE = 2048
B = 1_728_000


@pytest.fixture(scope="module")
def prod():
    import torch

    from tod_generation_mapbased_beam_amd import inputs
    from tod_generation_mapbased_beam_amd import tod_rotations as R
    from tod_generation_mapbased_beam_amd.device import DeviceMap, to_dev

    dev = torch.device("cuda", 0)
    vec, bv, _ = inputs.gaussian_beam(4096, 30.0, 0.99)
    th, ph, ps = inputs.ring_scan(B, 20.0)
    params = R.rotation_params_for_pointing(th, ph, ps)
    m1 = inputs.white_map(NSIDE, seed=11, fields=3, device=dev)
    m2 = inputs.white_map(NSIDE, seed=12, fields=3, device=dev)
    f32 = torch.float32
    return {
        "torch": torch, "dev": dev, "vec": vec, "bv": bv, "params": params, "m1": m1, "m2": m2,
        "dvec": to_dev(vec, f32), "dbv": to_dev(bv, f32), "dparams": [to_dev(p, f32) for p in params],
        "DeviceMap": DeviceMap, "S": int(vec.shape[0]),
    }


def _gpu(P, mp, c_q=1, c_u=2, mode="bilinear"):
    from tod_generation_mapbased_beam_amd import kernels

    torch = P["torch"]
    tod = torch.zeros((mp.shape[0], B), dtype=torch.float64, device=P["dev"])
    fn = kernels._gather_accum_fused_jit if mode == "bilinear" else kernels._gather_accum_nearest_jit
    fn(P["dvec"], *P["dparams"], NSIDE, P["DeviceMap"](mp, c_q, c_u), P["dbv"], B, P["S"], tod, c_q, c_u)
    torch.cuda.synchronize()
    return tod


def test_production_shape(prod):
    assert prod["S"] == 6364


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_production_spot_check_vs_oracle(prod, mode):
    """1000 random samples of the full-day launch against the oracle on the
    same samples."""
    g = _gpu(prod, prod["m1"], mode=mode).cpu().numpy()
    idx = np.sort(np.random.default_rng(5).choice(B, 1000, replace=False))
    sub = tuple(np.ascontiguousarray(p[idx]) for p in prod["params"])
    mp = prod["m1"].cpu().numpy()
    r = np.zeros((3, idx.size))
    fn = O.gather_accum_fused if mode == "bilinear" else O.gather_accum_nearest
    fn(prod["vec"], *sub, NSIDE, mp, prod["bv"], idx.size, prod["S"], r, 1, 2)
    assert rel_err(g[:, idx], r) <= TOL


def test_production_linear_in_the_map(prod):
    a = 0.37
    m = a * prod["m1"] + prod["m2"]
    t = _gpu(prod, m)
    t1 = _gpu(prod, prod["m1"])
    t2 = _gpu(prod, prod["m2"])
    assert rel_err((a * t1 + t2).cpu().numpy(), t.cpu().numpy()) <= TOL


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_production_deterministic(prod, mode):
    """Two launches agree bit for bit (the reference is deterministic and
    thread-count independent, tests/test_tod_bilinear.py:830-870): every
    weight-grid deposit is rounded onto the beam's grid, on which binary64
    cell sums are exact whatever the order of the LDS atomics, and pixels
    outside a sample's layout are summed in pixel order."""
    t1 = _gpu(prod, prod["m2"], mode=mode).cpu().numpy()
    t2 = _gpu(prod, prod["m2"], mode=mode).cpu().numpy()
    n_diff = int(np.count_nonzero(t1 != t2))
    print(f"{mode}: {n_diff} of {t1.size} entries differ between two launches")
    assert n_diff == 0


def test_production_constant_map(prod):
    """An I-only map of constant value c gives c * sum(beam) for every sample
    (the interpolation weights of each beam pixel sum to one)."""
    torch = prod["torch"]
    c = 2.5
    mp = torch.full((1, 12 * NSIDE * NSIDE), c, dtype=torch.float32, device=prod["dev"])
    t = _gpu(prod, mp, -1, -1).cpu().numpy()
    want = c * float(np.sum(prod["bv"].astype(np.float64)))
    assert np.max(np.abs(t - want)) <= TOL * abs(want)
