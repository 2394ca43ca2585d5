"""Spherical-harmonic synthesis of the synthetic skies on the GPU: the ring
sums of todg_ring_sums against the same sums as torch tensor operations on
the same device (the kernel performs the same binary64 operations in the
same order), and whole maps against the host synthesis.  The host synthesis
itself is pinned in test_host.py (point evaluation of the a_lm series)."""

import numpy as np
import pytest
import torch

from tod_generation_mapbased_beam_amd import healpix_np as hpx
from tod_generation_mapbased_beam_amd import inputs

pytestmark = pytest.mark.gpu


def _alms(nside, K, seed):
    lmax = 3 * nside - 1
    rng = np.random.default_rng(seed)
    tt, _, _ = inputs.cl_spectrum(lmax)
    return lmax, [inputs._alm_realisation(tt, lmax, rng) for _ in range(K)]


@pytest.mark.parametrize("nside,K", [(1, 1), (4, 3), (16, 2), (64, 4), (256, 5)])
def test_ring_sums_match_torch(nside, K):
    """K = 5 takes two launches (at most four fields each); nside 256 reaches
    m = 767 on rings within 0.2 deg of the poles, where lambda_mm underflows
    binary64 without the rescaled mantissa."""
    lmax, alms = _alms(nside, K, seed=nside)
    _, _, _, z = hpx.ring_layout(nside)
    th = np.arccos(z)
    got = inputs._ring_sums_gpu(alms, lmax, th, torch.device("cuda"))
    ref = inputs._ring_sums_torch(alms, lmax, th, "cuda")
    assert len(got) == K
    for g, r in zip(got, ref):
        assert tuple(g.shape) == (z.size, lmax + 1)
        d = (g - r).abs().max().item()
        s = r.abs().max().item()
        print(f"nside {nside}: max |gpu - torch| {d:.3e} of max {s:.3e}")
        assert d <= 1e-13 * s


def test_synthesized_map_matches_host():
    nside = 32
    lmax, alms = _alms(nside, 3, seed=11)
    got = inputs.synthesize_map(nside, alms, lmax, device="cuda")
    ref = inputs.synthesize_map(nside, alms, lmax)
    err = np.max(np.abs(got - ref)) / np.sqrt(np.mean(ref ** 2))
    print(f"map max err / rms {err:.3e}")
    assert err < 1e-12
