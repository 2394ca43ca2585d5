"""The allocation-free B_ell entry (todg_beam_bell_noalloc) against the
workspace entry (todg_beam_bell), through the C ABI: the same kernels, so the
results must agree bit for bit; two calls in a row, the second smaller (the
shared scratch keeps the first call's partial sums, which the second must
not see), then a size beyond the scratch, which must be refused."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bell_pair(N, lmax, seed):
    import torch

    from tod_generation_mapbased_beam_amd import _lib
    from tod_generation_mapbased_beam_amd.device import stream_handle

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda", 0)
    x = torch.as_tensor(rng.uniform(-1.0, 1.0, N), device=dev)
    w = torch.as_tensor(rng.uniform(0.0, 1.0, N) / N, device=dev)
    ws = torch.empty(int(_lib.lib().todg_beam_bell_workspace(N, lmax)), dtype=torch.float64, device=dev)
    ref = torch.full((lmax + 1,), np.nan, dtype=torch.float64, device=dev)
    got = torch.full((lmax + 1,), np.nan, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().todg_beam_bell(_lib.ptr(x), _lib.ptr(w), N, lmax, _lib.ptr(ws), _lib.ptr(ref),
                                         stream_handle()))
    _lib.check(_lib.lib().todg_beam_bell_noalloc(_lib.ptr(x), _lib.ptr(w), N, lmax, _lib.ptr(got),
                                                 stream_handle()))
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy(), float(w.sum().item())


def test_noalloc_matches_workspace_entry_twice():
    for N, lmax, seed in ((40000, 1500, 1), (3000, 300, 2), (0, 50, 3)):
        ref, got, wsum = _bell_pair(N, lmax, seed)
        assert np.all(np.isfinite(got))
        assert np.array_equal(ref, got), f"N={N} lmax={lmax}"
        assert abs(got[0] - wsum) <= 1e-12 * max(wsum, 1.0)  # B_0 = sum of the weights (P_0 = 1)


def test_noalloc_refuses_sizes_beyond_its_scratch():
    import torch

    from tod_generation_mapbased_beam_amd import _lib
    from tod_generation_mapbased_beam_amd.device import stream_handle

    dev = torch.device("cuda", 0)
    N, lmax = 1 << 20, 1000  # 4096 blocks x 1001 multipoles > 2^20 doubles
    x = torch.zeros(N, dtype=torch.float64, device=dev)
    out = torch.zeros(lmax + 1, dtype=torch.float64, device=dev)
    rc = _lib.lib().todg_beam_bell_noalloc(_lib.ptr(x), _lib.ptr(x), N, lmax, _lib.ptr(out),
                                           stream_handle())
    assert rc == _lib.TODG_ERR_ARG
    assert "scratch" in _lib.last_error()
