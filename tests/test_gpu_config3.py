"""Learned interpolation corrector training at its full size: an MLP 21 ->
64 -> 64 -> 16 (GELU) trained by this library's HIP kernels (BF16 MFMA
forward and backward, FP32 master weights, Adam lr 1e-3 with cosine decay,
8 epochs of batch 4096) on 1e7 (bilinear input, bicubic target) pairs at
uniformly random directions on 8 seeded nside-512 skies (spectrum of
inputs.cl_spectrum, lmax 3 nside - 1), held out: 1e6 directions on a new
seed.

Checks: the training loss falls every epoch after the first and the held-out
RMS residual against the bicubic judge is at most 0.05 of bilinear's (a run
of this recipe measured 0.021).
"""

import math

import pytest

pytestmark = pytest.mark.gpu


def test_full_training_run():
    from tod_generation_mapbased_beam_amd import learned_eval as L

    c, res = L.train(nside=512, n_maps=8, n_pairs=10_000_000, n_held=1_000_000, epochs=8, batch=4096,
                     lr=1e-3, seed=0)
    print(f"train pairs {res['train_pairs']}, held-out {res['held_out_pairs']}; {res['train_seconds']:.1f}s "
          f"training, {res['data_seconds']:.1f}s pair synthesis; loss per epoch "
          f"{[float(f'{v:.3g}') for v in res['loss_per_epoch']]}")
    print(f"held-out RMS residual vs bicubic: corrector / bilinear = {res['held_out_rms_ratio']:.4f} "
          f"(bilinear {math.sqrt(res['held_out_ms_residual_bilinear']):.4e}, corrector "
          f"{math.sqrt(res['held_out_ms_residual_corrector']):.4e})")
    # (directions whose 4 x 4 stencil is incomplete, next to the poles, are dropped)
    assert res["train_pairs"] >= 0.9999 * 10_000_000 and res["held_out_pairs"] >= 0.9999 * 1_000_000
    hist = res["loss_per_epoch"]
    assert all(b < a for a, b in zip(hist[1:], hist[2:]))
    assert res["held_out_rms_ratio"] <= 0.05
