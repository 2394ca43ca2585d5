"""GPU rotation preparation (todg_rotation_params) against the numpy path
that restates the reference's (tod_rotations.py:219-333 with the driver's
float32 casts, sample_based_tod_generation_gridint.py:277-281).

Bar (all operands lie in [-1, 1]): axes, cos_a, sin_a and ax_pts within
4 x 2^-24 of the numpy path's, cos_p and sin_p within 2^-21 + 2 x 2^-24, one
float32 ulp of the roll angle psi_b - beta (|angle| < 8) plus two of the
result: numpy's own float32 sin / cos are off the correctly rounded value by
one ulp on ~15 % of inputs, the GPU's are the binary64 functions rounded
once, and the binary64 beta that both round to float32 differs by a few
binary64 ulps.  The rotation angle is arccos(c . t) of the binary32
boresight t in both paths, so the angle, its axis and the parallactic offset
carry that rounding amplified by 1 / sin(angle) (pointings near the beam
centre's antipode): those bounds add 2 x 2^-24 / sin(angle).  Checked on a
day of ring scan, on pointings over the whole
sphere and for a beam grid centred off the pole; the TOD from the GPU
operands equals the oracle's on those operands to 1e-6, and differs from the
TOD of the numpy operands by no more than such operand changes move it.
"""

import math

import numpy as np
import pytest

import oracle as O
from cases import rel_err, small_config
from tod_generation_mapbased_beam_amd import inputs
from tod_generation_mapbased_beam_amd import tod_rotations as R

pytestmark = pytest.mark.gpu

NORTH = (np.zeros((1, 1)), np.full((1, 1), math.pi / 2))


def _ulps(a, b):
    """Distance in float32 units in the last place (values below 2^-24 in
    magnitude count as equal to zero)."""
    a = np.asarray(a, np.float32).ravel()
    b = np.asarray(b, np.float32).ravel()
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)  # monotone integer order
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    d = np.abs(ia - ib)
    tiny = (np.abs(a) < 2.0 ** -24) & (np.abs(b) < 2.0 ** -24)
    return np.where(tiny, 0, d)


def _both(th, ph, ps, ra=NORTH[0], dec=NORTH[1], center_idx=(0, 0)):
    from tod_generation_mapbased_beam_amd import kernels

    host = R.rotation_params_for_pointing(th, ph, ps, ra=ra, dec=dec, center_idx=center_idx)
    ctr = R.beam_centre(ra, dec, center_idx)
    dev = [p.cpu().numpy() for p in kernels.rotation_params(th, ph, ps, ctr)]
    return host, dev


NAMES = ["axes", "cos_a", "sin_a", "ax_pts", "cos_p", "sin_p"]
BASE = {"axes": 4 * 2.0 ** -24, "cos_a": 4 * 2.0 ** -24, "sin_a": 4 * 2.0 ** -24, "ax_pts": 4 * 2.0 ** -24,
        "cos_p": 2.0 ** -21 + 2 * 2.0 ** -24, "sin_p": 2.0 ** -21 + 2 * 2.0 ** -24}
CONDITIONED = ("axes", "cos_a", "sin_a", "cos_p", "sin_p")


def _bound(name, sin_a):
    """Per-sample bound (broadcast over the vector components)."""
    b = np.full(sin_a.shape, BASE[name])
    if name in CONDITIONED:
        b = b + 2 * 2.0 ** -24 / np.maximum(np.abs(sin_a.astype(np.float64)), 1e-6)
    return b


@pytest.mark.parametrize("kind", ["ring scan day", "whole sphere", "offset beam centre"])
def test_operands_match_the_numpy_path(kind):
    rng = np.random.default_rng(1)
    if kind == "ring scan day":
        th, ph, ps = inputs.ring_scan(1_728_000, 20.0)
        host, dev = _both(th, ph, ps)
    elif kind == "whole sphere":
        n = 1_000_000
        th = np.arccos(rng.uniform(-1, 1, n))
        ph = rng.uniform(0, 2 * math.pi, n)
        ps = rng.uniform(0, 2 * math.pi, n)
        host, dev = _both(th, ph, ps)
    else:  # a beam grid whose centre is off the pole (a pixelised beam map)
        n = 200_000
        th = np.arccos(rng.uniform(-0.9, 0.9, n))
        ph = rng.uniform(0, 2 * math.pi, n)
        ps = rng.uniform(0, 2 * math.pi, n)
        ra = np.linspace(-0.01, 0.01, 5)[None, :].repeat(5, 0) + 0.3
        dec = np.linspace(-0.01, 0.01, 5)[:, None].repeat(5, 1) + 0.2
        host, dev = _both(th, ph, ps, ra=ra, dec=dec, center_idx=None)
    bad = {}
    sin_a = np.asarray(host[2])
    for name, h, d in zip(NAMES, host, dev):
        u = _ulps(h, d)
        diff = np.abs(np.asarray(h, np.float64) - np.asarray(d, np.float64)).reshape(sin_a.shape[0], -1)
        err = float(diff.max())
        over = diff / _bound(name, sin_a)[:, None]
        print(f"{kind}: {name}: max |diff| {err:.2e} ({err / 2.0 ** -24:.1f} x 2^-24), max {int(u.max())} ulp, "
              f"{100 * float(np.mean(u > 0)):.1f}% differ, max diff / bound {float(over.max()):.2f}")
        if over.max() > 1.0:
            bad[name] = float(over.max())
    assert not bad, bad


def test_tod_parity_with_device_operands():
    from tod_generation_mapbased_beam_amd import kernels as K

    c = small_config(n_samples=4000)
    ctr = R.beam_centre(*NORTH, (0, 0))
    dev = K.rotation_params(c["theta"], c["phi"], c["psi"], ctr)
    devh = [p.cpu().numpy() for p in dev]
    got = np.zeros((3, c["B"]))
    K._gather_accum_fused_jit(c["vec"], *dev, c["nside"], c["map"], c["bv"], c["B"], c["S"], got, 1, 2)
    ref = np.zeros((3, c["B"]))
    O.gather_accum_fused(c["vec"], *devh, c["nside"], c["map"], c["bv"], c["B"], c["S"], ref, 1, 2)
    err = rel_err(got, ref)
    host = np.zeros((3, c["B"]))
    K._gather_accum_fused_jit(c["vec"], *c["params"], c["nside"], c["map"], c["bv"], c["B"], c["S"], host, 1, 2)
    shift = rel_err(got, host)
    print(f"device operands: TOD vs oracle {err:.2e}; vs the numpy operands' TOD {shift:.2e}")
    assert err <= 1e-6
    # one float32 ulp of a pointing angle (~6e-8 rad) moves a 60' beam's TOD
    # by ~1e-5 of its RMS at most on this smooth sky
    assert shift <= 1e-4
