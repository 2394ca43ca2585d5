"""MI355X-native (gfx950) beam-convolution TOD generation.

Drop-in GPU replacement for the hot path of a map-based beam TOD generator:
``tod[c, t] = sum_s beam_vals[s] * spin2_c(map[R(theta_t, phi_t, psi_t) . beam_pixel[s]])``
with bilinear or nearest HEALPix RING interpolation.  Public entry points
mirror the reference's names and signatures:

* :func:`beam_tod_batch` (tod_core.py:52)
* :func:`_gather_accum_fused_jit` (tod_bilinear.py:205)
* :func:`_gather_accum_nearest_jit` (tod_nearest.py:21)
* :func:`precompute_rotation_vector_batch`, :func:`_rotation_params`
  (tod_rotations.py:219, :306)

All computation runs in ``libtodgpu.so`` (hand-written HIP for gfx950, C ABI
in ``include/todgpu.h``); torch supplies device memory and streams.
"""

from .tod_rotations import precompute_rotation_vector_batch, _rotation_params  # noqa: F401
from .tod_core import beam_tod_batch  # noqa: F401
from .kernels import _gather_accum_fused_jit, _gather_accum_nearest_jit  # noqa: F401
from .device import DeviceMap  # noqa: F401

__all__ = [
    "beam_tod_batch",
    "_gather_accum_fused_jit",
    "_gather_accum_nearest_jit",
    "precompute_rotation_vector_batch",
    "_rotation_params",
    "DeviceMap",
]
