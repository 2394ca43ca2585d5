"""Many detectors in one persistent launch (``todg_gather_accum_batch``).

Each item is one detector's beam (uneven kept-pixel counts after the power
cut and tail clustering) and one scan chunk of its pointings (uneven
lengths).  All items share one packed map.  The plan uploads a device array
of ``todg_work_item`` structs once; ``launch`` enqueues the whole batch on
the current stream, and ``capture`` records it into a HIP graph (the work
queue is cleared by a kernel inside the launch, so replays are valid).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import DeviceMap, device, stream_handle, to_dev

# Field order and widths of todg_work_item (include/todgpu.h).
ITEM_DTYPE = np.dtype([
    ("vec_orig", "<u8"), ("beam_vals", "<u8"), ("axes", "<u8"), ("cos_a", "<u8"),
    ("sin_a", "<u8"), ("ax_pts", "<u8"), ("cos_p", "<u8"), ("sin_p", "<u8"),
    ("tod", "<u8"), ("B", "<i8"), ("S", "<i8"), ("first", "<i8"),
])


@dataclass
class Detector:
    """One work item: a beam and a chunk of per-sample rotation operands
    (the six float32 arrays of ``_rotation_params``)."""

    vec_orig: object
    beam_vals: object
    params: tuple
    tod: torch.Tensor | None = None


class BatchPlan:
    def __init__(self, detectors, dmap: DeviceMap, mode: str = "bilinear", chunk: int = 8,
                 variant: int = 0):
        modes = {"bilinear": 0, "nearest": 1, "bicubic": 2}
        if mode not in modes:
            raise ValueError(f"interp_mode {mode!r} not available; use 'bilinear', 'nearest' "
                             "or 'bicubic'")
        self.mode = modes[mode]
        self.dmap = dmap
        self.chunk = int(chunk)
        self.variant = int(variant)
        from .kernels import check_tod

        for i, d in enumerate(detectors):  # before anything is uploaded
            if d.tod is not None:
                if not isinstance(d.tod, torch.Tensor):
                    raise ValueError(f"detector {i}: tod must be a CUDA tensor (or None)")
                check_tod(d.tod, dmap.C, int(d.params[1].shape[0]), f"detector {i} tod")
        dev = device()
        f32 = torch.float32
        self._keep = []  # device tensors referenced by the item array
        rows = np.zeros(len(detectors), dtype=ITEM_DTYPE)
        first = 0
        self.tods = []
        for i, d in enumerate(detectors):
            v = to_dev(d.vec_orig, f32)
            bv = to_dev(d.beam_vals, f32)
            ps = [to_dev(p, f32) for p in d.params]
            B = int(ps[1].numel())
            S = int(bv.numel())
            if S > (1 << 28):
                raise ValueError(f"detector {i}: at most 2^28 beam pixels")
            tod = d.tod if d.tod is not None else torch.zeros((dmap.C, B), dtype=torch.float64, device=dev)
            self._keep += [v, bv, *ps, tod]
            self.tods.append(tod)
            rows[i] = (v.data_ptr(), bv.data_ptr(), *(p.data_ptr() for p in ps), tod.data_ptr(),
                       B, S, first)
            first += B
        self.total = first
        self.n_items = len(detectors)
        self.items = torch.from_numpy(rows.view(np.uint8).copy()).to(dev)
        self.queue = torch.zeros(8, dtype=torch.int32, device=dev)
        self.status = torch.zeros(max(1, self.n_items), dtype=torch.int32, device=dev)
        self.work = float(sum(int(t.shape[1]) * int(k.numel()) for t, k in
                              zip(self.tods, self._keep[1::9])))

    def launch(self):
        """Enqueue the whole batch (accumulates into each detector's tod)."""
        _lib.check(_lib.lib().todg_gather_accum_batch(
            self.mode, _lib.ptr(self.items), self.n_items, self.total, self.dmap.nside,
            _lib.ptr(self.dmap.packed), self.dmap.C, self.dmap.c_q, self.dmap.c_u, self.chunk,
            _lib.ptr(self.queue), _lib.ptr(self.status), self.variant, stream_handle()))

    def zero(self):
        for t in self.tods:
            t.zero_()

    def capture(self):
        """Record zero + launch into a HIP graph; returns the torch graph.
        Call ``launch`` once first so one-time setup (ring table) happens
        outside the capture."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.zero()
                self.launch()
        torch.cuda.current_stream().wait_stream(s)
        return g

    def failed_items(self):
        """Indices of items whose pointings held non-finite values."""
        return np.nonzero(self.status.cpu().numpy()[: self.n_items])[0].tolist()


def capture_plans(plans):
    """Record zero + launch of several plans (e.g. one per Stokes component)
    into one HIP graph, in order; returns the torch graph.  Launch each plan
    once first so that one-time setup happens outside the capture."""
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            for p in plans:
                p.zero()
                p.launch()
    torch.cuda.current_stream().wait_stream(s)
    return g
