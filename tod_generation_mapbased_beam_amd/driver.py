"""Per-day TOD driver on one GPU: scan memmaps in, ``tod_day_<i>.npy`` out.

The counterpart of the reference driver's ``tod_exact_gen_batched`` and
``_process_day`` (sample_based_tod_generation_gridint.py:208-333) and of
``open_scan_day`` (tod_io.py:70-95).  Each batch follows the reference's
casts and calls (pointings to float32, ``precompute_rotation_vector_batch``,
``psis_b = -betas + psi_b``, one ``beam_tod_batch`` per beam entry, float32
results summed into a float64 day array).  The map stays packed in HBM for
the whole run (it replaces the reference's POSIX shared-memory copy,
sample_based_tod_generation_gridint.py:478-507), so a batch can be a whole
day.  The rotation operands are computed on the GPU from the uploaded
pointings (``prep="device"``, the default: ``todg_rotation_params``, within
one float32 ulp of the numpy path), in chunks whose memmap reads, uploads
and downloads overlap the previous chunk's kernels; with
``prep="host"`` they come from the numpy path, split over host threads, the
host preparing the next batch while the GPU works on the current one.  A failing day is reported and
the others go on (sample_based_tod_generation_gridint.py:319-333).
"""

from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .tod_core import beam_tod_batch
from .tod_rotations import _rotation_params, beam_centre, precompute_rotation_vector_batch


def open_scan_day(folder_scan: str, day_index: int):
    """(theta, phi, psi) of one day as read-only memory maps (tod_io.py:70-95)."""
    return tuple(np.load(os.path.join(folder_scan, f"{name}_{day_index}.npy"), mmap_mode="r")
                 for name in ("theta", "phi", "psi"))


def write_scan_day(folder_scan: str, day_index: int, theta, phi, psi) -> None:
    """Write one day of boresight pointings in the layout open_scan_day reads."""
    os.makedirs(folder_scan, exist_ok=True)
    for name, arr in (("theta", theta), ("phi", phi), ("psi", psi)):
        np.save(os.path.join(folder_scan, f"{name}_{day_index}.npy"), np.asarray(arr, dtype=np.float64))


def beam_entry(vec_orig, beam_vals, comp_indices, mp, ra=None, dec=None):
    """A beam_data entry (sample_based_tod_generation_gridint.py:148-156, 447-450)
    for beams prepared in memory: ``mp_stacked`` holds the entry's components
    of the (3, npix) map as float32."""
    comp = list(comp_indices)
    stacked = np.ascontiguousarray(np.stack([np.asarray(mp[c]) for c in comp]), np.float32)
    stacked.setflags(write=False)  # read-only: its packed device copy is made once and reused
    return {
        "vec_orig": np.ascontiguousarray(vec_orig, np.float32),
        "beam_vals": np.ascontiguousarray(beam_vals, np.float32),
        "comp_indices": comp,
        "mp_stacked": stacked,
        "ra": np.zeros((1, 1)) if ra is None else ra,
        "dec": np.full((1, 1), np.pi / 2) if dec is None else dec,
        "n_sel": int(np.asarray(beam_vals).shape[0]),
    }


def _host_batch(theta_mmap, phi_mmap, psi_mmap, bs, be, ra0, dec0):
    theta_b = np.array(theta_mmap[bs:be], dtype=np.float32)
    phi_b = np.array(phi_mmap[bs:be], dtype=np.float32)
    psi_b = np.array(psi_mmap[bs:be], dtype=np.float32)
    rot_vecs, betas = precompute_rotation_vector_batch(ra0, dec0, phi_b, theta_b)
    psis_b = -betas + psi_b
    return rot_vecs, phi_b, theta_b, psis_b


# Samples below which a batch's host preparation is not split over threads.
HOST_SPLIT_MIN = 1 << 16


def _host_params(theta_mmap, phi_mmap, psi_mmap, bs, be, ra0, dec0, pool=None, parts=1):
    """The six kernel operands of samples [bs, be) by the numpy path
    (_host_batch, then _rotation_params), the range split into ``parts``
    pieces computed by ``pool``'s threads (numpy's loops release the GIL).
    Every operation is per sample, so the split changes no value."""

    def one(a, b):
        return _rotation_params(*_host_batch(theta_mmap, phi_mmap, psi_mmap, a, b, ra0, dec0))

    n = be - bs
    parts = max(1, min(int(parts), n // HOST_SPLIT_MIN))
    if pool is None or parts == 1:
        return one(bs, be)
    edges = [bs + (n * i) // parts for i in range(parts + 1)]
    pieces = list(pool.map(lambda i: one(edges[i], edges[i + 1]), range(parts)))
    return tuple(np.concatenate([p[j] for p in pieces]) for j in range(6))


# Samples per pipelined chunk of the device path (see _device_day).
PIPE_CHUNK = 1 << 19


def _device_day(beam_data, nside, mp, theta_mmap, phi_mmap, psi_mmap, n_samples, batch_size, ra0, dec0,
                interp_mode):
    """The device path of one day, pipelined over chunks of at most
    PIPE_CHUNK samples within each batch: while the GPU computes chunk k on
    the current stream, the host reads chunk k + 1 from the memmaps into
    pinned memory and a second stream (default priority) uploads it, and
    chunk k - 1's TOD goes back to pinned host memory on that stream.  Two
    device buffers alternate; events order each buffer's reuse.  Every
    sample's TOD is computed on its own, so the chunking changes no value."""
    import torch

    from . import kernels
    from .device import device, to_dev

    if n_samples == 0:
        return np.zeros((3, 0))
    dev = device()
    ctr = beam_centre(ra0, dec0)
    # the beam arrays on the device once (a pageable upload inside the loop
    # would wait for the compute stream to drain)
    beams = [{**d, "vec_orig": to_dev(d["vec_orig"], torch.float32),
              "beam_vals": to_dev(d["beam_vals"], torch.float32)} for d in beam_data.values()]
    spans = [(cs, min(cs + PIPE_CHUNK, bs + batch_size, n_samples))
             for bs in range(0, n_samples, batch_size)
             for cs in range(bs, min(bs + batch_size, n_samples), PIPE_CHUNK)]
    width = max(ce - cs for cs, ce in spans)
    compute = torch.cuda.current_stream(dev)
    copy = torch.cuda.Stream(dev)
    out = torch.empty((3, n_samples), dtype=torch.float64, pin_memory=True)
    stage = [torch.empty((3, width), dtype=torch.float64, pin_memory=True) for _ in range(2)]
    pts = [torch.empty((3, width), dtype=torch.float64, device=dev) for _ in range(2)]
    tods = [torch.empty((3, width), dtype=torch.float64, device=dev) for _ in range(2)]
    uploaded, consumed, downloaded = [None, None], [None, None], [None, None]

    def upload(k):
        cs, ce = spans[k]
        i, n = k % 2, ce - cs
        if uploaded[i] is not None:
            uploaded[i].synchronize()  # the staging buffer's last upload is done
        host = stage[i].numpy()
        for row, m in enumerate((theta_mmap, phi_mmap, psi_mmap)):
            np.copyto(host[row, :n], m[cs:ce], casting="same_kind")
        with torch.cuda.stream(copy):
            if consumed[i] is not None:
                copy.wait_event(consumed[i])  # chunk k - 2 has read pts[i]
            for row in range(3):
                pts[i][row, :n].copy_(stage[i][row, :n], non_blocking=True)
            uploaded[i] = torch.cuda.Event()
            uploaded[i].record(copy)

    def run(k):
        cs, ce = spans[k]
        i, n = k % 2, ce - cs
        compute.wait_event(uploaded[i])
        if downloaded[i] is not None:
            compute.wait_event(downloaded[i])  # chunk k - 2's TOD has left tods[i]
        tod = tods[i][:, :n]
        tod.zero_()
        p = pts[i]
        params = kernels.rotation_params(p[0, :n], p[1, :n], p[2, :n], ctr)
        for data in beams:
            contrib = beam_tod_batch(nside, mp, data, None, None, None, None, interp_mode=interp_mode,
                                     params=params)
            for comp, vals in contrib.items():
                tod[comp] += vals
        consumed[i] = torch.cuda.Event()
        consumed[i].record(compute)
        with torch.cuda.stream(copy):
            copy.wait_event(consumed[i])
            for row in range(3):
                out[row, cs:ce].copy_(tod[row], non_blocking=True)
            downloaded[i] = torch.cuda.Event()
            downloaded[i].record(copy)

    upload(0)
    for k in range(len(spans)):
        if k + 1 < len(spans):
            upload(k + 1)
        run(k)
    copy.synchronize()
    compute.synchronize()
    return out.numpy()


def tod_exact_gen_batched(beam_data, day_index, mp, batch_size, process_name=None, folder_scan=".",
                          interp_mode="bilinear", log=print, prep="device"):
    """TOD of one observation day, shape (3, n_samples) float64 (axis 0 =
    I, Q, U), in ceil(n_samples / batch_size) batches; ``prep`` = where the
    rotation operands are computed ("device" or "host")."""
    if prep not in ("device", "host"):
        raise ValueError("prep must be 'device' or 'host'")
    prefix = f"[{process_name}] " if process_name else ""
    npix = int(np.asarray(mp[0]).shape[0])
    nside = int(round((npix / 12) ** 0.5))
    theta_mmap, phi_mmap, psi_mmap = open_scan_day(folder_scan, day_index)
    n_samples = len(phi_mmap)
    first = next(iter(beam_data.values()))
    ra0, dec0 = first["ra"], first["dec"]
    batch_size = max(1, min(int(batch_size), n_samples))
    n_batches = (n_samples + batch_size - 1) // batch_size
    if log:
        log(prefix + f"Day {day_index} - {n_samples} samples, batch_size={batch_size}, n_batches={n_batches}")
    t0 = time.time()
    if prep == "device":
        tod_day = _device_day(beam_data, nside, mp, theta_mmap, phi_mmap, psi_mmap, n_samples, batch_size,
                              ra0, dec0, interp_mode)
        if log:
            total = time.time() - t0
            log(prefix + f"Done - {n_samples} samples in {total:.2f}s ({total / n_batches:.3f}s/batch)")
        return tod_day
    tod_day = np.zeros((3, n_samples))
    workers = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=1) as ahead, ThreadPoolExecutor(max_workers=workers) as pool:
        nxt = ahead.submit(_host_params, theta_mmap, phi_mmap, psi_mmap, 0, min(batch_size, n_samples),
                           ra0, dec0, pool, workers)
        for k in range(n_batches):
            bs = k * batch_size
            be = min(bs + batch_size, n_samples)
            params = nxt.result()
            if k + 1 < n_batches:  # the host prepares batch k+1 while the GPU runs batch k
                nb = min(be + batch_size, n_samples)
                nxt = ahead.submit(_host_params, theta_mmap, phi_mmap, psi_mmap, be, nb, ra0, dec0, pool,
                                   workers)
            for data in beam_data.values():
                contrib = beam_tod_batch(nside, mp, data, None, None, None, None, interp_mode=interp_mode,
                                         params=params)
                for comp, vals in contrib.items():
                    tod_day[comp, bs:be] += vals.cpu().numpy()
    if log:
        total = time.time() - t0
        log(prefix + f"Done - {n_samples} samples in {total:.2f}s ({total / n_batches:.3f}s/batch)")
    return tod_day


def run_days(beam_data, mp, days, batch_size, folder_scan, folder_out, interp_mode="bilinear",
             log=print, prep="device"):
    """Generate and save tod_day_<i>.npy for every day; a day that fails is
    reported as (day, False, message) and the others go on.  Returns the list
    of (day, ok, error) tuples."""
    os.makedirs(folder_out, exist_ok=True)
    results = []
    for d in days:
        try:
            tod = tod_exact_gen_batched(beam_data, d, mp, batch_size, folder_scan=folder_scan,
                                        interp_mode=interp_mode, log=log, prep=prep)
            np.save(os.path.join(folder_out, f"tod_day_{d}.npy"), tod)
            results.append((d, True, None))
        except Exception as e:  # failure isolation per day, as the reference's worker
            if log:
                log(f"Error on day {d}: {e}")
            results.append((d, False, str(e)))
    return results


__all__ = ["open_scan_day", "write_scan_day", "beam_entry", "tod_exact_gen_batched", "run_days"]
