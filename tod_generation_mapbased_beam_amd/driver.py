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
one float32 ulp of the numpy path) and the day's TOD is summed in HBM; with
``prep="host"`` they come from the numpy path, the host preparing the next
batch while the GPU works on the current one.  A failing day is reported and
the others go on (sample_based_tod_generation_gridint.py:319-333).
"""

from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .tod_core import beam_tod_batch
from .tod_rotations import beam_centre, precompute_rotation_vector_batch


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


def _device_day(beam_data, nside, mp, theta_mmap, phi_mmap, psi_mmap, n_samples, batch_size, ra0, dec0,
                interp_mode):
    import torch

    from . import kernels
    from .device import device

    dev = device()
    ctr = beam_centre(ra0, dec0)
    tod_day = torch.zeros((3, n_samples), dtype=torch.float64, device=dev)
    for bs in range(0, n_samples, batch_size):
        be = min(bs + batch_size, n_samples)
        pts = [torch.from_numpy(np.array(m[bs:be], dtype=np.float64)).to(dev)  # (a writeable copy)
               for m in (theta_mmap, phi_mmap, psi_mmap)]
        params = kernels.rotation_params(*pts, ctr)
        for data in beam_data.values():
            contrib = beam_tod_batch(nside, mp, data, None, None, None, None, interp_mode=interp_mode,
                                     params=params)
            for comp, vals in contrib.items():
                tod_day[comp, bs:be] += vals
    return tod_day.cpu().numpy()


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
    with ThreadPoolExecutor(max_workers=1) as pool:
        nxt = pool.submit(_host_batch, theta_mmap, phi_mmap, psi_mmap, 0, min(batch_size, n_samples),
                          ra0, dec0)
        for k in range(n_batches):
            bs = k * batch_size
            be = min(bs + batch_size, n_samples)
            rot_vecs, phi_b, theta_b, psis_b = nxt.result()
            if k + 1 < n_batches:  # the host prepares batch k+1 while the GPU runs batch k
                nb = min(be + batch_size, n_samples)
                nxt = pool.submit(_host_batch, theta_mmap, phi_mmap, psi_mmap, be, nb, ra0, dec0)
            for data in beam_data.values():
                contrib = beam_tod_batch(nside, mp, data, rot_vecs, phi_b, theta_b, psis_b,
                                         interp_mode=interp_mode)
                for comp, vals in contrib.items():
                    tod_day[comp, bs:be] += vals
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
