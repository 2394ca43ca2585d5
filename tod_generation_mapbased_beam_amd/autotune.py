"""Launch-shape autotuner with a persistent cache.

The GPU counterpart of the reference's empirical runtime calibration
(tod_calibrate.py:198-364, which sweeps processes x threads x batch size and
writes the winner back to the config, sample_based_tod_generation_gridint.py:
339-362): here the knobs are the kernel's launch and tile shape -- the
footprint kernel compiled for 2 or 4 waves per SIMD, its tile variants
(beam pixels per phase-1 claim 128/256/512 x 64-cell chunks per lane in the
footprint sweep 3/4/5), or the direct kernel -- and, for batched launches,
the build and the number of samples per work-queue claim.  Candidates are
timed with HIP events on a bounded prefix of the real workload; the winner is
stored in a JSON cache keyed by device and workload shape and picked up by
beam_tod_batch and bench.py.
"""

from __future__ import annotations

import json
import os
import threading

import torch

CACHE_ENV = "TODG_TUNE_CACHE"
SINGLE_VARIANTS = (0, 8, 2)
# footprint-kernel tile variants (include/todgpu.h): 16 + 3 i + j, claims of
# {128, 256, 512}[i] beam pixels, {3, 4, 5}[j] sweep chunks; built for I/Q/U
# with spin-2 and for one field
TILE_VARIANTS = tuple(range(16, 25))
BATCH_VARIANTS = (0, 8)
BATCH_CHUNKS = (1, 4, 16, 64)

_mu = threading.Lock()


def cache_path() -> str:
    p = os.environ.get(CACHE_ENV)
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "tod_generation_mapbased_beam_amd",
                        "tune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, value: dict) -> None:
    with _mu:
        data = _load()
        data[key] = value
        path = cache_path()
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def _bucket(n: int) -> int:
    """Power-of-two bucket of a size, so nearby workloads share a choice."""
    b = 1
    while b < n:
        b <<= 1
    return b


def workload_key(kind: str, nside: int, S: int, mode: str, C: int, qu: bool) -> str:
    dev = torch.cuda.get_device_name() if torch.cuda.is_available() else "cpu"
    return f"{dev}|{kind}|nside={nside}|S~{_bucket(int(S))}|{mode}|C={C}|qu={int(bool(qu))}"


def lookup(key: str):
    return _load().get(key)


def _time_ms(fn, reps: int) -> float:
    fn()  # warm-up (first launch also builds the ring table)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    ts.sort()
    return ts[len(ts) // 2]


def single_candidates(mode: str, dmap):
    """Variants worth timing for a workload: the 2/4-waves builds and the
    direct kernel, plus the tile variants where they are built."""
    qu = dmap.c_q >= 0 and dmap.c_u >= 0
    cand = list(SINGLE_VARIANTS)
    if mode in ("bilinear", "nearest") and ((qu and dmap.cs == 4) or (not qu and dmap.cs == 1)):
        cand += [v for v in TILE_VARIANTS if v != 20]  # (20 is variant 0's tile)
    return tuple(cand)


def tune_single(vec_orig, beam_vals, params, dmap, mode: str = "bilinear",
                variants=None, max_samples: int = 16384, reps: int = 3,
                persist: bool = True):
    """Time each kernel variant on the first ``max_samples`` samples of a
    single-detector workload; returns (best_variant, {variant: ms})."""
    from . import kernels
    from .device import to_dev

    if variants is None:
        variants = single_candidates(mode, dmap)

    f32 = torch.float32
    B = min(int(params[1].shape[0]), int(max_samples))
    ps = [to_dev(p, f32)[:B].contiguous() for p in params]
    v = to_dev(vec_orig, f32)
    bv = to_dev(beam_vals, f32)
    S = int(bv.numel())
    tod = torch.zeros((dmap.C, B), dtype=torch.float64, device=v.device)
    fn = kernels._gather_accum_fused_jit if mode == "bilinear" else kernels._gather_accum_nearest_jit
    times = {}
    for var in variants:
        times[var] = _time_ms(lambda: fn(v, *ps, dmap.nside, dmap, bv, B, S, tod, dmap.c_q,
                                         dmap.c_u, variant=var), reps)
    best = min(times, key=times.get)
    if persist:
        qu = dmap.c_q >= 0 and dmap.c_u >= 0
        _store(workload_key("single", dmap.nside, S, mode, dmap.C, qu),
               {"variant": int(best), "ms": {str(k): t for k, t in times.items()}, "samples": B})
    return best, times


def tune_batch(make_plan, variants=BATCH_VARIANTS, chunks=BATCH_CHUNKS, reps: int = 3,
               persist: bool = True, key_info=None):
    """Time (variant, chunk) pairs for a batched launch.  ``make_plan(variant,
    chunk)`` returns a BatchPlan; returns ((variant, chunk), {pair: ms})."""
    times = {}
    for var in variants:
        for ch in chunks:
            plan = make_plan(var, ch)
            times[(var, ch)] = _time_ms(plan.launch, reps)
    best = min(times, key=times.get)
    if persist and key_info is not None:
        _store(workload_key("batch", *key_info),
               {"variant": int(best[0]), "chunk": int(best[1]),
                "ms": {f"{k[0]}/{k[1]}": t for k, t in times.items()}})
    return best, times


def tuned_variant(nside: int, S: int, mode: str, C: int, qu: bool, default: int = 0) -> int:
    hit = lookup(workload_key("single", nside, S, mode, C, qu))
    return int(hit["variant"]) if hit else default
