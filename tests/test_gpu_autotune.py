"""The launch-shape autotuner on the GPU: single-detector variants (tile
shapes included) and batched (build, chunk) pairs are timed, the winner is
persisted to the cache and read back by beam_tod_batch's lookup."""

import json

import numpy as np
import pytest

from cases import small_config

pytestmark = pytest.mark.gpu


@pytest.fixture()
def cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("TODG_TUNE_CACHE", str(path))
    return path


def test_tune_single_persists_and_is_used(cache):
    from tod_generation_mapbased_beam_amd import autotune
    from tod_generation_mapbased_beam_amd.device import DeviceMap

    c = small_config()
    dmap = DeviceMap(c["map"], 1, 2)
    cand = autotune.single_candidates("bilinear", dmap)
    assert set(autotune.TILE_VARIANTS) - {20} <= set(cand)
    best, times = autotune.tune_single(c["vec"], c["bv"], c["params"], dmap, "bilinear", max_samples=4096,
                                       reps=2)
    assert set(times) == set(cand) and best == min(times, key=times.get)
    assert all(t > 0 for t in times.values())
    data = json.loads(cache.read_text())
    (key, entry), = data.items()
    assert entry["variant"] == best and "|single|" in key
    assert autotune.tuned_variant(c["nside"], c["S"], "bilinear", 3, True, default=-7) == best
    # a layout without tile variants only times the builds and the direct kernel
    assert autotune.single_candidates("bilinear", DeviceMap(c["map"], -1, -1)) == autotune.SINGLE_VARIANTS


def test_tune_batch_persists(cache):
    from tod_generation_mapbased_beam_amd import autotune
    from tod_generation_mapbased_beam_amd.batch import BatchPlan, Detector
    from tod_generation_mapbased_beam_amd.device import DeviceMap

    c = small_config()
    dmap = DeviceMap(c["map"], 1, 2)
    dets = [Detector(c["vec"], c["bv"], tuple(p[i * 2000:(i + 1) * 2000] for p in c["params"])) for i in range(4)]
    best, times = autotune.tune_batch(lambda v, ch: BatchPlan(dets, dmap, "bilinear", ch, v), reps=2,
                                      key_info=(c["nside"], c["S"], "bilinear", 3, True))
    assert set(times) == {(v, ch) for v in autotune.BATCH_VARIANTS for ch in autotune.BATCH_CHUNKS}
    assert best == min(times, key=times.get)
    entry = next(v for k, v in json.loads(cache.read_text()).items() if "|batch|" in k)
    assert (entry["variant"], entry["chunk"]) == best
