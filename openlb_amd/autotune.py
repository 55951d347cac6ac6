"""Launch-shape autotuner for the step kernels.

For a lattice it sweeps threads per block and an occupancy cap (waves per
SIMD, enforced by an LDS reservation); for an ensemble also the number of
tiles each persistent block walks.  Every candidate is timed with HIP events
(olbgfx_time_steps / olbgfx_ensemble_time) after one untimed step, and the
fastest is applied.  Results are cached in a JSON file keyed by shape,
storage and dynamics mix, so a later run with the same key skips the sweep.

The trial steps advance the lattice: tune during warm-up.
"""
from __future__ import annotations

import hashlib
import json
import os

BLOCKS = (64, 128, 256)
ORDERS = (-1, 0, 2, 4, 6)  # block orders: contiguous per XCD, runs of 2^k blocks
WAVES = (0, 2, 3, 4, 5, 6)
TILES = (1, 2, 4, 8)


def _cache_path() -> str:
    return os.environ.get("OLBGFX_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "olbgfx", "autotune.json"))


def _load() -> dict:
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(key: str, value: dict) -> None:
    path = _cache_path()
    data = _load()
    data[key] = value
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + ".tmp", "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(path + ".tmp", path)
    except OSError:
        pass


def lattice_key(lat, mask_digest: str = "") -> str:
    return f"lattice:{lat.extent}:{lat.padding}:{lat.periodic}:{lat.storage}:{mask_digest}"


def tune_lattice(lat, steps: int = 4, stream=None, key: str | None = None, use_cache: bool = True,
                 blocks=BLOCKS, waves=WAVES, orders=ORDERS) -> dict:
    """pick (block_threads, waves_per_simd), then the block order, for one
    lattice and apply them"""
    key = key or lattice_key(lat)
    if use_cache:
        hit = _load().get(key)
        if hit:
            lat.set_launch(hit["block"], hit["waves"])
            lat.set_block_order(hit.get("order", -1))
            return dict(hit, cached=True)
    trials = []
    lat.set_block_order(-1)
    for b in blocks:
        for w in waves:
            lat.set_launch(b, w)
            lat.collide_and_stream(1, stream)
            ms = lat.time_steps(steps, stream) / steps
            trials.append({"block": b, "waves": w, "order": -1, "ms": ms})
    best = min(trials, key=lambda t: t["ms"])
    lat.set_launch(best["block"], best["waves"])
    for o in orders:
        if o == -1:
            continue
        lat.set_block_order(o)
        lat.collide_and_stream(1, stream)
        ms = lat.time_steps(steps, stream) / steps
        trials.append({"block": best["block"], "waves": best["waves"], "order": o, "ms": ms})
    best = min(trials, key=lambda t: t["ms"])
    lat.set_launch(best["block"], best["waves"])
    lat.set_block_order(best["order"])
    pick = {k: best[k] for k in ("block", "waves", "order", "ms")}
    _store(key, pick)
    return dict(pick, trials=trials)


def ensemble_key(extents, storage: str) -> str:
    h = hashlib.sha1(repr(sorted(extents)).encode()).hexdigest()[:12]
    return f"ensemble:{len(extents)}:{h}:{storage}"


def tune_ensemble(ens, steps: int = 4, stream=None, key: str | None = None, use_cache: bool = True,
                  blocks=(128, 256), waves=(0, 3, 4, 5), tiles=(1, 2, 4),
                  group_bytes=(0, 96 << 20, 160 << 20), steps_per_group=(1, 4, 8),
                  tickets=(False, True)) -> dict:
    """pick (block_threads, waves_per_simd, tiles_per_block), then the group
    schedule (group_bytes, steps_per_group), for an ensemble"""
    key = key or ensemble_key([m.extent for m in ens.members], ens.members[0].storage)
    if use_cache:
        hit = _load().get(key)
        if hit:
            ens.set_launch(hit["block"], hit["waves"], hit["tiles"])
            ens.set_schedule(hit["group_bytes"], hit["steps_per_group"], hit.get("ticket", False))
            return dict(hit, cached=True)
    trials = []
    span = max(8, steps)
    ens.set_schedule(0, 1, False)
    for b in blocks:
        for w in waves:
            for t in tiles:
                ens.set_launch(b, w, t)
                ens.step(1, stream)
                ms = ens.time("persistent", span, stream) / span
                trials.append({"block": b, "waves": w, "tiles": t, "group_bytes": 0, "steps_per_group": 1,
                               "ticket": False, "ms": ms})
    best = min(trials, key=lambda t: t["ms"])
    ens.set_launch(best["block"], best["waves"], best["tiles"])
    for gb in group_bytes:
        for k in steps_per_group:
            for tk in tickets:
                ens.set_schedule(gb, k, tk)
                ens.step(k, stream)
                ms = ens.time("persistent", span, stream) / span
                trials.append({"block": best["block"], "waves": best["waves"], "tiles": best["tiles"],
                               "group_bytes": gb, "steps_per_group": k, "ticket": tk, "ms": ms})
    best = min(trials, key=lambda t: t["ms"])
    ens.set_launch(best["block"], best["waves"], best["tiles"])
    ens.set_schedule(best["group_bytes"], best["steps_per_group"], best["ticket"])
    keep = ("block", "waves", "tiles", "group_bytes", "steps_per_group", "ticket", "ms")
    _store(key, {k: best[k] for k in keep})
    return dict({k: best[k] for k in keep}, trials=trials)
