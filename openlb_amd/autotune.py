"""Python face of the library's launch autotuner (olbgfx_autotune_lattice,
olbgfx_autotune_ensemble).

For a lattice the library sweeps threads per block and an occupancy cap
(waves per SIMD, enforced by an LDS reservation), then the order in which
blocks visit the cells; for an ensemble the tiles each persistent block walks
and the group schedule.  Every candidate is timed with HIP events on the
lattice itself, the fastest is applied, and the populations, offsets and
statistics are put back as they were, so tuning does not advance the
simulation.  Choices are cached in a text file keyed by shape, storage and
dynamics mix ($OLBGFX_TUNE_CACHE, default ~/.cache/olbgfx/autotune.txt).
"""
from __future__ import annotations

import ctypes
import os

from . import _check, library

CHOICES = 8
TRIALS = 128


class TuneSpace(ctypes.Structure):
    _fields_ = [
        ("n_block_threads", ctypes.c_uint32), ("block_threads", ctypes.c_uint32 * CHOICES),
        ("n_waves", ctypes.c_uint32), ("waves_per_simd", ctypes.c_uint32 * CHOICES),
        ("n_orders", ctypes.c_uint32), ("block_order", ctypes.c_int32 * CHOICES),
        ("n_tiles", ctypes.c_uint32), ("tiles_per_block", ctypes.c_uint32 * CHOICES),
        ("n_group_bytes", ctypes.c_uint32), ("group_bytes", ctypes.c_uint64 * CHOICES),
        ("n_steps_per_group", ctypes.c_uint32), ("steps_per_group", ctypes.c_uint32 * CHOICES),
        ("n_tickets", ctypes.c_uint32), ("ticket", ctypes.c_int32 * 2),
    ]


class TuneTrial(ctypes.Structure):
    _fields_ = [
        ("block_threads", ctypes.c_uint32), ("waves_per_simd", ctypes.c_uint32),
        ("block_order", ctypes.c_int32), ("tiles_per_block", ctypes.c_uint32),
        ("group_bytes", ctypes.c_uint64), ("steps_per_group", ctypes.c_uint32),
        ("ticket", ctypes.c_int32), ("ms_per_step", ctypes.c_double),
    ]


class TuneResult(ctypes.Structure):
    _fields_ = [("best", TuneTrial), ("cached", ctypes.c_int32), ("ntrials", ctypes.c_uint32),
                ("trials", TuneTrial * TRIALS)]


def cache_path() -> str:
    return os.environ.get("OLBGFX_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "olbgfx", "autotune.txt"))


def _space(**lists) -> TuneSpace:
    sp = TuneSpace()
    for name, count, values in (("block_threads", "n_block_threads", lists.get("blocks")),
                                ("waves_per_simd", "n_waves", lists.get("waves")),
                                ("block_order", "n_orders", lists.get("orders")),
                                ("tiles_per_block", "n_tiles", lists.get("tiles")),
                                ("group_bytes", "n_group_bytes", lists.get("group_bytes")),
                                ("steps_per_group", "n_steps_per_group", lists.get("steps_per_group")),
                                ("ticket", "n_tickets", lists.get("tickets"))):
        if values is None:
            continue
        arr = getattr(sp, name)
        if len(values) > len(arr):
            raise ValueError(f"at most {len(arr)} candidates for {name}")
        for k, v in enumerate(values):
            arr[k] = int(v)
        setattr(sp, count, len(values))
    return sp


def _trial(t: TuneTrial, ensemble: bool) -> dict:
    d = {"block": t.block_threads, "waves": t.waves_per_simd, "order": t.block_order, "ms": t.ms_per_step}
    if ensemble:
        d.update(tiles=t.tiles_per_block, group_bytes=t.group_bytes, steps_per_group=t.steps_per_group,
                 ticket=bool(t.ticket))
    return d


def _result(res: TuneResult, ensemble: bool) -> dict:
    out = _trial(res.best, ensemble)
    out["cached"] = bool(res.cached)
    out["trials"] = [_trial(res.trials[k], ensemble) for k in range(res.ntrials)]
    return out


def _cache_arg(use_cache: bool):
    if not use_cache:
        return None
    path = cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    return path.encode()


def tune_lattice(lat, steps: int = 4, stream=None, use_cache: bool = True,
                 blocks=None, waves=None, orders=None) -> dict:
    """pick (block_threads, waves_per_simd), then the block order, for one
    lattice and apply them; None takes the library's candidates"""
    res = TuneResult()
    sp = _space(blocks=blocks, waves=waves, orders=orders)
    _check(library().olbgfx_autotune_lattice(lat._h, int(steps), ctypes.byref(sp), _cache_arg(use_cache),
                                             stream, ctypes.byref(res)))
    return _result(res, False)


def tune_ensemble(ens, steps: int = 4, stream=None, use_cache: bool = True, blocks=None, waves=None,
                  tiles=None, orders=None, group_bytes=None, steps_per_group=None, tickets=None) -> dict:
    """pick (block_threads, waves_per_simd, tiles_per_block), then the block
    order, then the group schedule (group_bytes, steps_per_group, ticket), for
    an ensemble"""
    res = TuneResult()
    sp = _space(blocks=blocks, waves=waves, tiles=tiles, orders=orders, group_bytes=group_bytes,
                steps_per_group=steps_per_group, tickets=tickets)
    _check(library().olbgfx_autotune_ensemble(ens._h, int(steps), ctypes.byref(sp), _cache_arg(use_cache),
                                              stream, ctypes.byref(res)))
    return _result(res, True)
