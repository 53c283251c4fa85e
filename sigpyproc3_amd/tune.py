"""Autotuner for the sweep's block shapes.

One-plan sweeps: each ``spp_sweep_plan`` variant (DM rows per wave x waves x channels per
stage) gives the same bit-exact plane; their speed depends on nbits (field width), nchans
and the DM tile spreads. ``tune_plan`` times each variant with HIP events on the stream the
work runs on and keeps the fastest.

Batches: ``tune_batch`` searches DM rows per item x channels per stage (the batch kernel's
block shapes) x time tiles per item on the bound batch.

Choices are cached in a JSON file keyed by the device, the number of block shapes the library
offers, and the exact work: nbits, nchans, nsamps and a hash of the delay table of every plan.
"""

from __future__ import annotations

import hashlib
import json
import os
import threading
from pathlib import Path

from sigpyproc3_amd import _lib

_lock = threading.Lock()
_last: dict = {}

BATCH_TIME_TILES = (1, 2, 4)


def last_times() -> dict:
    """Per-candidate times (ms) of the most recent timed tuning."""
    return dict(_last)


def cache_path() -> Path:
    env = os.environ.get("SPP_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "sigpyproc3_amd" / "sweep_tune.json"


def _device_name() -> str:
    t = _lib.torch()
    return t.cuda.get_device_name() if t.cuda.is_available() else "cpu"


def _plan_sig(plan) -> str:
    h = hashlib.sha1(plan.table.tobytes()).hexdigest()[:16]
    return f"nbits={plan.nbits}|nchans={plan.nchans}|ndm={plan.ndm}|nsamps={plan.nsamps}|table={h}"


def _key(plan) -> str:
    # the shape count keeps a choice made by a library with another set of shapes from applying
    return f"{_device_name()}|shapes={plan.variant_count()}|{_plan_sig(plan)}"


def _batch_key(batch) -> str:
    h = hashlib.sha1("/".join(_plan_sig(p) for p in batch.plans).encode()).hexdigest()[:24]
    return (f"{_device_name()}|batch|shapes={batch.shape_count()}|n={len(batch.plans)}"
            f"|rounds={int(batch.info.rounds)}|{h}")


def _load() -> dict:
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store(entries: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(entries, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home: the choice still applies to this run


def _timed(fn, repeats: int, stream) -> float:
    """Average ms of fn() over `repeats`, events recorded on the stream fn enqueues on."""
    t = _lib.torch()
    st = stream if stream is not None else t.cuda.current_stream()
    fn()  # warm
    e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(repeats):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / repeats


def tune_plan(plan, packed, plane=None, repeats: int = 2, cache: bool = True, stream=None) -> int:
    key = _key(plan)
    with _lock:
        entries = _load() if cache else {}
    if key in entries:
        plan.variant = int(entries[key]["variant"])
        return plan.variant
    if plane is None:
        plane = _lib.empty((plan.ndm, plan.max_len), "float32")
    plan.build(packed, stream=stream)
    times = {}
    for v in range(plan.variant_count()):
        plan.variant = v
        times[v] = _timed(lambda: plan.run(plane, stream=stream), repeats, stream)
    _last.clear()
    _last.update({k: round(v, 3) for k, v in times.items()})
    best = min(times, key=times.get)
    plan.variant = best
    if cache:
        with _lock:
            entries = _load()
            entries[key] = {"variant": best, "ms": {str(k): round(v, 4) for k, v in times.items()}}
            _store(entries)
    return best


def tune_batch(batch, repeats: int = 2, cache: bool = True, stream=None) -> tuple[int, int]:
    """Pick (block shape, time tiles per item) for a bound batch by timing each."""
    if batch._bound is None:
        msg = "bind the batch before tuning it"
        raise ValueError(msg)
    key = _batch_key(batch)
    with _lock:
        entries = _load() if cache else {}
    if key in entries:
        batch.set_shape(int(entries[key]["shape"]), int(entries[key]["time_tiles"]), stream=stream)
        return batch.shape
    times = {}
    for s in range(batch.shape_count()):
        for tt in BATCH_TIME_TILES:
            batch.set_shape(s, tt, stream=stream)
            times[(s, tt)] = _timed(lambda: batch.run(stream=stream), repeats, stream)
    batch.check(stream=stream)
    _last.clear()
    _last.update({f"{s}x{tt}": round(v, 3) for (s, tt), v in times.items()})
    best = min(times, key=times.get)
    batch.set_shape(*best, stream=stream)
    if cache:
        with _lock:
            entries = _load()
            entries[key] = {"shape": best[0], "time_tiles": best[1],
                            "ms": {f"{s}x{tt}": round(v, 4) for (s, tt), v in times.items()}}
            _store(entries)
    return best
