"""Autotuner for the sweep's stage-2 block shape.

Each ``spp_sweep_plan`` variant (DM rows per wave x waves x channels per stage) gives the
same bit-exact plane; their speed depends on nbits (field width), nchans and the DM
tile spreads. ``tune_plan`` times each variant with HIP events on the caller's stream and
keeps the fastest, caching the winner per (device, nbits, nchans, ndm, nsamps) in JSON.
"""

from __future__ import annotations

import json
import os
import threading
from pathlib import Path

from sigpyproc3_amd import _lib

_lock = threading.Lock()
_last: dict[int, float] = {}


def last_times() -> dict[int, float]:
    """Per-variant stage-2 times (ms) of the most recent timed tuning."""
    return dict(_last)


def cache_path() -> Path:
    env = os.environ.get("SPP_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "sigpyproc3_amd" / "sweep_tune.json"


def _key(plan) -> str:
    t = _lib.torch()
    name = t.cuda.get_device_name() if t.cuda.is_available() else "cpu"
    return f"{name}|nbits={plan.nbits}|nchans={plan.nchans}|ndm={plan.ndm}|nsamps={plan.nsamps}"


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
        pass  # a read-only home: the choice still applies to this plan


def tune_plan(plan, packed, plane=None, repeats: int = 2, cache: bool = True, stream=None) -> int:
    key = _key(plan)
    with _lock:
        entries = _load() if cache else {}
    if key in entries:
        plan.variant = int(entries[key]["variant"])
        return plan.variant
    t = _lib.torch()
    if plane is None:
        plane = _lib.empty((plan.ndm, plan.max_len), "float32")
    plan.build(packed, stream=stream)
    times = {}
    for v in range(plan.variant_count()):
        plan.variant = v
        plan.run(plane, stream=stream)  # warm
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(repeats):
            plan.run(plane, stream=stream)
        e1.record()
        e1.synchronize()
        times[v] = e0.elapsed_time(e1) / repeats
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
