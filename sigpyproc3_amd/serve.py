"""Streaming DM sweep over many filterbank streams, gulps of varying length.

Each stream has a fixed DM list and therefore one ``SweepPlan``, created once with a capacity
of its largest gulp plus an overlap equal to the largest delay of its table. A gulp of ``n``
new samples is placed after the last ``overlap`` samples of the stream's previous gulp and
swept over ``overlap + n`` samples (``SweepPlan.execute_n``), so row ``i`` of the plane gains
exactly ``n`` new output samples, at plane offsets ``[overlap - md_i, overlap - md_i + n)``
(``md_i`` = the row's largest delay), continuing the previous gulp's row without a gap. The
first gulp of a stream has no overlap: its rows hold ``n - md_i`` samples.

Every stream owns a HIP stream (default priority), two device input buffers used in turn
(the overlap is copied from one to the other), a plane and its plan workspace. Work for a
gulp is enqueued on the stream's HIP stream (host-to-device copy of the new packed samples
from pinned memory, overlap copy, sweep) and completes in order per stream; streams run
concurrently. ``run_arrivals`` replays an arrival schedule in real time (open loop) and
returns the latency of every gulp, from its arrival to the completion of its sweep.
"""

from __future__ import annotations

import threading
import time
from dataclasses import dataclass

import numpy as np

from sigpyproc3_amd import _lib
from sigpyproc3_amd.sweep import SweepPlan


@dataclass
class Gulp:
    stream: int
    nsamps: int
    arrival: float = 0.0     # seconds from the start of the run
    done: float = float("nan")


class StreamState:
    """Device state of one stream: plan, input buffers, plane, HIP stream."""

    def __init__(self, table: np.ndarray, nbits: int, big: int, max_gulp: int) -> None:
        t = _lib.torch()
        self.table = np.ascontiguousarray(table, dtype=np.int32)
        self.overlap = int(self.table.max())
        self.capacity = int(max_gulp) + self.overlap
        self.nbits = nbits
        self.row_bytes = self.table.shape[1] * nbits // 8
        self.stream = t.cuda.Stream()
        self.plan = SweepPlan(self.table, self.capacity, nbits, big, stream=self.stream)
        nb = self.capacity * self.row_bytes
        self.inbuf = [_lib.empty((nb,), "uint8"), _lib.empty((nb,), "uint8")]
        self.plane = _lib.empty((self.plan.ndm * self.plan.max_len,), "float32")
        self.cur = 0
        self.have = 0            # samples of the last gulp held in inbuf[cur]
        self.first = True

    def submit(self, host_src, nsamps: int):
        """Enqueue one gulp (``nsamps`` new packed samples from pinned host memory); returns
        (plane, ld, row_offset, event): the new samples of row i are plane rows[i] at
        ``[row_offset[i], row_offset[i] + nsamps)`` once ``event`` has completed."""
        t = _lib.torch()
        ov = 0 if self.first else self.overlap
        n = ov + int(nsamps)
        if n > self.capacity or n <= self.overlap:
            msg = f"gulp of {nsamps} samples outside (0, {self.capacity - self.overlap}] (overlap {self.overlap})"
            raise ValueError(msg)
        rb = self.row_bytes
        with t.cuda.stream(self.stream):
            dst = self.inbuf[self.cur ^ 1]
            if ov:
                src = self.inbuf[self.cur]
                dst[: ov * rb].copy_(src[(self.have - ov) * rb : self.have * rb], non_blocking=True)
            dst[ov * rb : n * rb].copy_(host_src[: nsamps * rb], non_blocking=True)
            ld = int(self.plan.row_len_for(n).max())
            self.plan.execute_n(dst, n, self.plane, ld=ld, stream=self.stream)
            ev = t.cuda.Event()
            ev.record(self.stream)
        self.cur ^= 1
        self.have = n
        md = self.table.max(axis=1).astype(np.int64)
        row_offset = (ov - md) if ov else np.zeros_like(md)
        self.first = False
        return self.plane, ld, row_offset, ev


def make_source(max_gulp: int, row_bytes: int, seed: int = 5):
    """Pinned host buffer of random packed samples, the payload of every gulp."""
    t = _lib.torch()
    host = t.empty(max_gulp * row_bytes, dtype=t.uint8, pin_memory=True)
    g = t.Generator().manual_seed(seed)
    host.copy_(t.randint(0, 256, (host.numel(),), dtype=t.uint8, generator=g))
    return host


def poisson_schedule(nstreams: int, ngulps: int, rate: float, gulp_lo: int, gulp_hi: int,
                     seed: int) -> list[Gulp]:
    """``ngulps`` arrivals of a Poisson process of ``rate`` gulps/s, each on a uniformly
    chosen stream, with lengths ~ U[gulp_lo, gulp_hi]."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, ngulps) if rate > 0 else np.zeros(ngulps)
    arr = np.cumsum(gaps) - (gaps[0] if ngulps else 0.0)
    streams = rng.integers(0, nstreams, ngulps)
    sizes = rng.integers(gulp_lo, gulp_hi + 1, ngulps)
    return [Gulp(int(s), int(n), float(a)) for s, n, a in zip(streams, sizes, arr)]


def run_arrivals(states: list[StreamState], source, schedule: list[Gulp]) -> float:
    """Replay ``schedule`` in real time; fills ``Gulp.done`` (seconds from the start) and
    returns the wall time until the last gulp completed."""
    pending: list[tuple[Gulp, object]] = []
    lock = threading.Lock()
    finished = threading.Event()
    t0 = time.perf_counter()

    def watch() -> None:
        left = len(schedule)
        while left:
            with lock:
                cur = list(pending)
            for g, ev in cur:
                if ev.query():
                    g.done = time.perf_counter() - t0
                    with lock:
                        pending.remove((g, ev))
                    left -= 1
            time.sleep(2e-5)
        finished.set()

    th = threading.Thread(target=watch, daemon=True)
    th.start()
    for g in schedule:
        wait = g.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        *_, ev = states[g.stream].submit(source, g.nsamps)
        with lock:
            pending.append((g, ev))
    finished.wait()
    th.join()
    return max(g.done for g in schedule)


def latency_summary(schedule: list[Gulp]) -> dict:
    lat = np.array([g.done - g.arrival for g in schedule]) * 1e3
    return {"n": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "max_ms": float(lat.max())}
