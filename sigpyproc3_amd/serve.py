"""Streaming DM-sweep service: continuous batching of gulps from many streams under an HBM
budget.

The reference's only streaming model is one stream read in overlapping gulps
(``FilReader.read_plan``, readers.py:207-233, with ``Filterbank.dedisperse``'s skip-back of
the largest delay, base.py:402-411). The server keeps that model per stream and runs many
streams at once:

* ``open_stream`` registers a stream with its own DM list (delay table), bit depth and
  largest gulp. Its plan comes from a cache keyed by the table, nbits, bit order and
  capacity, and its tables stay resident on the device, together with the stream's overlap
  buffer (the last ``overlap`` = largest-delay samples of its previous gulp's input).
* ``submit`` queues a gulp of ``n`` new packed samples (pinned host memory). A gulp's input
  is the stream's overlap followed by the new samples, so row ``i`` gains exactly ``n``
  output samples, continuing the previous gulp's row without a gap (the first gulp of a
  stream has no overlap: its rows hold ``n - md_i`` samples).
* A scheduler thread admits queued gulps in arrival order whenever their working set (input,
  stage-1 copy, plane) fits the HBM budget (one device buffer, sub-allocated: no device
  allocator calls while serving) and their stream has no other gulp on the device (a
  stream's gulps run in order; memory is not held by gulps that could not start yet), and enqueues each one's copies and sweep on its
  stream's HIP stream (default priority): gulps of different streams run concurrently, and
  gulps that arrive while others are on the device join as memory frees up. A gulp's input
  and stage-1 copy are returned to the budget when it completes; its plane when the caller
  releases the result. ``max_queued`` bounds the queue (``submit`` then raises QueueFull:
  backpressure instead of unbounded latency).

Results do not depend on what else is in flight: every gulp reads only its own input and
its stream's overlap and resident tables, and integer sums are exact.
"""

from __future__ import annotations

import collections
import ctypes
import hashlib
import threading
import time
from dataclasses import dataclass

import numpy as np

from sigpyproc3_amd import _lib
from sigpyproc3_amd.sweep import SweepPlan


class QueueFull(RuntimeError):
    """The server's queue holds ``max_queued`` gulps: the caller must retry later."""


class Arena:
    """First-fit allocator of byte ranges in one device buffer of ``total`` bytes: the HBM
    budget (host bookkeeping only; offsets and sizes are multiples of 256). Freed ranges
    coalesce with their neighbours."""

    def __init__(self, total: int) -> None:
        self.total = int(total) // 256 * 256
        self.free_list: list[list[int]] = [[0, self.total]] if self.total > 0 else []  # [offset, size]
        self.used = 0
        self.peak = 0

    def alloc(self, nbytes: int) -> int | None:
        n = _align(nbytes)
        for k, (off, size) in enumerate(self.free_list):
            if size >= n:
                if size == n:
                    del self.free_list[k]
                else:
                    self.free_list[k] = [off + n, size - n]
                self.used += n
                self.peak = max(self.peak, self.used)
                return off
        return None

    def free(self, off: int, nbytes: int) -> None:
        n = _align(nbytes)
        if n == 0:
            return
        fl = self.free_list
        k = 0
        while k < len(fl) and fl[k][0] < off:
            k += 1
        if (k > 0 and fl[k - 1][0] + fl[k - 1][1] > off) or (k < len(fl) and off + n > fl[k][0]):
            msg = f"double free or overlap at offset {off}"
            raise RuntimeError(msg)
        fl.insert(k, [off, n])
        if k + 1 < len(fl) and fl[k][0] + fl[k][1] == fl[k + 1][0]:
            fl[k][1] += fl[k + 1][1]
            del fl[k + 1]
        if k > 0 and fl[k - 1][0] + fl[k - 1][1] == fl[k][0]:
            fl[k - 1][1] += fl[k][1]
            del fl[k]
        self.used -= n


def admit_in_order(pending: collections.deque, arena: Arena, busy=None) -> list:
    """Pop the queued requests (objects with ``need`` bytes and a ``handle``) to admit now, in
    arrival order. A request waits while its stream (``handle``) has a gulp on the device or
    an older gulp still queued (``busy``: the set of streams with a gulp in flight; None
    disables the check), and the first request that may start but does not fit the arena
    stops the scan, so a large gulp is never overtaken and starved. Sets each admitted
    request's ``offset`` and adds its stream to ``busy``."""
    out = []
    seen = set()
    keep = collections.deque()
    blocked = False
    while pending:
        req = pending.popleft()
        h = getattr(req, "handle", None)
        if blocked or (busy is not None and (h in busy or h in seen)):
            keep.append(req)
            seen.add(h)
            continue
        off = arena.alloc(req.need)
        if off is None:
            keep.append(req)
            seen.add(h)
            blocked = True
            continue
        req.offset = off
        out.append(req)
        seen.add(h)
        if busy is not None:
            busy.add(h)
    pending.extend(keep)
    return out


def _align(n: int) -> int:
    return (int(n) + 255) // 256 * 256


@dataclass
class _PlanEntry:
    plan: SweepPlan
    tables: object     # view of the arena holding the uploaded plan tables
    offset: int        # its arena offset
    refs: int = 0


class StreamHandle:
    """One registered stream: its plan, resident overlap buffer and HIP stream."""

    def __init__(self, server, key, entry: _PlanEntry, nbits: int, max_gulp: int, ov_offset: int) -> None:
        t = _lib.torch()
        self.server = server
        self.key = key
        self.entry = entry
        self.plan = entry.plan
        self.nbits = nbits
        self.nchans = self.plan.nchans
        self.row_bytes = self.nchans * nbits // 8
        self.overlap = int(self.plan.table.max())
        self.max_gulp = int(max_gulp)
        self.md = self.plan.table.max(axis=1).astype(np.int64)
        self.stream = t.cuda.Stream()
        self.ov_bytes = self.overlap * self.row_bytes
        self.ov_offset = ov_offset
        self.overlap_buf = server._view(ov_offset, self.ov_bytes)
        self.first = True
        self.closed = False


class GulpRequest:
    """A submitted gulp. After ``wait()``: the new samples of DM row i are
    ``plane[i, row_offset[i] : row_offset[i] + nsamps]`` (first gulp of a stream: the
    row's ``row_len[i]`` samples from 0). ``release()`` returns the plane to the budget."""

    def __init__(self, handle: StreamHandle, host, nsamps: int, first: bool) -> None:
        self.handle = handle
        self.host = host
        self.nsamps = int(nsamps)
        self.first = first
        ov = 0 if first else handle.overlap
        self.n = ov + self.nsamps
        self.ld = int(self.n - handle.md.min())
        self.row_len = (self.n - handle.md).astype(np.int64)
        self.row_offset = np.zeros_like(handle.md) if first else (ov - handle.md)
        g = ctypes.c_int64()
        _lib.check(_lib.load().spp_sweep_g_bytes(handle.plan._handle, self.n, ctypes.byref(g)))
        self.g_bytes = int(g.value)
        self.in_bytes = self.n * handle.row_bytes
        self.plane_bytes = handle.plan.ndm * self.ld * 4
        # one arena block: [plane | input | stage-1 copy]; the tail is freed at completion
        self.work_bytes = _align(self.in_bytes) + _align(self.g_bytes)
        self.need = _align(self.plane_bytes) + self.work_bytes
        self.offset = -1
        self.t_submit = time.perf_counter()
        self.t_admit = float("nan")
        self.t_done = float("nan")
        self.plane = None
        self._event = None
        self._done = threading.Event()
        self._released = False
        self.error: BaseException | None = None

    @property
    def done(self) -> bool:
        return self._done.is_set()

    def wait(self, timeout: float | None = None):
        if not self._done.wait(timeout):
            msg = "gulp not complete"
            raise TimeoutError(msg)
        if self.error is not None:
            raise self.error
        return self

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit

    def rows(self) -> list[np.ndarray]:
        """The new output samples of every DM row (host copies)."""
        self.wait()
        plane = self.plane.cpu().numpy()
        n = self.nsamps if not self.first else None
        out = []
        for i in range(plane.shape[0]):
            a = int(self.row_offset[i])
            b = a + (n if n is not None else int(self.row_len[i]))
            out.append(plane[i, a:b])
        return out

    def release(self) -> None:
        """Return the plane to the budget (after the caller's reads of it have completed)."""
        self.wait()
        self.handle.server._release_plane(self)


class Server:
    """Continuous-batching sweep service under an HBM budget (see the module docstring).

    The budget is one device buffer allocated up front; every plan table, overlap buffer and
    gulp working set is a range of it (``Arena``), so serving never calls the device
    allocator."""

    def __init__(self, budget_bytes: int, max_queued: int = 1024) -> None:
        _lib.load()
        _lib.require_gpu()
        self.arena = Arena(budget_bytes)
        self.mem = _lib.empty((max(self.arena.total, 256),), "uint8")
        self.max_queued = int(max_queued)
        self._plans: dict = {}
        self._pending: collections.deque = collections.deque()
        self._inflight: list = []
        self._busy: set = set()   # streams with a gulp on the device
        self._cv = threading.Condition()
        self._stop = False
        self.admitted = 0
        self.rejected = 0
        self._device = _lib.torch().cuda.current_device()
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    @property
    def peak_used(self) -> int:
        return self.arena.peak

    @property
    def used(self) -> int:
        return self.arena.used

    def _view(self, offset: int, nbytes: int):
        return self.mem[offset : offset + int(nbytes)]

    # ------------------------------------------------------------------ streams
    def open_stream(self, table: np.ndarray, nbits: int, big: int, max_gulp: int) -> StreamHandle:
        table = np.ascontiguousarray(np.atleast_2d(table), dtype=np.int32)
        overlap = int(table.max())
        capacity = int(max_gulp) + overlap
        key = (hashlib.sha1(table.tobytes()).hexdigest(), table.shape, int(nbits), int(big), capacity)
        with self._cv:
            entry = self._plans.get(key)
            new_entry = entry is None
            if new_entry:
                plan = SweepPlan(table, capacity, nbits, big, allocate=False)
                tbytes = int(plan.info.tables_bytes)
                toff = self.arena.alloc(tbytes)
                if toff is None:
                    msg = "HBM budget too small for the stream's resident plan tables"
                    raise RuntimeError(msg)
                tables = self._view(toff, tbytes)
                _lib.check(_lib.load().spp_sweep_plan_upload(plan._handle, _lib.ptr(tables), tbytes, None))
                entry = _PlanEntry(plan, tables, toff)
            ooff = self.arena.alloc(overlap * table.shape[1] * nbits // 8)
            if ooff is None:
                if new_entry:
                    self.arena.free(entry.offset, int(entry.plan.info.tables_bytes))
                msg = "HBM budget too small for the stream's overlap buffer"
                raise RuntimeError(msg)
            if new_entry:
                self._plans[key] = entry
            entry.refs += 1
            return StreamHandle(self, key, entry, int(nbits), max_gulp, ooff)

    def close_stream(self, h: StreamHandle) -> None:
        """Release the stream's overlap buffer and, with its last user, the plan tables (after
        its queued gulps have completed)."""
        while True:
            with self._cv:
                busy = any(r.handle is h for r in self._pending) or any(r.handle is h for r in self._inflight)
                if not busy:
                    break
            time.sleep(1e-4)
        with self._cv:
            if h.closed:
                return
            h.closed = True
            h.stream.synchronize()
            self.arena.free(h.ov_offset, h.ov_bytes)
            h.overlap_buf = None
            h.entry.refs -= 1
            if h.entry.refs == 0:
                self.arena.free(h.entry.offset, int(h.entry.plan.info.tables_bytes))
                del self._plans[h.key]

    # ------------------------------------------------------------------ gulps
    def submit(self, h: StreamHandle, host, nsamps: int) -> GulpRequest:
        """Queue one gulp of ``nsamps`` new packed samples (a pinned host uint8 tensor of at
        least nsamps * row_bytes bytes, kept by the request until its copy completes)."""
        nsamps = int(nsamps)
        if h.closed:
            msg = "stream is closed"
            raise ValueError(msg)
        if nsamps > h.max_gulp or nsamps <= 0:
            msg = f"gulp of {nsamps} samples outside (0, {h.max_gulp}]"
            raise ValueError(msg)
        if h.first and nsamps <= h.overlap:
            msg = f"the first gulp needs more samples than the largest delay ({h.overlap})"
            raise ValueError(msg)
        if int(host.numel()) < nsamps * h.row_bytes:
            msg = "host buffer shorter than the gulp"
            raise ValueError(msg)
        with self._cv:
            if len(self._pending) >= self.max_queued:
                self.rejected += 1
                raise QueueFull(f"{len(self._pending)} gulps queued")
            req = GulpRequest(h, host, nsamps, h.first)
            if req.need > self.arena.total:
                msg = f"gulp needs {req.need} bytes, more than the whole budget ({self.arena.total})"
                raise ValueError(msg)
            h.first = False
            self._pending.append(req)
            self._cv.notify()
        return req

    def _release_plane(self, req: GulpRequest) -> None:
        with self._cv:
            if req._released:
                return
            req._released = True
            req.plane = None
            self.arena.free(req.offset, req.plane_bytes)
            self._cv.notify()

    # ------------------------------------------------------------------ scheduler
    def _launch(self, req: GulpRequest) -> None:
        t = _lib.torch()
        h = req.handle
        rb = h.row_bytes
        ov = 0 if req.first else h.overlap
        off = req.offset
        plane = self._view(off, req.plane_bytes).view(t.float32).view(h.plan.ndm, req.ld)
        off += _align(req.plane_bytes)
        inp = self._view(off, req.in_bytes)
        off += _align(req.in_bytes)
        g = self._view(off, req.g_bytes)
        with t.cuda.stream(h.stream):
            if ov:
                inp[: ov * rb].copy_(h.overlap_buf[: ov * rb], non_blocking=True)
            inp[ov * rb :].copy_(req.host[: req.nsamps * rb], non_blocking=True)
            if h.overlap:
                h.overlap_buf.copy_(inp[(req.n - h.overlap) * rb :], non_blocking=True)
            _lib.call("spp_sweep_execute_g", h.plan._handle, _lib.ptr(inp), req.n, _lib.ptr(plane), req.ld,
                      _lib.ptr(h.entry.tables), int(h.plan.info.tables_bytes), _lib.ptr(g), req.g_bytes,
                      _lib.stream_handle(h.stream))
            ev = t.cuda.Event()
            ev.record(h.stream)
        req.plane, req._event = plane, ev

    def _loop(self) -> None:
        t = _lib.torch()
        t.cuda.set_device(self._device)
        while True:
            with self._cv:
                if self._stop and not self._pending and not self._inflight:
                    return
                # completions: the input and stage-1 copy go back to the arena
                still = []
                for req in self._inflight:
                    if req._event.query():
                        req.t_done = time.perf_counter()
                        req.host = None
                        self.arena.free(req.offset + _align(req.plane_bytes), req.work_bytes)
                        self._busy.discard(req.handle)
                        req._done.set()
                    else:
                        still.append(req)
                self._inflight = still
                # admissions: arrival order, one gulp per stream on the device
                for req in admit_in_order(self._pending, self.arena, self._busy):
                    req.t_admit = time.perf_counter()
                    try:
                        self._launch(req)
                    except BaseException as exc:  # noqa: BLE001 - reported to the waiter
                        self.arena.free(req.offset, req.need)
                        self._busy.discard(req.handle)
                        req._released = True
                        req.error = exc
                        req.t_done = time.perf_counter()
                        req._done.set()
                        continue
                    self.admitted += 1
                    self._inflight.append(req)
                if not self._inflight and not self._stop:
                    self._cv.wait(0.05 if not self._pending else 1e-3)
                    continue
            time.sleep(2e-5)

    def shutdown(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc) -> None:
        self.shutdown()


# ---------------------------------------------------------------------- load generation
@dataclass
class Gulp:
    stream: int
    nsamps: int
    arrival: float = 0.0     # seconds from the start of the run
    done: float = float("nan")
    submitted: float = float("nan")  # seconds from the start when submit() returned
    rejected: bool = False
    request: GulpRequest | None = None


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


def run_arrivals(server: Server, handles: list[StreamHandle], source, schedule: list[Gulp],
                 on_done=None, on_submit=None) -> float:
    """Replay ``schedule`` in real time (open loop): each arrival is submitted at its time
    (a full queue rejects it). A consumer thread takes every gulp as it completes, calls
    ``on_done(gulp)`` (before its plane is released) and releases it. Fills ``Gulp.done``
    (seconds from the start) and returns the wall time until the last admitted gulp
    completed."""
    t0 = time.perf_counter()
    submitted: list[Gulp] = []
    lock = threading.Lock()
    all_submitted = threading.Event()
    end = [0.0]
    failure: list[BaseException] = []

    def consume() -> None:
        waiting: list[Gulp] = []
        try:
            while True:
                with lock:
                    waiting.extend(submitted)
                    submitted.clear()
                    last = all_submitted.is_set()
                rest = []
                for g in waiting:
                    if g.request.done:
                        g.request.wait()
                        g.done = g.request.t_done - t0
                        end[0] = max(end[0], g.done)
                        if on_done is not None:
                            on_done(g)
                        g.request.release()
                    else:
                        rest.append(g)
                waiting = rest
                if last and not waiting:
                    return
                time.sleep(1e-4)
        except BaseException as exc:  # noqa: BLE001 - re-raised by the caller
            failure.append(exc)

    th = threading.Thread(target=consume, daemon=True)
    th.start()
    for g in schedule:
        wait = g.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        try:
            req = server.submit(handles[g.stream], source, g.nsamps)
        except QueueFull:
            g.rejected = True
            continue
        g.request = req
        g.submitted = time.perf_counter() - t0
        if on_submit is not None:
            on_submit(g)
        with lock:
            submitted.append(g)
    with lock:
        all_submitted.set()
    th.join()
    if failure:
        raise failure[0]
    return end[0]


def latency_summary(schedule: list[Gulp], tsamp: float | None = None) -> dict:
    ok = [g for g in schedule if not g.rejected]
    lat = np.array([g.done - g.arrival for g in ok]) * 1e3
    out = {"n": int(lat.size), "rejected": len(schedule) - len(ok),
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}
    lag = np.array([g.submitted - g.arrival for g in ok]) * 1e3
    out["submit_lag_ms_p50"] = float(np.percentile(lag, 50))
    out["submit_lag_ms_max"] = float(lag.max())
    out["arrival_span_s"] = float(max(g.arrival for g in schedule) - min(g.arrival for g in schedule))
    out["last_done_s"] = float(max(g.done for g in ok))
    queue = np.array([g.request.t_admit - g.request.t_submit for g in ok]) * 1e3
    service = np.array([g.request.t_done - g.request.t_admit for g in ok]) * 1e3
    out["queue_ms_p50"] = float(np.percentile(queue, 50))
    out["queue_ms_p99"] = float(np.percentile(queue, 99))
    out["service_ms_p50"] = float(np.percentile(service, 50))
    out["service_ms_p99"] = float(np.percentile(service, 99))
    if tsamp is not None:
        dur = np.array([g.nsamps * tsamp for g in ok]) * 1e3
        ratio = lat / dur
        out["latency_over_duration_p50"] = float(np.percentile(ratio, 50))
        out["latency_over_duration_p99"] = float(np.percentile(ratio, 99))
        out["within_duration"] = float((ratio < 1.0).mean())
    return out
