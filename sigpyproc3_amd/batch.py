"""Many sweeps in one persistent launch (``spp_batch_*``).

``SweepBatch`` takes a list of ``SweepPlan`` objects -- filterbanks of uneven nbits, nchans,
nsamps and DM lists -- and runs all of them, stage-1 copies included, as one kernel launch
over a device work queue. It replaces the reference's per-call loop (base.py:405-420, one
``kernels.dedisperse`` call per gulp and DM) for a whole set of observations at once: every
plane equals ``SweepPlan.execute`` of its own plan bit for bit.

A bound batch reads its streams and writes its planes at fixed device addresses, so a
caller that copies the next gulps into the same buffers can ``capture`` the launch once as
a HIP graph and ``replay`` it per gulp set.
"""

from __future__ import annotations

import ctypes

import numpy as np

from sigpyproc3_amd import _lib


class SweepBatch:
    """Host handle + device workspace of a batch of sweep plans."""

    def __init__(self, plans, g_budget: int = 0, *, stream=None) -> None:
        lib = _lib.load()
        _lib.require_gpu()
        self.plans = list(plans)  # the C batch reads the plans' host tables: keep them alive
        if not self.plans:
            msg = "a batch needs at least one plan"
            raise ValueError(msg)
        arr = (ctypes.c_void_p * len(self.plans))(*[p._handle.value for p in self.plans])
        handle = ctypes.c_void_p()
        _lib.check(lib.spp_batch_create(arr, len(self.plans), int(g_budget), ctypes.byref(handle)))
        self._handle = handle
        self._refresh_info()
        self.workspace = _lib.empty((int(self.info.workspace_bytes),), "uint8")
        _lib.check(lib.spp_batch_upload(self._handle, _lib.ptr(self.workspace), int(self.info.workspace_bytes),
                                        _lib.stream_handle(stream)))
        self._bound = None

    def _refresh_info(self) -> None:
        info = _lib.BatchInfo()
        _lib.check(_lib.load().spp_batch_get_info(self._handle, ctypes.byref(info)))
        self.info = info

    @property
    def rounds(self) -> np.ndarray:
        """Round (G pool generation) of each plan."""
        out = np.zeros(len(self.plans), dtype=np.int32)
        _lib.check(_lib.load().spp_batch_rounds(self._handle, _lib.ptr(out)))
        return out

    @property
    def total_samples(self) -> int:
        """DM-trial samples one run produces (sum over plans of the row lengths)."""
        return int(sum(p.total_samples for p in self.plans))

    @staticmethod
    def shape_count() -> int:
        return int(_lib.load().spp_batch_shape_count())

    @property
    def shape(self) -> tuple[int, int]:
        """(block shape, time tiles per stage-2 item)."""
        return int(self.info.shape), int(self.info.time_tiles_per_item)

    def set_shape(self, shape: int, time_tiles: int = 1, stream=None) -> None:
        _lib.check(_lib.load().spp_batch_set_shape(self._handle, int(shape), int(time_tiles)))
        self._refresh_info()
        if self._bound is not None:
            self.bind(*self._bound, stream=stream)

    def bind(self, packed, planes=None, ld=None, stream=None):
        """Fix the device streams (one packed uint8 tensor per plan) and planes (float32
        ``[ndm, ld]``, allocated when not given). Returns the planes."""
        n = len(self.plans)
        if len(packed) != n:
            msg = f"{len(packed)} streams for {n} plans"
            raise ValueError(msg)
        if ld is None:
            ld = [p.max_len for p in self.plans]
        ld = [int(v) for v in ld]
        if planes is None:
            planes = [_lib.empty((p.ndm, w), "float32") for p, w in zip(self.plans, ld)]
        for p, x, pl, w in zip(self.plans, packed, planes, ld):
            p._check_packed(x)
            if tuple(pl.shape) != (p.ndm, w) or not pl.is_contiguous():
                msg = f"plane must be a contiguous [{p.ndm}, {w}] float32 tensor"
                raise ValueError(msg)
        pk = (ctypes.c_void_p * n)(*[_lib.ptr(x).value for x in packed])
        pl = (ctypes.c_void_p * n)(*[_lib.ptr(x).value for x in planes])
        lds = np.ascontiguousarray(ld, dtype=np.int64)
        _lib.check(_lib.load().spp_batch_bind(self._handle, pk, pl, _lib.ptr(lds), _lib.ptr(self.workspace),
                                              int(self.info.workspace_bytes), _lib.stream_handle(stream)))
        self._bound = (list(packed), list(planes), ld)
        return list(planes)

    def run(self, stream=None) -> None:
        """One persistent launch over every plan (enqueued; no synchronisation)."""
        _lib.call("spp_batch_run", self._handle, _lib.ptr(self.workspace), int(self.info.workspace_bytes),
                  _lib.stream_handle(stream))

    def errors(self, stream=None) -> int:
        """Device error flags since the last call (0: every queue wait completed)."""
        flags = ctypes.c_int32(0)
        _lib.check(_lib.load().spp_batch_errors(self._handle, _lib.ptr(self.workspace),
                                                int(self.info.workspace_bytes), _lib.stream_handle(stream),
                                                ctypes.byref(flags)))
        return int(flags.value)

    def check(self, stream=None) -> None:
        if self.errors(stream) != 0:
            msg = "batch queue wait gave up: the planes are incomplete"
            raise RuntimeError(msg)

    def capture(self) -> None:
        """Capture one run as a HIP graph owned by the batch (after ``bind``)."""
        _lib.call("spp_batch_capture", self._handle, _lib.ptr(self.workspace), int(self.info.workspace_bytes))

    def replay(self, stream=None) -> None:
        _lib.call("spp_batch_replay", self._handle, _lib.stream_handle(stream))

    def execute(self, packed, planes=None, ld=None, stream=None):
        """bind + run + error check; returns the planes."""
        planes = self.bind(packed, planes, ld, stream=stream)
        self.run(stream=stream)
        self.check(stream=stream)
        return planes

    def tune(self, repeats: int = 2, cache: bool = True, stream=None) -> tuple[int, int]:
        """Time every (block shape, time tiles per item) on the bound data; keep the fastest."""
        from sigpyproc3_amd import tune

        return tune.tune_batch(self, repeats=repeats, cache=cache, stream=stream)

    def __del__(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.load().spp_batch_destroy(h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._handle = None
