"""Filterbank operations on the GPU (mirrors ``sigpyproc.base.Filterbank``).

``dedisperse`` and ``subband`` follow the reference's gulp orchestration
(base.py:360-424, 1030-1107) exactly -- same delays, gulp size, read schedule and
output offsets -- but move each packed gulp to the device and run the fused
unpack + dedisperse / subband kernel there. ``dedisperse_sweep`` extends the path to a
whole DM-trial list in one pass over a device-resident stream.
"""

from __future__ import annotations

import logging
from abc import ABC, abstractmethod

import numpy as np

from sigpyproc3_amd import _lib
from sigpyproc3_amd.timeseries import TimeSeries

logger = logging.getLogger(__name__)


class _GulpUploader:
    """Host->device copies of packed gulps through one pinned buffer.

    ``read_plan`` reads every gulp into the single buffer its ``allocator`` returns; here
    that buffer is pinned host memory. Each gulp is copied asynchronously on the compute
    stream and the host waits for the copy only (not for the kernel), so the next file
    read overlaps the previous gulp's kernel. Stream order keeps the device buffer safe:
    a copy is enqueued after the kernel that read the previous contents.
    """

    def __init__(self, nbytes: int) -> None:
        t = _lib.torch()
        self.nbytes = nbytes
        self.host = t.empty(nbytes, dtype=t.uint8, pin_memory=True)
        self.dev = _lib.empty((nbytes,), "uint8")

    def allocator(self):
        buf = self.host.numpy()

        def alloc(n: int):
            if n != self.nbytes:
                msg = f"unexpected buffer request {n} != {self.nbytes}"
                raise ValueError(msg)
            return buf

        return alloc

    def upload(self, host_view: np.ndarray):
        t = _lib.torch()
        n = host_view.size
        src = self.host[:n] if host_view.ctypes.data == self.host.data_ptr() else t.from_numpy(host_view)
        self.dev[:n].copy_(src, non_blocking=True)
        ev = t.cuda.Event()
        ev.record()
        ev.synchronize()  # the host buffer is read again by the next file read
        return self.dev


class Filterbank(ABC):
    """Base class of frequency-major filterbank data readers (base.py:36-60)."""

    def __init__(self) -> None:
        self.logger = logger

    @property
    @abstractmethod
    def header(self):
        """Observation metadata."""

    @abstractmethod
    def read_plan(self, *args, **kwargs):
        """Gulp iterator of unpacked samples."""

    @abstractmethod
    def read_plan_packed(self, *args, **kwargs):
        """Gulp iterator of packed bytes."""

    def _kernel_nbits(self) -> tuple[int, int]:
        nbits = self.header.nbits
        if nbits == 16:
            msg = "No matching definition for 16-bit samples (the reference kernels take uint8 or float32)"
            raise TypeError(msg)
        big = 1 if (nbits in (2, 4)) else 0  # default bit orders (io/bits.py:175-182)
        return nbits, big

    def dedisperse(self, dm: float, ref_freq: str | float = "fch1", gulp: int = 16384,
                   start: int = 0, nsamps: int | None = None, *, parallel: bool = False,
                   **plan_kwargs) -> TimeSeries:
        """Dedisperse to a time series (base.py:360-424) on the GPU.

        ``parallel`` is accepted for signature compatibility; the GPU kernel serves both.
        """
        del parallel
        hdr = self.header
        chan_delays = hdr.get_dmdelays(dm, ref_freq)
        max_delay = int(np.max(np.abs(chan_delays)))
        gulp = max(2 * max_delay, gulp)
        tim_len = hdr.nsamples - max_delay
        if tim_len < 0:
            msg = "negative dimensions are not allowed"
            raise ValueError(msg)
        nbits, big = self._kernel_nbits()
        _lib.require_gpu()
        tim = _lib.zeros((max(tim_len, 1),), "float32")
        d_dev = _lib.to_device(np.ascontiguousarray(chan_delays, dtype=np.int32))
        stream = _lib.stream_handle()
        nsamps_eff = (hdr.nsamples - start) if nsamps is None else nsamps
        gulp_eff = min(nsamps_eff, gulp)
        up = _GulpUploader(max(1, gulp_eff * self.samp_stride))
        plan_kwargs.pop("description", None)
        plan_kwargs.pop("quiet", None)
        allocator = plan_kwargs.pop("allocator", None) or up.allocator()
        for nsamps_r, ii, raw in self.read_plan_packed(
                gulp=gulp, start=start, nsamps=nsamps, skipback=max_delay, allocator=allocator):
            dev = up.upload(raw)
            _lib.call("spp_dedisperse", _lib.ptr(dev), nbits, big, hdr.nchans, nsamps_r,
                      _lib.ptr(d_dev), max_delay, _lib.ptr(tim), ii * (gulp - max_delay), stream)
        data = tim[:tim_len].cpu().numpy() if tim_len > 0 else np.zeros(0, dtype=np.float32)
        return TimeSeries(data, hdr.new_header({"nchans": 1, "dm": dm, "nsamples": tim_len}))

    def subband(self, dm: float, nsub: int, ref_freq: str | float = "fch1",
                outfile_name: str | None = None, gulp: int = 16384, start: int = 0,
                nsamps: int | None = None, **plan_kwargs) -> str:
        """Write dedispersed subbands to a SIGPROC file (base.py:1030-1107) on the GPU.

        Each gulp's output buffer is zeroed before the kernel accumulates into it; the
        reference accumulates into an uninitialised buffer (base.py:1073-1074).
        """
        hdr = self.header
        subfactor = hdr.nchans // nsub
        chan_delays = hdr.get_dmdelays(dm, ref_freq)
        max_delay = int(np.max(np.abs(chan_delays)))
        gulp = max(2 * max_delay, gulp)
        out_len = (gulp - max_delay) * nsub
        new_foff = hdr.foff * hdr.nchans // nsub
        new_fch1 = hdr.ftop - new_foff / 2
        chan_to_sub = np.arange(hdr.nchans, dtype=np.int32) // subfactor
        updates = {"fch1": new_fch1, "foff": new_foff, "refdm": dm, "nchans": nsub, "nbits": 32}
        if outfile_name is None:
            outfile_name = f"{hdr.basename}_DM{dm:06.2f}.subbands"
        out_file = hdr.prep_outfile(outfile_name, updates=updates, nbits=32)
        nbits, big = self._kernel_nbits()
        _lib.require_gpu()
        d_dev = _lib.to_device(np.ascontiguousarray(chan_delays, dtype=np.int32))
        cts_dev = _lib.to_device(chan_to_sub)
        out_dev = _lib.zeros((max(out_len, 1),), "float32")
        stream = _lib.stream_handle()
        nsamps_eff = (hdr.nsamples - start) if nsamps is None else nsamps
        up = _GulpUploader(max(1, min(nsamps_eff, gulp) * self.samp_stride))
        for k in ("description", "quiet"):
            plan_kwargs.pop(k, None)
        allocator = plan_kwargs.pop("allocator", None) or up.allocator()
        try:
            for nsamps_r, _ii, raw in self.read_plan_packed(
                    gulp=gulp, start=start, nsamps=nsamps, skipback=max_delay, allocator=allocator):
                dev = up.upload(raw)
                _lib.call("spp_subband", _lib.ptr(dev), nbits, big, hdr.nchans, nsamps_r,
                          _lib.ptr(d_dev), _lib.ptr(cts_dev), max_delay, nsub, _lib.ptr(out_dev),
                          out_len, stream)
                nwrite = (nsamps_r - max_delay) * nsub
                out_file.cwrite(out_dev[:nwrite].cpu().numpy())
        finally:
            out_file.close()
        return outfile_name

    def dedisperse_sweep(self, dms, ref_freq: str | float = "fch1", start: int = 0,
                         nsamps: int | None = None):
        """Dedisperse at every DM of ``dms`` in one device pass.

        Row ``i`` of the result equals ``self.dedisperse(dms[i], ref_freq).data`` (for the
        default ``start``/``nsamps``; otherwise the dedispersion of samples
        ``[start, start + nsamps)``), computed from the packed stream resident on the GPU.
        Returns a :class:`~sigpyproc3_amd.sweep.SweepResult`.
        """
        from sigpyproc3_amd.sweep import SweepPlan, SweepResult, load_packed_stream

        hdr = self.header
        nbits, big = self._kernel_nbits()
        if nbits == 32:
            msg = "the sweep runs on integer (1/2/4/8-bit) data"
            raise ValueError(msg)
        dms = np.atleast_1d(np.asarray(dms, dtype=np.float64))
        table = np.atleast_2d(hdr.get_dmdelays(dms, ref_freq)).astype(np.int32)
        nsamps = (hdr.nsamples - start) if nsamps is None else nsamps
        packed = load_packed_stream(self, start, nsamps)
        plan = SweepPlan(table, nsamps, nbits, big)
        plane = plan.execute(packed)
        return SweepResult(plane=plane, row_len=plan.row_len, dms=dms, header=hdr, plan=plan,
                           table=table)
