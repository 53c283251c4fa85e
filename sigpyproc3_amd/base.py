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


class _GulpPipeline:
    """Packed gulps into HBM with the file read of the next gulp overlapping the current one.

    Two pinned host buffers and two device buffers alternate. A reader thread runs
    ``read_plan_packed_into`` (file reads release the GIL; each gulp is read in parallel
    pieces by the shared read pool) and reads gulp k + 1 while gulp k is copied to the device
    and its kernels run. A host buffer is handed back to the reader
    once its copy has completed; a device buffer is rewritten only by a copy enqueued on the
    same stream after the kernels that read it.
    """

    def __init__(self, reader, gulp: int, start: int, nsamps: int | None, skipback: int) -> None:
        import queue
        import threading

        t = _lib.torch()
        nsamps_eff = (reader.header.nsamples - start) if nsamps is None else nsamps
        nbytes = max(1, min(nsamps_eff, gulp) * reader.samp_stride)
        self.host = [t.empty(nbytes, dtype=t.uint8, pin_memory=True) for _ in range(2)]
        self.dev = [_lib.empty((nbytes,), "uint8") for _ in range(2)]
        self.free = [threading.Semaphore(1), threading.Semaphore(1)]
        self.items: queue.Queue = queue.Queue()  # at most 2 entries: one per acquired buffer
        self._stop = False
        bufs = [h.numpy() for h in self.host]
        from sigpyproc3_amd.io.fileio import read_pool

        pool = read_pool()

        def produce() -> None:
            try:
                gen = reader.read_plan_packed_into(bufs, gulp=gulp, start=start, nsamps=nsamps,
                                                   skipback=skipback, pool=pool)
                k = 0
                while True:
                    self.free[k % 2].acquire()
                    if self._stop:
                        return
                    item = next(gen, None)
                    if item is None:
                        break
                    self.items.put(item)
                    k += 1
                self.items.put(None)
            except BaseException as exc:  # noqa: BLE001 - re-raised in the consumer
                self.items.put(exc)

        self.thread = threading.Thread(target=produce, daemon=True)
        self.thread.start()

    def __iter__(self):
        t = _lib.torch()
        try:
            while True:
                item = self.items.get()
                if item is None:
                    return
                if isinstance(item, BaseException):
                    raise item
                nsamps_r, ii, i, nbytes = item
                dev = self.dev[i]
                dev[:nbytes].copy_(self.host[i][:nbytes], non_blocking=True)
                ev = t.cuda.Event()
                ev.record()
                yield nsamps_r, ii, dev, nbytes
                ev.synchronize()          # the copy has read host buffer i
                self.free[i].release()
        finally:
            self._stop = True
            for f in self.free:
                f.release()
            self.thread.join()


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
        for nsamps_r, ii, dev, _nb in self._device_gulps(gulp, start, nsamps, plan_kwargs, max_delay):
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
        try:
            for nsamps_r, _ii, dev, _nb in self._device_gulps(gulp, start, nsamps, plan_kwargs, max_delay):
                _lib.call("spp_subband", _lib.ptr(dev), nbits, big, hdr.nchans, nsamps_r,
                          _lib.ptr(d_dev), _lib.ptr(cts_dev), max_delay, nsub, _lib.ptr(out_dev),
                          out_len, stream)
                nwrite = (nsamps_r - max_delay) * nsub
                out_file.cwrite(out_dev[:nwrite].cpu().numpy())
        finally:
            out_file.close()
        return outfile_name

    def _device_gulps(self, gulp: int, start: int, nsamps: int | None, plan_kwargs: dict,
                      skipback: int = 0):
        """(nsamps_r, ii, device bytes, nbytes) of every packed gulp of the read_plan schedule.

        Without a caller ``allocator`` the gulps go through ``_GulpPipeline`` (double-buffered
        pinned memory, reads overlapping the device work); with one, every gulp is read into
        the caller's buffer and copied from there.
        """
        for k in ("description", "quiet"):
            plan_kwargs.pop(k, None)
        allocator = plan_kwargs.pop("allocator", None)
        if allocator is None and hasattr(self, "read_plan_packed_into"):
            yield from _GulpPipeline(self, gulp, start, nsamps, skipback)
            return
        nsamps_eff = (self.header.nsamples - start) if nsamps is None else nsamps
        up = _GulpUploader(max(1, min(nsamps_eff, gulp) * self.samp_stride))
        for nsamps_r, ii, raw in self.read_plan_packed(gulp=gulp, start=start, nsamps=nsamps,
                                                       skipback=skipback,
                                                       allocator=allocator or up.allocator()):
            yield nsamps_r, ii, up.upload(raw), raw.size

    def collapse(self, method: str = "sum", gulp: int = 16384, start: int = 0,
                 nsamps: int | None = None, *, parallel: bool = False, **plan_kwargs) -> TimeSeries:
        """Zero-DM time series: the sum over channels per sample (base.py:256-306).

        Runs ``spp_collapse`` (``kernels.extract_tim``: the float64 channel sum of each sample,
        exact for integer data) per gulp. Gulp ``ii`` is written at offset ``ii * nsamps_r``, as
        the reference does (base.py:297-303).
        """
        del parallel
        hdr = self.header
        tim_len = (hdr.nsamples - start) if nsamps is None else nsamps
        nbits, big = self._kernel_nbits()
        _lib.require_gpu()
        tim = _lib.zeros((max(tim_len, 1),), "float32")
        stream = _lib.stream_handle()
        for nsamps_r, ii, dev, _nb in self._device_gulps(gulp, start, nsamps, plan_kwargs):
            if ii * nsamps_r + nsamps_r > tim_len:
                msg = "output offset beyond the time series"
                raise ValueError(msg)
            _lib.call("spp_collapse", _lib.ptr(dev), nbits, big, hdr.nchans, nsamps_r, _lib.ptr(tim),
                      ii * nsamps_r, stream)
        tim_ar = tim[:tim_len].cpu().numpy() if tim_len > 0 else np.zeros(0, dtype=np.float32)
        if method == "mean":
            tim_ar /= hdr.nchans
        return TimeSeries(tim_ar, hdr.new_header({"nchans": 1, "dm": 0}))

    def bandpass(self, method: str = "sum", gulp: int = 16384, start: int = 0, *,
                 parallel: bool = False, nsamps: int | None = None, **plan_kwargs) -> TimeSeries:
        """Per-channel sums over time (base.py:308-358) on the GPU (``spp_bandpass``).

        ``parallel=False`` accumulates in float32 in time order per channel, as
        ``kernels.extract_bpass`` does; ``parallel=True`` adds each gulp's float64 sum once,
        as ``kernels.extract_bpass_par`` does.
        """
        hdr = self.header
        nbits, big = self._kernel_nbits()
        _lib.require_gpu()
        bpass = _lib.zeros((hdr.nchans,), "float32")
        stream = _lib.stream_handle()
        ws = None
        num_samples = 0
        for nsamps_r, _ii, dev, _nb in self._device_gulps(gulp, start, nsamps, plan_kwargs):
            need = _lib.workspace_size("spp_bandpass_workspace_size", nbits, hdr.nchans, nsamps_r)
            if ws is None or ws.numel() < need:
                ws = _lib.workspace(need)
            _lib.call("spp_bandpass", _lib.ptr(dev), nbits, big, hdr.nchans, nsamps_r,
                      _lib.ptr(bpass), 1 if parallel else 0, _lib.ptr(ws), ws.numel(), stream)
            num_samples += nsamps_r
        bpass_ar = bpass.cpu().numpy()
        if method == "mean":
            bpass_ar /= num_samples
        return TimeSeries(bpass_ar, hdr.new_header({"nchans": 1, "nsamples": len(bpass_ar)}))

    def _mask_tables(self, chan_mask: np.ndarray, mask_value) -> tuple[np.ndarray, np.ndarray]:
        """keep/set byte tables of one row such that (b & keep) | set equals masking the
        unpacked samples and packing them again the reference's way (no field masking)."""
        hdr = self.header
        nbits = hdr.nbits
        nchans = hdr.nchans
        if nbits == 32:
            keep = np.full(nchans * 4, 0xFF, dtype=np.uint8)
            setb = np.zeros(nchans * 4, dtype=np.uint8)
            vb = np.frombuffer(np.float32(mask_value).tobytes(), dtype=np.uint8)
            for c in np.flatnonzero(chan_mask[:nchans]):
                keep[4 * c : 4 * c + 4] = 0
                setb[4 * c : 4 * c + 4] = vb
            return keep, setb
        if nbits == 8:
            m = chan_mask[:nchans]
            keep = np.where(m, 0, 0xFF).astype(np.uint8)
            setb = np.where(m, int(mask_value), 0).astype(np.uint8)
            return keep, setb
        per = 8 // nbits
        big = nbits in (2, 4)
        fmask = (1 << nbits) - 1
        keep = np.zeros(nchans // per, dtype=np.uint32)
        setb = np.zeros(nchans // per, dtype=np.uint32)
        v = int(mask_value)
        for k in range(per):
            shift = (8 - (k + 1) * nbits) if big else k * nbits
            m = chan_mask[k:nchans:per]
            keep |= np.where(m, 0, fmask << shift).astype(np.uint32)
            setb |= np.where(m, v << shift, 0).astype(np.uint32)
        return (keep & 0xFF).astype(np.uint8), (setb & 0xFF).astype(np.uint8)

    def apply_channel_mask(self, chan_mask, mask_value: float = 0, outfile_name: str | None = None,
                           gulp: int = 16384, start: int = 0, nsamps: int | None = None, *,
                           parallel: bool = False, **plan_kwargs) -> str:
        """Replace masked channels and write a new file (base.py:530-583).

        The mask is applied to the file's own bytes on the GPU (``spp_mask_bytes``); the
        result equals masking the unpacked samples and packing them as the reference does.
        """
        del parallel
        hdr = self.header
        if outfile_name is None:
            outfile_name = f"{hdr.basename}_masked.fil"
        chan_mask = np.array(chan_mask).astype("bool")
        if chan_mask.size < hdr.nchans:
            msg = f"chan_mask has {chan_mask.size} entries < nchans = {hdr.nchans}"
            raise ValueError(msg)
        self._kernel_nbits()
        mask_value = np.float32(mask_value).astype(hdr.dtype)
        keep, setb = self._mask_tables(chan_mask, mask_value)
        _lib.require_gpu()
        keep_d, set_d = _lib.to_device(keep), _lib.to_device(setb)
        stream = _lib.stream_handle()
        out_file = hdr.prep_outfile(outfile_name)
        try:
            for _nr, _ii, dev, nb in self._device_gulps(gulp, start, nsamps, plan_kwargs):
                _lib.call("spp_mask_bytes", _lib.ptr(dev), nb, self.samp_stride, _lib.ptr(keep_d),
                          _lib.ptr(set_d), stream)
                out_file.write(dev[:nb].cpu().numpy().tobytes())
        finally:
            out_file.close()
        return outfile_name

    def remove_zerodm(self, outfile_name: str | None = None, gulp: int = 16384, start: int = 0,
                      nsamps: int | None = None, *, parallel: bool = False, **plan_kwargs) -> str:
        """Remove the channel-weighted zero-DM signal and write a new file (base.py:961-1028).

        The bandpass is ``self.bandpass(**plan_kwargs)`` (whole file, float32 sums) and the
        weights ``bpass / bpass.sum()`` in float32 on the host, as in the reference; each
        gulp then runs ``spp_remove_zerodm`` on the GPU.
        """
        del parallel
        hdr = self.header
        if outfile_name is None:
            outfile_name = f"{hdr.basename}_noZeroDM.fil"
        nbits, big = self._kernel_nbits()
        bpass = self.bandpass(**dict(plan_kwargs)).data
        chanwts = bpass / bpass.sum()
        _lib.require_gpu()
        w_d = _lib.to_device(np.ascontiguousarray(chanwts, dtype=np.float32))
        b_d = _lib.to_device(np.ascontiguousarray(bpass, dtype=np.float32))
        out_dtype = "float32" if nbits == 32 else "uint8"
        nsamps_eff = (hdr.nsamples - start) if nsamps is None else nsamps
        out_dev = _lib.empty((max(1, min(nsamps_eff, gulp) * hdr.nchans),), out_dtype)
        stream = _lib.stream_handle()
        out_file = hdr.prep_outfile(outfile_name, nbits=hdr.nbits)
        try:
            for nsamps_r, _ii, dev, _nb in self._device_gulps(gulp, start, nsamps, plan_kwargs):
                _lib.call("spp_remove_zerodm", _lib.ptr(dev), nbits, big, hdr.nchans, nsamps_r,
                          _lib.ptr(w_d), _lib.ptr(b_d), _lib.ptr(out_dev), stream)
                out_file.cwrite(out_dev[: nsamps_r * hdr.nchans].cpu().numpy())
        finally:
            out_file.close()
        return outfile_name

    def fold(self, period: float, dm: float, accel: float = 0, nbins: int = 50, nints: int = 32,
             nbands: int = 32, gulp: int = 16384, start: int = 0, nsamps: int | None = None,
             **plan_kwargs):
        """Fold into a (subint, subband, phase) cube (base.py:1109-1207) on the GPU.

        Per gulp the host lists the valid samples by (subint, phase bin) with the
        reference's float64 arithmetic (``foldedcube.fold_cells``); ``spp_fold`` then sums
        every cell to the reference's float32 value (exact integer totals while a cell stays
        below 2^24, the reference's order beyond). Returns ``FoldedData``.
        """
        from sigpyproc3_amd.foldedcube import FoldedData, band_starts, fold_cells

        hdr = self.header
        if np.isclose(np.modf(period / hdr.tsamp)[0], 0.001, atol=0):
            self.logger.warning("Foldng interval is an integer multiple of the sampling time")
        if nbins > period / hdr.tsamp:
            self.logger.warning("Number of phase bins is greater than period/sampling time")
        if (hdr.nsamples * hdr.nchans) // (nbands * nints * nbins) < 10:
            msg = f"nbands x nints x nbins is too large: {nbands * nints * nbins}"
            raise ValueError(msg)
        nbands = min(nbands, hdr.nchans)
        chan_delays = np.atleast_1d(hdr.get_dmdelays(dm)).astype(np.int32)
        max_delay = int(chan_delays.max())
        gulp = max(2 * max_delay, gulp)
        nbits, big = self._kernel_nbits()
        _lib.require_gpu()
        ncell = nbins * nints * nbands
        fold_d = _lib.zeros((ncell,), "float32")
        count_d = _lib.zeros((ncell,), "int32")
        d_dev = _lib.to_device(chan_delays)
        bands_dev = _lib.to_device(band_starts(hdr.nchans, nbands))
        stream = _lib.stream_handle()
        keep = []
        ws = None
        for nsamps_r, ii, dev, _nb in self._device_gulps(gulp, start, nsamps, plan_kwargs, max_delay):
            nvalid = nsamps_r - max_delay
            if nvalid <= 0:
                continue
            if int(np.abs(chan_delays).max()) >= nsamps_r:
                msg = "channel delays exceed the gulp"
                raise ValueError(msg)
            subint, phasebin = fold_cells(ii * (gulp - max_delay), nvalid, hdr.nsamples, hdr.tsamp,
                                          period, accel, nbins, nints)
            key = subint * nbins + phasebin
            order = np.argsort(key, kind="stable").astype(np.int32)
            keys, first = np.unique(key[order], return_index=True)
            key_start = np.append(first, nvalid).astype(np.int32)
            key_pos = ((keys // nbins) * nbins * nbands + keys % nbins).astype(np.int32)
            bufs = [_lib.to_device(order), _lib.to_device(key_start), _lib.to_device(key_pos)]
            keep.append(bufs)
            need = _lib.workspace_size("spp_fold_workspace_size", nbits, hdr.nchans, nsamps_r, nbands)
            if ws is None or ws.numel() < need:
                ws = _lib.workspace(need)
            _lib.call("spp_fold", _lib.ptr(dev), nbits, big, hdr.nchans, nsamps_r, _lib.ptr(d_dev),
                      _lib.ptr(bufs[0]), nvalid, _lib.ptr(bufs[1]), _lib.ptr(bufs[2]), len(keys),
                      _lib.ptr(bands_dev), nbands, nbins, _lib.ptr(fold_d), _lib.ptr(count_d),
                      _lib.ptr(ws), ws.numel(), stream)
        fold_ar = fold_d.cpu().numpy()
        count_ar = count_d.cpu().numpy()
        del keep
        with np.errstate(divide="ignore", invalid="ignore"):
            fold_ar /= count_ar
        fold_ar = fold_ar.reshape(nints, nbands, nbins)
        return FoldedData(fold_ar, hdr.new_header(), period, dm, accel)

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
