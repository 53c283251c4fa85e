"""DM-trial sweep engine and the matched filter that follows it.

``SweepPlan`` wraps the C planner (``spp_sweep_plan_*``): it groups DM rows into tiles,
validates the delay table and sizes a device workspace, which it owns (a torch uint8
tensor) together with the uploaded plan tables. ``execute`` runs the two sweep kernels
on a packed stream already in HBM. ``SweepResult`` holds the DM x time plane on the
device and converts rows to reference ``TimeSeries``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from sigpyproc3_amd import _lib
from sigpyproc3_amd.core.kernels import good_size_real

DEFAULT_WIDTHS = (1, 2, 4, 8, 16, 32, 64, 128, 256)  # get_box_width_spacing(256, 2.0)


class SweepPlan:
    """Host plan + device workspace of one sweep configuration."""

    def __init__(self, table: np.ndarray, nsamps: int, nbits: int, big: int, *, stream=None,
                 allocate: bool = True) -> None:
        lib = _lib.load()
        _lib.require_gpu()
        self.table = np.ascontiguousarray(np.atleast_2d(table), dtype=np.int32)
        self.ndm, self.nchans = self.table.shape
        self.nsamps = int(nsamps)
        self.nbits = int(nbits)
        self.big = int(big)
        handle = ctypes.c_void_p()
        _lib.check(lib.spp_sweep_plan_create(_lib.ptr(self.table), self.ndm, self.nchans, self.nsamps,
                                             self.nbits, self.big, ctypes.byref(handle)))
        self._handle = handle
        info = _lib.SweepInfo()
        _lib.check(lib.spp_sweep_plan_info(self._handle, ctypes.byref(info)))
        self.info = info
        rl = np.zeros(self.ndm, dtype=np.int64)
        _lib.check(lib.spp_sweep_plan_row_lengths(self._handle, _lib.ptr(rl)))
        self.row_len = rl
        self.workspace = None
        if allocate:
            self.allocate(stream=stream)

    def allocate(self, stream=None) -> None:
        """Allocate the device workspace (tables + stage-1 copy) and upload the tables."""
        if self.workspace is not None:
            return
        nbytes = int(self.info.workspace_bytes)
        self.workspace = _lib.empty((nbytes,), "uint8")
        _lib.check(_lib.load().spp_sweep_plan_upload(self._handle, _lib.ptr(self.workspace), nbytes,
                                                     _lib.stream_handle(stream)))

    def release(self) -> None:
        """Free the device workspace (a batch holds its own copy of the plan's tables)."""
        self.workspace = None

    @property
    def max_len(self) -> int:
        return int(self.info.max_len)

    @property
    def variant(self) -> int:
        """Block shape of the stage-2 kernel (see ``spp_sweep_plan_set_variant``)."""
        return int(self.info.mode) // 64

    @variant.setter
    def variant(self, v: int) -> None:
        lib = _lib.load()
        _lib.check(lib.spp_sweep_plan_set_variant(self._handle, int(v)))
        _lib.check(lib.spp_sweep_plan_info(self._handle, ctypes.byref(self.info)))

    @staticmethod
    def variant_count() -> int:
        return int(_lib.load().spp_sweep_variant_count())

    def tune(self, packed, plane=None, repeats: int = 2, cache: bool = True, stream=None) -> int:
        """Time every stage-2 block shape on this plan and keep the fastest.

        The choice is cached in a JSON file keyed by (device, nbits, nchans, ndm, nsamps)
        (``SPP_TUNE_CACHE``, default ``~/.cache/sigpyproc3_amd/sweep_tune.json``); a cached
        key is applied without timing. Returns the chosen variant.
        """
        from sigpyproc3_amd import tune

        return tune.tune_plan(self, packed, plane=plane, repeats=repeats, cache=cache, stream=stream)

    @property
    def total_samples(self) -> int:
        """DM-trial samples one execution produces (sum of row lengths)."""
        return int(self.info.total_samples)

    def build(self, packed, stream=None) -> None:
        """Stage 1 only: the channel-major interleaved copy of ``packed`` (workspace)."""
        self._check_packed(packed)
        _lib.call("spp_sweep_build", self._handle, _lib.ptr(packed), _lib.ptr(self.workspace),
                  int(self.info.workspace_bytes), _lib.stream_handle(stream))

    def run(self, plane, ld: int | None = None, stream=None):
        """Stage 2 only: the DM-tiled sums from the staged copy into ``plane``."""
        ld = self.max_len if ld is None else int(ld)
        _lib.call("spp_sweep_run", self._handle, _lib.ptr(plane), ld, _lib.ptr(self.workspace),
                  int(self.info.workspace_bytes), _lib.stream_handle(stream))
        return plane

    def _check_packed(self, packed) -> None:
        need = self.nsamps * self.nchans * self.nbits // 8
        if int(packed.numel()) < need:
            msg = f"packed stream holds {int(packed.numel())} bytes < {need}"
            raise ValueError(msg)

    def execute(self, packed, plane=None, ld: int | None = None, stream=None):
        """Sweep the packed time-major stream ``packed`` (device uint8, nsamps*nchans*nbits/8
        bytes); returns the float32 plane ``[ndm, ld]`` (allocated if not given)."""
        self._check_packed(packed)
        ld = self.max_len if ld is None else int(ld)
        if plane is None:
            plane = _lib.empty((self.ndm, ld), "float32")
        elif tuple(plane.shape) != (self.ndm, ld) or not plane.is_contiguous():
            msg = f"plane must be a contiguous [{self.ndm}, {ld}] float32 tensor"
            raise ValueError(msg)
        _lib.call("spp_sweep_execute", self._handle, _lib.ptr(packed), _lib.ptr(plane), ld,
                  _lib.ptr(self.workspace), int(self.info.workspace_bytes), _lib.stream_handle(stream))
        return plane

    def row_len_for(self, nsamps: int) -> np.ndarray:
        """Row lengths of an execution on the first ``nsamps`` samples."""
        return self.row_len - (self.nsamps - int(nsamps))

    def execute_n(self, packed, nsamps: int, plane, ld: int | None = None, stream=None):
        """``execute`` on the first ``nsamps`` samples of ``packed`` (max delay < nsamps <=
        the plan's nsamps): rows are ``row_len_for(nsamps)`` long. ``plane`` is a contiguous
        float32 tensor of at least ``ndm * ld`` elements, ``ld`` >= the longest row."""
        nsamps = int(nsamps)
        need = nsamps * self.nchans * self.nbits // 8
        if int(packed.numel()) < need:
            msg = f"packed stream holds {int(packed.numel())} bytes < {need}"
            raise ValueError(msg)
        ld = self.max_len - (self.nsamps - nsamps) if ld is None else int(ld)
        if int(plane.numel()) < self.ndm * ld or not plane.is_contiguous():
            msg = f"plane must be a contiguous float32 tensor of >= {self.ndm} x {ld} elements"
            raise ValueError(msg)
        _lib.call("spp_sweep_execute_n", self._handle, _lib.ptr(packed), nsamps, _lib.ptr(plane), ld,
                  _lib.ptr(self.workspace), int(self.info.workspace_bytes), _lib.stream_handle(stream))
        return plane

    def __del__(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.load().spp_sweep_plan_destroy(h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._handle = None


def load_packed_stream(fil, start: int, nsamps: int):
    """Packed bytes of samples ``[start, start + nsamps)`` of a reader, on the device."""
    t = _lib.torch()
    nbytes = nsamps * fil.samp_stride
    host = t.empty(nbytes, dtype=t.uint8, pin_memory=True)
    if hasattr(fil, "read_packed_to_device"):  # pool-read pieces, each copied while the next is read
        from sigpyproc3_amd.io.fileio import read_pool

        dev = _lib.empty((nbytes,), "uint8")
        got = fil.read_packed_to_device(host, dev, start, nsamps, read_pool())
        if got != nbytes:
            msg = f"short read: {got} != {nbytes}"
            raise ValueError(msg)
        return dev
    buf = host.numpy()
    got = 0
    for _nr, _ii, raw in fil.read_plan_packed(gulp=nsamps, start=start, nsamps=nsamps,
                                              allocator=lambda n: buf):
        got = raw.size
    if got != nbytes:
        msg = f"short read: {got} != {nbytes}"
        raise ValueError(msg)
    return host.to(_lib.device(), non_blocking=False)


def dedisperse_resident(packed, header, dm: float, ref_freq: str | float = "fch1", *, stream=None):
    """``Filterbank.dedisperse(dm, ref_freq)`` (base.py:360-424) of a whole stream already in
    HBM (``packed``: the file's packed samples, e.g. from ``load_packed_stream``), as one
    ``spp_dedisperse`` launch over the stream instead of one per gulp.

    With non-negative delays every output sample sums the same inputs whatever the gulp
    schedule, so one launch gives the reference's time series bit for bit. Negative delays
    (an ascending band, or a reference frequency below the top) read before each gulp and
    wrap inside it (numba's negative indexing, kernels.py:361), so those keep the reference's
    gulp schedule (gulp 16384, skip-back of the largest delay) over the resident bytes.
    """
    from sigpyproc3_amd.readers import plan_blocks
    from sigpyproc3_amd.timeseries import TimeSeries

    _lib.require_gpu()
    hdr = header
    nbits = int(hdr.nbits)
    if nbits not in (1, 2, 4, 8):
        msg = "dedisperse_resident takes 1/2/4/8-bit data"
        raise ValueError(msg)
    big = 1 if nbits in (2, 4) else 0
    delays = np.ascontiguousarray(hdr.get_dmdelays(dm, ref_freq), dtype=np.int32)
    md = int(np.abs(delays).max())
    n = int(hdr.nsamples)
    tim_len = n - md
    if tim_len < 0:
        msg = "negative dimensions are not allowed"
        raise ValueError(msg)
    row_bytes = hdr.nchans * nbits // 8
    if int(packed.numel()) < n * row_bytes:
        msg = f"packed stream holds {int(packed.numel())} bytes < {n * row_bytes}"
        raise ValueError(msg)
    tim = _lib.zeros((max(tim_len, 1),), "float32")
    d_dev = _lib.to_device(delays)
    st = _lib.stream_handle(stream)
    if tim_len > 0 and int(delays.min()) >= 0:
        _lib.call("spp_dedisperse", _lib.ptr(packed), nbits, big, hdr.nchans, n, _lib.ptr(d_dev), md,
                  _lib.ptr(tim), 0, st)
    elif tim_len > 0:
        gulp = max(2 * md, 16384)
        _, blocks = plan_blocks(n, gulp, md)
        pos = 0
        for ii, nr, skip in blocks:
            _lib.call("spp_dedisperse", _lib.ptr(packed[pos * row_bytes :]), nbits, big, hdr.nchans, nr,
                      _lib.ptr(d_dev), md, _lib.ptr(tim), ii * (gulp - md), st)
            pos += nr + skip
    data = tim[:tim_len].cpu().numpy() if tim_len > 0 else np.zeros(0, dtype=np.float32)
    return TimeSeries(data, hdr.new_header({"nchans": 1, "dm": dm, "nsamples": tim_len}))


def running_normalise(x, row_len, window: int = 4096, *, ld: int | None = None, stream=None):
    """Running mean/std normalisation of each row (``spp_running_normalise``).

    ``x`` is a device tensor or host array ``[nrows, >= max(row_len)]``; returns the same
    kind. Rows are normalised over their first ``row_len[i]`` samples.
    """
    host = isinstance(x, np.ndarray)
    xd = _lib.to_device(np.ascontiguousarray(x, dtype=np.float32)) if host else x.contiguous()
    nrows, ldx = int(xd.shape[0]), int(xd.shape[1])
    rl = np.ascontiguousarray(row_len, dtype=np.int64)
    z = _lib.empty((nrows, ldx if ld is None else ld), "float32")
    need = np.zeros(1, dtype=np.int64)
    _lib.call("spp_running_normalise_workspace_size", nrows, _lib.ptr(rl), int(window),
              need.ctypes.data_as(_lib._P64))
    ws = _lib.empty((int(need[0]),), "uint8")
    _lib.call("spp_running_normalise", _lib.ptr(xd), ldx, nrows, _lib.ptr(rl), int(window),
              _lib.ptr(z), int(z.shape[1]), _lib.ptr(ws), int(need[0]), _lib.stream_handle(stream))
    return z.cpu().numpy() if host else z


def normalise_boxcar_mf(x, row_len, window: int = 4096, widths=DEFAULT_WIDTHS, *, stream=None):
    """``boxcar_mf(running_normalise(x, row_len, window), row_len, widths)`` in one pass over
    ``x`` (``spp_normalise_boxcar_mf``): the normalised series stays on chip.

    Returns host arrays ``(snr float32, peak int32, width_index int32)``.
    """
    xd = _lib.to_device(np.ascontiguousarray(x, dtype=np.float32)) if isinstance(x, np.ndarray) else x
    if not xd.is_contiguous():
        xd = xd.contiguous()
    nrows, ldx = int(xd.shape[0]), int(xd.shape[1])
    rl = np.ascontiguousarray(row_len, dtype=np.int64)
    w = np.ascontiguousarray(widths, dtype=np.int32)
    need = np.zeros(1, dtype=np.int64)
    _lib.call("spp_normalise_boxcar_mf_workspace_size", nrows, _lib.ptr(rl), _lib.ptr(w), int(w.size),
              need.ctypes.data_as(_lib._P64))
    ws = _lib.empty((int(need[0]),), "uint8")
    out = _lib.empty((3, nrows), "int32")  # float32 snr bits, peak, width index
    _lib.call("spp_normalise_boxcar_mf", _lib.ptr(xd), ldx, nrows, _lib.ptr(rl), int(window), _lib.ptr(w),
              int(w.size), _lib.ptr(out[0]), _lib.ptr(out[1]), _lib.ptr(out[2]), _lib.ptr(ws), int(need[0]),
              _lib.stream_handle(stream))
    host = out.cpu().numpy()
    return host[0].view(np.float32).copy(), host[1].copy(), host[2].copy()


def boxcar_mf(z, row_len, widths=DEFAULT_WIDTHS, *, stream=None):
    """Boxcar matched filter + per-row argmax (``spp_boxcar_mf``).

    Returns host arrays ``(snr float32, peak int32, width_index int32)``.
    """
    zd = _lib.to_device(np.ascontiguousarray(z, dtype=np.float32)) if isinstance(z, np.ndarray) else z
    nrows, ldz = int(zd.shape[0]), int(zd.shape[1])
    rl = np.ascontiguousarray(row_len, dtype=np.int64)
    w = np.ascontiguousarray(widths, dtype=np.int32)
    need = np.zeros(1, dtype=np.int64)
    _lib.call("spp_boxcar_mf_workspace_size", nrows, _lib.ptr(rl), _lib.ptr(w), int(w.size),
              need.ctypes.data_as(_lib._P64))
    ws = _lib.empty((int(need[0]),), "uint8")
    snr = _lib.empty((nrows,), "float32")
    peak = _lib.empty((nrows,), "int32")
    widx = _lib.empty((nrows,), "int32")
    _lib.call("spp_boxcar_mf", _lib.ptr(zd), ldz, nrows, _lib.ptr(rl), _lib.ptr(w), int(w.size),
              _lib.ptr(snr), _lib.ptr(peak), _lib.ptr(widx), _lib.ptr(ws), int(need[0]),
              _lib.stream_handle(stream))
    return snr.cpu().numpy(), peak.cpu().numpy(), widx.cpu().numpy()


class SweepResult:
    """A DM x time plane on the device plus the row metadata."""

    def __init__(self, plane, row_len, dms, header, plan=None, table=None) -> None:
        self.plane = plane
        self.row_len = np.asarray(row_len, dtype=np.int64)
        self.dms = np.asarray(dms)
        self.header = header
        self.plan = plan
        self.table = table

    @property
    def ndm(self) -> int:
        return int(self.plane.shape[0])

    def row(self, i: int) -> np.ndarray:
        return self.plane[i, : int(self.row_len[i])].cpu().numpy()

    def timeseries(self, i: int):
        """Row ``i`` as the ``TimeSeries`` ``Filterbank.dedisperse(dms[i])`` returns."""
        from sigpyproc3_amd.timeseries import TimeSeries

        n = int(self.row_len[i])
        return TimeSeries(self.row(i), self.header.new_header({"nchans": 1, "dm": float(self.dms[i]),
                                                                "nsamples": n}))

    def matched_filter(self, window: int = 4096, widths=DEFAULT_WIDTHS):
        """Running normalisation over ``window`` samples, then boxcar MF per DM row, in one
        pass over the plane (``normalise_boxcar_mf``).

        Returns ``(snr, peak_bin, width_index)`` host arrays of length ndm.
        """
        return normalise_boxcar_mf(self.plane, self.row_len, window, widths)


def mf_good_sizes(row_len) -> np.ndarray:
    return np.array([good_size_real(int(n)) for n in row_len], dtype=np.int64)
