"""In-memory filterbank blocks (mirrors ``sigpyproc.block``, block.py:25-425).

A block is float32 ``[nchans, nsamples]`` (channel-major) with a header. ``dedisperse``
rolls every channel by its negative delay and ``dmt_transform`` sums channels rolled by
their positive delays at each of ``dmsteps`` DMs (the sign the reference uses,
core/kernels.py:1155-1193); both run on the GPU through ``spp_roll_rows`` / ``spp_dmt``.
The roll offsets are built here on the host from the host delay table.
"""

from __future__ import annotations

import numpy as np

from sigpyproc3_amd import _lib
from sigpyproc3_amd.header import Header


def _roll_offsets(shifts: np.ndarray, ncols: int, *, valid: bool) -> tuple[np.ndarray, int]:
    """Column offsets that read a roll by ``shifts`` (kernels.roll_block[_valid],
    core/kernels.py:1059-1152) and the output width.

    Circular: ``res[t] = arr[(t - s) mod n]`` -> offset ``-s mod n``, width n. Valid:
    ``res[t] = arr[start - s + t]`` with ``start = max(0, max s)`` and width
    ``n + min(0, min s) - start`` (ValueError if not positive)."""
    s = np.asarray(shifts, dtype=np.int64)
    if not valid:
        return np.mod(-s, ncols), ncols
    start = max(0, int(s.max()))
    end = ncols + min(0, int(s.min()))
    width = end - start
    if width <= 0:
        msg = f"Not enough samples. Required at least {start - min(0, int(s.min()))} samples, given {ncols}."
        raise ValueError(msg)
    return start - s, width


class BaseBlock:
    """A 2-D float32 block with a header (block.py:25-165)."""

    def __init__(self, data: np.ndarray, header: Header) -> None:
        self._data = np.asarray(data, dtype=np.float32)
        self._header = header
        self._check_input()

    @property
    def data(self) -> np.ndarray:
        return self._data

    @property
    def header(self) -> Header:
        return self._header

    @property
    def nsamples(self) -> int:
        return int(self.data.shape[1])

    def _check_input(self) -> None:
        if not isinstance(self.header, Header):
            msg = "Input header is not a Header instance"
            raise TypeError(msg)
        if self.data.ndim != 2:
            msg = "Input data is not 2 dimensional"
            raise ValueError(msg)
        if self.nsamples != self.header.nsamples:
            msg = (f"Input data length ({self.nsamples}) does not match header nsamples "
                   f"({self.header.nsamples})")
            raise ValueError(msg)


class FilterbankBlock(BaseBlock):
    """Filterbank data ``[nchans, nsamples]`` at a given DM (block.py:167-367)."""

    def __init__(self, data: np.ndarray, header: Header, dm: float = 0) -> None:
        super().__init__(data, header)
        self._dm = dm

    @property
    def dm(self) -> float:
        return self._dm

    @property
    def nchans(self) -> int:
        return int(self.data.shape[0])

    def get_tim(self) -> np.ndarray:
        """Sum over channels (block.py:239-246)."""
        return self.data.sum(axis=0)

    def get_bandpass(self) -> np.ndarray:
        """Sum over time (block.py:248-265)."""
        return self.data.sum(axis=1)

    def dedisperse(self, dm: float, ref_freq: str = "fch1", *,
                   only_valid_samples: bool = False) -> FilterbankBlock:
        """Roll each channel by minus its delay (block.py:267-307), on the GPU."""
        delays = np.atleast_1d(self.header.get_dmdelays(dm, ref_freq=ref_freq)).astype(np.int64)
        off, width = _roll_offsets(-delays, self.nsamples, valid=only_valid_samples)
        out = _run_roll(self.data, off, width)
        return FilterbankBlock(out, self.header.new_header({"nsamples": out.shape[1]}), dm)

    def dmt_transform(self, dm: float, dmsteps: int = 512, ref_freq: str | float = "fch1", *,
                      only_valid_samples: bool = False) -> DMTBlock:
        """DM-time transform over ``dm + linspace(-dm, dm, dmsteps)`` (block.py:309-343).

        Each DM sums the channels rolled by **plus** their delays, as kernels.dmt_block does
        (the opposite direction of ``dedisperse``). With ``only_valid_samples`` the output
        width comes from the whole delay table, and every DM's own valid range must have
        that width, as in kernels.dmt_block_valid (ValueError otherwise).
        """
        dm_arr = dm + np.linspace(-dm, dm, dmsteps)
        dm_delays = np.atleast_2d(self.header.get_dmdelays(dm_arr, ref_freq=ref_freq)).astype(np.int64)
        if dm_delays.shape[1] != self.nchans:
            msg = "Number of chans must be same in both arrays."
            raise ValueError(msg)
        n = self.nsamples
        if only_valid_samples:
            _, width = _roll_offsets(dm_delays.reshape(-1), n, valid=True)
            offs = []
            for row in dm_delays:
                off, w = _roll_offsets(row, n, valid=True)
                if w != width:
                    msg = (f"could not broadcast input array from shape ({w},) into shape "
                           f"({width},)")
                    raise ValueError(msg)
                offs.append(off)
            off2 = np.stack(offs)
        else:
            width = n
            off2 = np.mod(-dm_delays, n)
        out = _run_dmt(self.data, off2, width)
        # The header keeps nsamples (block.py:343), so a valid-samples transform narrower
        # than the block fails the length check, as the reference's does.
        return DMTBlock(out, self.header.new_header({"nchans": 1}), dm_arr)


class DMTBlock(BaseBlock):
    """DM-time transform ``[ndms, nsamples]`` (block.py:370-425)."""

    def __init__(self, data: np.ndarray, header: Header, dms: np.ndarray) -> None:
        super().__init__(data, header)
        self._dms = np.asarray(dms, dtype=np.float32)
        if self.ndms != self.dms.size:
            msg = f"Number of DMs ({self.ndms}) does not match number of DM values ({self.dms.size})"
            raise ValueError(msg)

    @property
    def dms(self) -> np.ndarray:
        return self._dms

    @property
    def ndms(self) -> int:
        return int(self.data.shape[0])


def _run_roll(data: np.ndarray, off: np.ndarray, width: int) -> np.ndarray:
    _lib.require_gpu()
    nrows, ncols = data.shape
    src = _lib.to_device(np.ascontiguousarray(data, dtype=np.float32))
    off_h = np.ascontiguousarray(off, dtype=np.int64)
    off_d = _lib.to_device(off_h)
    out = _lib.empty((nrows, width), "float32")
    for r0 in range(0, nrows, 65535):
        r1 = min(nrows, r0 + 65535)
        _lib.call("spp_roll_rows", _lib.ptr(src[r0:r1]), r1 - r0, ncols, _lib.ptr(off_h[r0:r1]),
                  _lib.ptr(off_d[r0:r1]), width, _lib.ptr(out[r0:r1]), _lib.stream_handle())
    return out.cpu().numpy()


def _run_dmt(data: np.ndarray, off: np.ndarray, width: int) -> np.ndarray:
    _lib.require_gpu()
    nchans, ncols = data.shape
    ndm = off.shape[0]
    src = _lib.to_device(np.ascontiguousarray(data, dtype=np.float32))
    off_h = np.ascontiguousarray(off, dtype=np.int64)
    off_d = _lib.to_device(off_h)
    out = _lib.empty((ndm, width), "float32")
    for i0 in range(0, ndm, 65535):
        i1 = min(ndm, i0 + 65535)
        _lib.call("spp_dmt", _lib.ptr(src), nchans, ncols, _lib.ptr(off_h[i0:i1]), _lib.ptr(off_d[i0:i1]),
                  i1 - i0, width, _lib.ptr(out[i0:i1]), _lib.stream_handle())
    return out.cpu().numpy()
