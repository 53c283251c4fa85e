"""Folded data cubes (mirrors ``sigpyproc.foldedcube``, foldedcube.py:23-352) and the host
side of folding: the cell key of every sample, computed with the reference's arithmetic.

``Filterbank.fold`` lists each gulp's samples by (subintegration, phase bin) with
``fold_cells`` and runs ``spp_fold`` on the GPU; the cube and its slices are host numpy.
"""

from __future__ import annotations

import numpy as np

from sigpyproc3_amd import params
from sigpyproc3_amd.header import Header

SPEED_OF_LIGHT = 299792458.0  # m/s (core/kernels.py:17)


def float_floordiv(vx: np.ndarray, wx: float) -> np.ndarray:
    """Elementwise ``vx // wx`` for floats as Python and numba evaluate it: remainder by
    fmod, quotient ``(vx - mod) / wx`` stepped down when the remainder's sign differs from
    the divisor's, then snapped to the nearest integer (CPython float_divmod)."""
    vx = np.asarray(vx, dtype=np.float64)
    mod = np.fmod(vx, wx)
    div = (vx - mod) / wx
    div = np.where((mod != 0) & ((wx < 0) != (mod < 0)), div - 1.0, div)
    flo = np.floor(div)
    flo = np.where(div - flo > 0.5, flo + 1.0, flo)
    return np.where(div != 0, flo, 0.0)


def fold_cells(index: int, nvalid: int, total_nsamps: int, tsamp: float, period: float,
               accel: float, nbins: int, nints: int) -> tuple[np.ndarray, np.ndarray]:
    """(subint, phase bin) of samples ``index .. index + nvalid`` (kernels.fold,
    core/kernels.py:425-437). tsamp, period and accel enter as float32 values (the kernel's
    signature) and everything else is float64, in the reference's operation order."""
    ts = float(np.float32(tsamp))
    per = float(np.float32(period))
    acc = float(np.float32(accel))
    k = np.arange(index, index + nvalid, dtype=np.int64).astype(np.float64)
    tj = k * ts
    tobs = float(total_nsamps) * ts
    phase = ((nbins * tj) * (1.0 + (acc * (tj - tobs)) / (2 * SPEED_OF_LIGHT))) / per + 0.5
    phasebin = np.abs(np.trunc(phase).astype(np.int64)) % nbins
    subint = float_floordiv(k, total_nsamps / nints).astype(np.int64)
    return subint, phasebin


def band_starts(nchans: int, nsubs: int) -> np.ndarray:
    """First channel of each subband, ``ichan // (nchans / nsubs)`` (core/kernels.py:439-441),
    plus nchans at the end."""
    band = float_floordiv(np.arange(nchans, dtype=np.float64), nchans / nsubs).astype(np.int64)
    return np.searchsorted(band, np.arange(nsubs + 1), side="left").astype(np.int32)


class Profile:
    """A 1-D pulse profile (foldedcube.py:23-62)."""

    def __init__(self, data, tsamp: float) -> None:
        self._data = np.asarray(data, dtype=np.float32)
        self._tsamp = tsamp

    @property
    def data(self) -> np.ndarray:
        return self._data

    @property
    def tsamp(self) -> float:
        return self._tsamp

    def compute_mf(self, temp_kind: str = "boxcar", nbins_max: int = 32, spacing_factor: float = 1.5):
        from sigpyproc3_amd.core.filters import MatchedFilter

        return MatchedFilter(self.data, temp_kind=temp_kind, nbins_max=nbins_max,
                             spacing_factor=spacing_factor)


class FoldSlice:
    """A 2-D (rows x phase) slice of a folded cube (foldedcube.py:65-118)."""

    def __init__(self, data, tsamp: float) -> None:
        self._data = np.asarray(data, dtype=np.float32)
        self._tsamp = tsamp

    @property
    def data(self) -> np.ndarray:
        return self._data

    @property
    def tsamp(self) -> float:
        return self._tsamp

    @property
    def nbins(self) -> int:
        return int(self.data.shape[1])

    def normalize(self) -> FoldSlice:
        return FoldSlice(self.data / np.mean(self.data, axis=1, keepdims=True), self.tsamp)

    def get_profile(self) -> Profile:
        return Profile(np.sum(self.data, axis=0), self.tsamp)


class FoldedData:
    """Folded cube ``[nsubints, nsubbands, nbins]`` (foldedcube.py:121-352)."""

    def __init__(self, data, hdr: Header, period: float, dm: float, accel: float = 0) -> None:
        self._data = np.asarray(data, dtype=np.float32)
        self._hdr = hdr
        self._period = period
        self._dm = dm
        self._accel = accel
        if not isinstance(hdr, Header):
            msg = "Input header is not a Header instance"
            raise TypeError(msg)
        if self._data.ndim != 3:
            msg = "Input data is not 3 dimensional"
            raise ValueError(msg)
        self._tph_shifts = np.zeros(self.nsubints, dtype=np.int32)
        self._fph_shifts = np.zeros(self.nsubbands, dtype=np.int32)

    @property
    def data(self) -> np.ndarray:
        return self._data

    @property
    def header(self) -> Header:
        return self._hdr

    @property
    def period(self) -> float:
        return self._period

    @property
    def dm(self) -> float:
        return self._dm

    @property
    def nsubints(self) -> int:
        return int(self.data.shape[0])

    @property
    def nsubbands(self) -> int:
        return int(self.data.shape[1])

    @property
    def nbins(self) -> int:
        return int(self.data.shape[2])

    def get_subint(self, nsubint: int) -> FoldSlice:
        return FoldSlice(self.data[nsubint], self.header.tsamp)

    def get_subband(self, nsubband: int) -> FoldSlice:
        return FoldSlice(self.data[:, nsubband], self.header.tsamp)

    def get_profile(self) -> Profile:
        return Profile(np.sum(self.data, axis=(0, 1)), self.header.tsamp)

    def get_time_phase(self) -> FoldSlice:
        return FoldSlice(np.sum(self.data, axis=1), self.header.tsamp)

    def get_freq_phase(self) -> FoldSlice:
        return FoldSlice(self.data.sum(axis=0), self.header.tsamp)

    def centre(self) -> FoldedData:
        """Roll the on-pulse region of the profile to the middle (foldedcube.py:260-266)."""
        on = self.get_profile().compute_mf().on_pulse
        pos = int(np.mean(on))
        return FoldedData(np.roll(self.data, -(pos - self.nbins // 2), axis=2), self.header,
                          self.period, self.dm)

    def replace_nan(self) -> None:
        self.data[np.isnan(self.data)] = np.nanmedian(self.data)

    def _roll_rows(self, shifts: np.ndarray, axis_rows: int) -> None:
        for a in range(self.nsubints):
            for b in range(self.nsubbands):
                s = shifts[a] if axis_rows == 0 else shifts[b]
                self.data[a][b] = np.roll(self.data[a][b], -s, axis=0)

    def update_dm(self, dm: float) -> None:
        """Re-align the subbands for a new DM (foldedcube.py:271-287)."""
        self._roll_rows(self._dm_drifts(dm), axis_rows=1)
        self._dm = dm

    def update_period(self, period: float) -> None:
        """Re-align the subintegrations for a new period (foldedcube.py:289-305)."""
        self._roll_rows(self._period_drifts(period), axis_rows=0)
        self._period = period

    def _dm_drifts(self, newdm: float) -> np.ndarray:
        if newdm - self.dm == 0:
            drifts = -1 * self._fph_shifts
            self._fph_shifts.fill(0)
            return drifts
        width = self.header.foff * self.header.nchans / self.nsubbands
        freqs = np.arange(self.nsubbands, dtype=np.float64) * width + self.header.fch1
        drifts = params.compute_dmdelays(freqs, newdm - self.dm, self.period / self.nbins,
                                         self.header.fch1, in_samples=True)
        out = drifts - self._fph_shifts
        self._fph_shifts = drifts
        return out

    def _period_drifts(self, newperiod: float) -> np.ndarray:
        dbins = (newperiod / self._period - 1) * self.header.tobs * self.nbins / self._period
        if dbins == 0:
            drifts = -1 * self._tph_shifts
            self._tph_shifts.fill(0)
            return drifts
        drifts = np.arange(self.nsubints, dtype=np.float32)
        drifts = np.round(drifts / (self.nsubints / dbins)).astype(np.int32)
        out = drifts - self._tph_shifts
        self._tph_shifts = drifts
        return out
