"""1-D time series container (mirrors ``sigpyproc.timeseries.TimeSeries``)."""

from __future__ import annotations

import numpy as np

from sigpyproc3_amd.core import stats


class TimeSeries:
    """A float32 time series and its header (timeseries.py:27-47, 484-499)."""

    def __init__(self, data, header) -> None:
        self._data = np.asarray(data, dtype=np.float32)
        self._header = header
        self._check_input()

    @property
    def data(self) -> np.ndarray:
        return self._data

    @property
    def header(self):
        return self._header

    @property
    def nsamples(self) -> int:
        return len(self.data)

    def normalise(self, loc_method: str = "mean", scale_method: str = "std") -> TimeSeries:
        """Subtract the location and divide by the scale (timeseries.py:82-105)."""
        z = stats.estimate_zscore(self.data, loc_method, scale_method)
        return TimeSeries(z.data, self.header.new_header())

    def running_normalise(self, window: int = 4096) -> TimeSeries:
        """Normalise by the running mean/std over ``window`` samples on the GPU.

        No reference counterpart (the pre-step of the DM-sweep matched filter); see
        ``sigpyproc3_amd.sweep.running_normalise``.
        """
        from sigpyproc3_amd.sweep import running_normalise

        z = running_normalise(self.data[None, :], np.array([self.nsamples]), window)
        return TimeSeries(z[0], self.header.new_header())

    def _check_input(self) -> None:
        from sigpyproc3_amd.header import Header

        if not isinstance(self.header, Header):
            msg = "Input header is not a Header instance"
            raise TypeError(msg)
        if self.data.ndim != 1:
            msg = "Input data is not 1D"
            raise ValueError(msg)
        if self.data.size == 0:
            msg = "Input data is empty"
            raise ValueError(msg)
        if len(self.data) != self.header.nsamples:
            msg = (f"Input data length ({len(self.data)}) does not match header nsamples "
                   f"({self.header.nsamples})")
            raise ValueError(msg)
