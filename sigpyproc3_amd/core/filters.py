"""Matched filter (mirrors ``sigpyproc.core.filters``): templates on the host, the
template-bank convolution on the GPU (``spp_convolve_templates``)."""

from __future__ import annotations

import attrs
import numpy as np

from sigpyproc3_amd.core import kernels
from sigpyproc3_amd.core.stats import ZScoreResult, estimate_zscore

GAUSSIAN_FWHM_TO_SIGMA = 1.0 / (8.0 * np.log(2.0)) ** 0.5


@attrs.define(auto_attribs=True, slots=True, frozen=True)
class Template:
    """1-D pulse template (filters.py:368-587)."""

    data: np.ndarray
    width: float
    ref_bin: int = attrs.field(default=0, validator=[attrs.validators.instance_of(int | np.integer),
                                                     attrs.validators.ge(0)])
    ref: str = attrs.field(default="start", validator=attrs.validators.in_({"start", "peak"}))
    kind: str = attrs.field(default="custom", validator=attrs.validators.instance_of(str))

    def __attrs_post_init__(self) -> None:
        if not self.data.size:
            msg = "Empty data array is not supported."
            raise ValueError(msg)
        if self.data.ndim != 1:
            msg = f"Only 1D data is supported, got {self.data.ndim}."
            raise ValueError(msg)
        if self.ref_bin >= self.data.size:
            msg = f"Reference bin {self.ref_bin} is out of bounds."
            raise ValueError(msg)

    def get_model(self, peak_bin: int, nbins: int) -> np.ndarray:
        padded = np.pad(self.data, (0, nbins - self.data.size))
        return np.roll(kernels.normalize_template(padded), peak_bin - self.ref_bin)

    def get_on_pulse(self, peak_bin: int, nbins: int) -> tuple[int, int]:
        if self.ref == "start":
            left, right = peak_bin, peak_bin + self.width
        else:
            left, right = peak_bin - round(self.width), peak_bin + round(self.width)
        return (max(0, left), int(min(nbins, right)))

    @classmethod
    def gen_boxcar(cls, width: int) -> Template:
        width = int(width)
        if width <= 0:
            msg = f"Width {width} must be greater than 0."
            raise ValueError(msg)
        return cls(np.ones(width, dtype=np.float32), width, ref_bin=0, ref="start", kind="boxcar")

    @classmethod
    def gen_gaussian(cls, width: float, extent: float = 3.5) -> Template:
        if width <= 0:
            msg = f"Width {width} must be greater than 0."
            raise ValueError(msg)
        stddev = GAUSSIAN_FWHM_TO_SIGMA * width
        size = int(np.ceil(extent * stddev))
        x = np.arange(-size, size + 1)
        arr = np.exp(-0.5 * x**2 / stddev**2)
        return cls(arr, width, ref_bin=len(x) // 2, ref="peak", kind="gaussian")

    @classmethod
    def gen_lorentzian(cls, width: float, extent: float = 3.5) -> Template:
        if width <= 0:
            msg = f"Width {width} must be greater than 0."
            raise ValueError(msg)
        gamma = width / 2
        size = int(np.ceil(extent * gamma))
        x = np.arange(-size, size + 1)
        arr = gamma**2 / (x**2 + gamma**2)
        return cls(arr, width, ref_bin=len(x) // 2, ref="peak", kind="lorentzian")

    def __str__(self) -> str:
        return f"Template(size={self.data.size}, kind={self.kind}, width={self.width:.3f})"

    __repr__ = __str__


class MatchedFilter:
    """Matched-filter pulse search on a 1-D series (filters.py:29-365)."""

    def __init__(self, data, loc_method: str = "median", scale_method: str = "iqr",
                 temp_kind: str = "boxcar", nbins_max: int = 32, spacing_factor: float = 1.5) -> None:
        data = np.asarray(data)
        if data.ndim != 1:
            msg = f"Data dimension {data.ndim} is not supported."
            raise ValueError(msg)
        self._temp_kind = temp_kind
        self._data = np.asarray(data, dtype=np.float32)
        self._zscores = estimate_zscore(self.data, loc_method=loc_method, scale_method=scale_method)
        self._setup_templates(nbins_max, spacing_factor)
        self._compute()

    data = property(lambda self: self._data)
    zscores: ZScoreResult = property(lambda self: self._zscores)
    temp_kind = property(lambda self: self._temp_kind)
    temp_widths = property(lambda self: self._temp_widths)
    temp_bank = property(lambda self: self._temp_bank)
    convs = property(lambda self: self._convs)
    best_temp = property(lambda self: self._best_temp)
    snr = property(lambda self: self._best_snr)

    @property
    def peak_bin(self) -> int:
        return int(self._peak_bin)

    @property
    def best_model(self) -> np.ndarray:
        return (self.best_temp.get_model(self.peak_bin, self.data.size) * self.snr
                * self.zscores.scale + self.zscores.loc)

    @property
    def on_pulse(self) -> tuple[int, int]:
        return self.best_temp.get_on_pulse(self.peak_bin, self.data.size)

    def _setup_templates(self, nbins_max: int, spacing_factor: float) -> None:
        if self.temp_kind == "boxcar":
            self._temp_widths = self.get_box_width_spacing(nbins_max, spacing_factor)
        else:
            if spacing_factor <= 1:
                msg = "Spacing factor must be greater than 1 for non-boxcar templates."
                raise ValueError(msg)
            npoints = int(np.ceil(np.log(nbins_max) / np.log(spacing_factor))) + 1
            self._temp_widths = np.geomspace(1, nbins_max, npoints)
        bank = []
        for width in self.temp_widths:
            temp = getattr(Template, f"gen_{self.temp_kind}")(width)
            if temp.data.size > self.data.size:
                msg = f"Template size ({temp.data.size}) is larger than the data size({self.data.size})."
                raise ValueError(msg)
            bank.append(temp)
        self._temp_bank = bank

    def _compute(self) -> None:
        self._convs = kernels.convolve_templates(self.zscores.data, [t.data for t in self.temp_bank],
                                                 [t.ref_bin for t in self.temp_bank])
        self._itemp, self._peak_bin = np.unravel_index(self._convs.argmax(), self._convs.shape)
        self._best_temp = self.temp_bank[self._itemp]
        self._best_snr = self._convs[self._itemp, self._peak_bin]

    @staticmethod
    def get_box_width_spacing(size_max: int, spacing_factor: float = 1.5) -> np.ndarray:
        """Boxcar widths (filters.py:340-365)."""
        widths = [1]
        while widths[-1] < size_max:
            nxt = int(max(widths[-1] + 1, spacing_factor * widths[-1]))
            if nxt > size_max:
                break
            widths.append(nxt)
        return np.array(widths, dtype=np.float32)
