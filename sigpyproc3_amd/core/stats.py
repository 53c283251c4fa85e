"""Location/scale estimates and z-scores (the parts of ``sigpyproc.core.stats`` that the
matched filter uses). Host numpy: these are per-series scalar statistics, computed
exactly as the reference computes them, ahead of the GPU convolution."""

from __future__ import annotations

import attrs
import numpy as np


@attrs.define(auto_attribs=True, slots=True, kw_only=True)
class ZScoreResult:
    """Z-scores and the location/scale used (stats.py:26-49)."""

    data: np.ndarray
    loc: np.ndarray
    scale: np.ndarray


def estimate_loc(data, method: str = "median", axis=None, *, keepdims: bool = False):
    """stats.estimate_loc (stats.py:284-324)."""
    data = np.asanyarray(data)
    if len(data) == 0:
        msg = "Cannot estimate loc from an empty array."
        raise ValueError(msg)
    if method == "mean":
        return np.mean(data, axis=axis, keepdims=keepdims, dtype=np.float64)
    if method == "median":
        return np.median(data, axis=axis, keepdims=keepdims)
    msg = f"Method {method} is not supported for estimating location."
    raise ValueError(msg)


def _scale_iqr(data, axis=None):
    data = np.asanyarray(data, dtype=np.float64)
    norm = 1.3489795003921634  # norm.ppf(0.75) - norm.ppf(0.25)
    p = np.percentile(data, [25, 75], axis=axis, keepdims=True)
    return np.squeeze(np.diff(p, axis=0) / norm)


def _scale_mad(data, axis=None):
    data = np.asanyarray(data, dtype=np.float64)
    norm = 0.6744897501960817  # norm.ppf(0.75)
    loc = np.median(data, axis=axis, keepdims=True)
    mad = np.median(np.abs(data - loc), axis=axis, keepdims=True) / norm
    zero = np.isclose(mad, 0)
    if np.any(zero):
        aad = np.mean(np.abs(data - loc), axis=axis, keepdims=True) / np.sqrt(2 / np.pi)
        mad = np.where(zero, aad, mad)
    return np.squeeze(mad)


_SCALES = {"iqr": _scale_iqr, "mad": _scale_mad}


def estimate_scale(data, method: str = "mad", axis=None, *, keepdims: bool = False):
    """stats.estimate_scale (stats.py:326-415) for std, iqr and mad."""
    data = np.asanyarray(data, dtype=np.float64)
    if len(data) == 0:
        msg = "Cannot estimate scale from an empty array."
        raise ValueError(msg)
    if method == "std":
        return np.std(data, axis=axis, keepdims=keepdims, dtype=np.float64)
    fn = _SCALES.get(method)
    if fn is None:
        msg = f"Method {method} is not supported for estimating scale."
        raise ValueError(msg)
    result = fn(data, axis)
    if keepdims:
        axes = tuple(range(data.ndim)) if axis is None else ((axis,) if isinstance(axis, int) else axis)
        result = np.expand_dims(result, axis=axes)
    if isinstance(result, np.ndarray) and result.size == 1 and not keepdims:
        result = result.flat[0]
    return result


def estimate_zscore(data, loc_method: str = "median", scale_method: str = "mad",
                    axis: int | None = 0) -> ZScoreResult:
    """stats.estimate_zscore (stats.py:418-477)."""
    data = np.asanyarray(data, dtype=np.float32)
    if data.size == 0:
        msg = "Cannot estimate Z-scores from an empty array."
        raise ValueError(msg)
    loc = (np.zeros(1, dtype=data.dtype) if loc_method == "norm"
           else estimate_loc(data, loc_method, axis, keepdims=True))
    scale = (np.ones(1, dtype=data.dtype) if scale_method == "norm"
             else estimate_scale(data, scale_method, axis, keepdims=True))
    zero = np.isclose(scale, 0)
    if np.any(zero):
        scale = np.where(zero, 1, scale)
    z = np.subtract(data, loc, dtype=np.float32)
    np.divide(z, scale, out=z)
    return ZScoreResult(data=z, loc=np.asarray(loc), scale=np.asarray(scale))
