"""GPU operators with the names and signatures of ``sigpyproc.core.kernels``.

The reference calls its numba kernels by name (``getattr(kernels, "unpack2_8_big")``,
``kernels.dedisperse_par if parallel else kernels.dedisperse``, ``kernels.subband``,
``kernels.convolve_templates``). These functions keep those names and argument
meanings and run on the GPU through ``libspp_hip.so``. Arguments may be host numpy
arrays (copied to the device and back; outputs are filled in place as the reference
does) or device tensors (no copies). There is no CPU path.
"""

from __future__ import annotations

import numpy as np

from sigpyproc3_amd import _lib


def _dev(x, dtype: str):
    if isinstance(x, np.ndarray):
        return _lib.to_device(np.ascontiguousarray(x, dtype=np.dtype(dtype)))
    if str(x.dtype) != f"torch.{dtype}":
        msg = f"expected a {dtype} tensor, got {x.dtype}"
        raise TypeError(msg)
    return x.contiguous()


def _size(x) -> int:
    return x.size if isinstance(x, np.ndarray) else int(x.numel())


def _unpack(array, unpacked, nbits: int, big: int) -> None:
    _lib.require_gpu()
    n = _size(array)
    if _size(unpacked) != n * (8 // nbits):
        msg = f"Unpacking array must be {8 // nbits} x input size, got {_size(unpacked)}"
        raise ValueError(msg)
    if n == 0:
        return
    src = _dev(array, "uint8")
    host_out = isinstance(unpacked, np.ndarray)
    dst = _lib.empty((n * (8 // nbits),), "uint8") if host_out else unpacked
    _lib.call("spp_unpack", _lib.ptr(src), n, _lib.ptr(dst), nbits, big, _lib.stream_handle())
    if host_out:
        unpacked[:] = dst.cpu().numpy()


# kernels.unpack{1,2,4}_8_{big,little} (core/kernels.py:20-69)
def unpack1_8_big(array, unpacked) -> None:
    _unpack(array, unpacked, 1, 1)


def unpack1_8_little(array, unpacked) -> None:
    _unpack(array, unpacked, 1, 0)


def unpack2_8_big(array, unpacked) -> None:
    _unpack(array, unpacked, 2, 1)


def unpack2_8_little(array, unpacked) -> None:
    _unpack(array, unpacked, 2, 0)


def unpack4_8_big(array, unpacked) -> None:
    _unpack(array, unpacked, 4, 1)


def unpack4_8_little(array, unpacked) -> None:
    _unpack(array, unpacked, 4, 0)


def max_abs_delay(delays) -> int:
    d = delays if isinstance(delays, np.ndarray) else delays.cpu().numpy()
    return int(np.max(np.abs(np.asarray(d, dtype=np.int64)))) if np.size(d) else 0


def dedisperse(in_arr, out_arr, delays, nchans: int, nsamps: int, out_start_index: int) -> None:
    """kernels.dedisperse (core/kernels.py:340-362) on one gulp of unpacked samples.

    ``in_arr`` is uint8 (one byte per sample) or float32, ``[nsamps][nchans]`` flat;
    ``out_arr[out_start_index + t]`` is written for ``t < nsamps - max|delays|``.
    """
    _lib.require_gpu()
    md = max_abs_delay(delays)
    nout = nsamps - md
    if nout <= 0:
        return
    is_f32 = (in_arr.dtype == np.float32) if isinstance(in_arr, np.ndarray) else str(in_arr.dtype) == "torch.float32"
    nbits = 32 if is_f32 else 8
    src = _dev(in_arr, "float32" if is_f32 else "uint8")
    if _size(src) < nsamps * nchans:
        msg = f"input holds {_size(src)} samples < nsamps*nchans = {nsamps * nchans}"
        raise ValueError(msg)
    d = _dev(np.asarray(delays, dtype=np.int32) if isinstance(delays, np.ndarray) else delays, "int32")
    if _size(out_arr) < out_start_index + nout:
        msg = "output array too short"
        raise ValueError(msg)
    host_out = isinstance(out_arr, np.ndarray)
    if host_out:
        dst = _lib.empty((nout,), "float32")
        base, start = dst, 0
    else:
        base, start = out_arr, out_start_index
    _lib.call("spp_dedisperse", _lib.ptr(src), nbits, 0, nchans, nsamps, _lib.ptr(d), md,
              _lib.ptr(base), start, _lib.stream_handle())
    if host_out:
        out_arr[out_start_index : out_start_index + nout] = dst.cpu().numpy()


# The reference's parallel variant (core/kernels.py:542) computes the same values.
dedisperse_par = dedisperse


def subband(inarray, outarray, delays, chan_to_sub, maxdelay: int, nchans: int, nsubs: int,
            nsamps: int) -> None:
    """kernels.subband (core/kernels.py:379-401) on one gulp.

    The output is zeroed over the written range before accumulating (the reference's
    buffer is never zeroed, base.py:1073-1074; that non-determinism is not reproduced).
    """
    _lib.require_gpu()
    nout = nsamps - maxdelay
    if nout <= 0:
        return
    is_f32 = (inarray.dtype == np.float32) if isinstance(inarray, np.ndarray) else str(inarray.dtype) == "torch.float32"
    src = _dev(inarray, "float32" if is_f32 else "uint8")
    d = _dev(np.asarray(delays, dtype=np.int32) if isinstance(delays, np.ndarray) else delays, "int32")
    cts = _dev(np.asarray(chan_to_sub, dtype=np.int32) if isinstance(chan_to_sub, np.ndarray) else chan_to_sub, "int32")
    host_out = isinstance(outarray, np.ndarray)
    size = _size(outarray)
    dst = _lib.empty((size,), "float32") if host_out else outarray
    if host_out:
        dst.copy_(_lib.torch().from_numpy(np.ascontiguousarray(outarray)))
    _lib.call("spp_subband", _lib.ptr(src), 32 if is_f32 else 8, 0, nchans, nsamps, _lib.ptr(d),
              _lib.ptr(cts), maxdelay, nsubs, _lib.ptr(dst), size, _lib.stream_handle())
    if host_out:
        outarray[:] = dst.cpu().numpy()


def good_size_real(n: int) -> int:
    """pocketfft good_size(n, real=True) (kernels.py:909-926): smallest 2^a 3^b 5^c >= n."""
    n = int(n)
    if n <= 6:
        return n
    best = 2 * n
    f5 = 1
    while f5 < best:
        f35 = f5
        while f35 < best:
            x = f35
            while x < n:
                x *= 2
            best = min(best, x)
            f35 *= 3
        f5 *= 5
    return best


def nb_fft_good_size(n: int, real: bool = False) -> int:  # noqa: FBT001, FBT002
    """kernels.nb_fft_good_size (core/kernels.py:909-926) for real transforms."""
    if not real:
        msg = "only real-transform good sizes are used on this path"
        raise NotImplementedError(msg)
    return good_size_real(n)


def normalize_template(arr: np.ndarray) -> np.ndarray:
    """kernels.normalize_template (core/kernels.py:1008-1027): zero mean, unit power."""
    mean = np.mean(arr)
    arr_norm = arr - mean
    norm = np.sqrt(np.sum(arr_norm**2))
    if norm == 0:
        return arr_norm
    return arr_norm / norm


def convolve_templates(data, temp_bank, ref_bin) -> np.ndarray:
    """kernels.convolve_templates (core/kernels.py:861-906) on the GPU.

    Circular convolution of ``data`` (padded circularly to the 5-smooth good size) with
    each zero-mean, unit-power, time-reversed template, evaluated directly in float64.
    Returns float32 ``[ntemps, len(data)]``.
    """
    _lib.require_gpu()
    n = _size(data)
    temps = [np.ascontiguousarray(t, dtype=np.float32).ravel() for t in temp_bank]
    lens = np.array([t.size for t in temps], dtype=np.int32)
    refs = np.array([int(r) for r in ref_bin], dtype=np.int32)
    taps = np.ascontiguousarray(np.concatenate(temps)) if temps else np.zeros(1, np.float32)
    src = _dev(data if not isinstance(data, np.ndarray) else np.asarray(data, dtype=np.float32), "float32")
    convs = _lib.empty((len(temps), n), "float32")
    need = np.zeros(1, dtype=np.int64)
    _lib.call("spp_convolve_templates_workspace_size", n, len(temps), int(taps.size),
              need.ctypes.data_as(_lib._P64))
    ws = _lib.empty((int(need[0]),), "uint8")
    _lib.call("spp_convolve_templates", _lib.ptr(src), n, _lib.ptr(taps), _lib.ptr(lens),
              _lib.ptr(refs), len(temps), _lib.ptr(convs), _lib.ptr(ws), int(need[0]),
              _lib.stream_handle())
    return convs.cpu().numpy()
