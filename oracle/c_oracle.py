"""ctypes loader for the C/OpenMP oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/_build/libspp_oracle.so`` (built by ``oracle/Makefile``). Used by
tests as a fast checker and by ``bench.py`` as the CPU baseline; never by the product.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libspp_oracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        p = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.spp_oracle_dedisperse.argtypes = [p, p, p, i64, i64, i64]
        L.spp_oracle_dedisperse_f32.argtypes = [p, p, p, i64, i64, i64]
        L.spp_oracle_sweep.argtypes = [p, i64, i64, p, i64, p, i64]
        L.spp_oracle_subband.argtypes = [p, p, i64, p, p, i64, i64, i64, i64]
        L.spp_oracle_unpack.argtypes = [p, i64, p, ctypes.c_int, ctypes.c_int]
        L.spp_oracle_set_threads.argtypes = [ctypes.c_int]
        L.spp_oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def set_threads(n: int) -> None:
    lib().spp_oracle_set_threads(int(n))


def max_threads() -> int:
    return int(lib().spp_oracle_max_threads())


def dedisperse(in_arr: np.ndarray, out_arr: np.ndarray, delays: np.ndarray, nchans: int,
               nsamps: int, out_start: int) -> None:
    """kernels.dedisperse_par restated in C (core/kernels.py:340-362, 542)."""
    d = np.ascontiguousarray(delays, dtype=np.int32)
    assert out_arr.dtype == np.float32 and out_arr.flags.c_contiguous
    if in_arr.dtype == np.uint8:
        a = np.ascontiguousarray(in_arr)
        lib().spp_oracle_dedisperse(_ptr(a), _ptr(out_arr), _ptr(d), nchans, nsamps, out_start)
    else:
        a = np.ascontiguousarray(in_arr, dtype=np.float32)
        lib().spp_oracle_dedisperse_f32(_ptr(a), _ptr(out_arr), _ptr(d), nchans, nsamps, out_start)


def sweep(stream: np.ndarray, nchans: int, table: np.ndarray, ld: int | None = None) -> np.ndarray:
    """Per-DM dedisperse of a whole unpacked stream; returns plane [ndm][ld]."""
    s = np.ascontiguousarray(stream, dtype=np.uint8).reshape(-1)
    t = np.ascontiguousarray(np.atleast_2d(table), dtype=np.int32)
    n = s.size // nchans
    ld = n if ld is None else ld
    plane = np.zeros((t.shape[0], ld), dtype=np.float32)
    lib().spp_oracle_sweep(_ptr(s), n, nchans, _ptr(t), t.shape[0], _ptr(plane), ld)
    return plane


def subband(in_arr, out_arr, delays, chan_to_sub, maxdelay, nchans, nsubs, nsamps) -> None:
    a = np.ascontiguousarray(in_arr, dtype=np.uint8)
    d = np.ascontiguousarray(delays, dtype=np.int32)
    cts = np.ascontiguousarray(chan_to_sub, dtype=np.int32)
    lib().spp_oracle_subband(_ptr(a), _ptr(out_arr), out_arr.size, _ptr(d), _ptr(cts),
                             maxdelay, nchans, nsubs, nsamps)


def unpack(arr: np.ndarray, nbits: int, bitorder: str) -> np.ndarray:
    a = np.ascontiguousarray(arr, dtype=np.uint8)
    out = np.empty(a.size * (8 // nbits), dtype=np.uint8)
    lib().spp_oracle_unpack(_ptr(a), a.size, _ptr(out), nbits, 1 if bitorder[0] == "b" else 0)
    return out
