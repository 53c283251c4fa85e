"""The allocation-free matched filter (spp_boxcar_mf_noalloc, scratch in a static device
array) against spp_boxcar_mf with a caller's workspace: identical outputs on two calls in a
row, the second on fewer and shorter rows, and a refused request beyond the scratch size."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(noalloc, z, row_len, widths):
    from sigpyproc3_amd import _lib

    zd = _lib.to_device(np.ascontiguousarray(z, dtype=np.float32))
    nrows, ldz = z.shape
    rl = np.ascontiguousarray(row_len, dtype=np.int64)
    w = np.ascontiguousarray(widths, dtype=np.int32)
    snr = _lib.empty((nrows,), "float32")
    peak = _lib.empty((nrows,), "int32")
    widx = _lib.empty((nrows,), "int32")
    if noalloc:
        _lib.call("spp_boxcar_mf_noalloc", _lib.ptr(zd), ldz, nrows, _lib.ptr(rl), _lib.ptr(w), int(w.size),
                  _lib.ptr(snr), _lib.ptr(peak), _lib.ptr(widx), _lib.stream_handle())
    else:
        need = np.zeros(1, dtype=np.int64)
        _lib.call("spp_boxcar_mf_workspace_size", nrows, _lib.ptr(rl), _lib.ptr(w), int(w.size),
                  need.ctypes.data_as(_lib._P64))
        ws = _lib.empty((int(need[0]),), "uint8")
        _lib.call("spp_boxcar_mf", _lib.ptr(zd), ldz, nrows, _lib.ptr(rl), _lib.ptr(w), int(w.size),
                  _lib.ptr(snr), _lib.ptr(peak), _lib.ptr(widx), _lib.ptr(ws), int(need[0]),
                  _lib.stream_handle())
    return snr.cpu().numpy(), peak.cpu().numpy(), widx.cpu().numpy()


def test_noalloc_matches_workspace_entry_twice():
    rng = np.random.default_rng(3)
    widths = [1, 2, 4, 8, 16, 32, 64, 128, 256]
    # first call: 96 rows of up to 200 000 samples (FFT good sizes even, and one odd: 370 -> 375)
    rows = np.array([200000 - 37 * i for i in range(95)] + [370], dtype=np.int64)
    z = rng.normal(size=(rows.size, int(rows.max()))).astype(np.float32)
    for i, n in enumerate(rows):
        z[i, n // 3 : n // 3 + 9] += 3.0
    print("no-alloc MF: 96 rows", flush=True)
    a = _call(False, z, rows, widths)
    b = _call(True, z, rows, widths)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    # second call: fewer, shorter rows (the scratch keeps the first call's contents)
    rows2 = np.array([50000, 40000, 30000, 4097, 3000], dtype=np.int64)
    z2 = rng.normal(size=(rows2.size, int(rows2.max()))).astype(np.float32)
    print("no-alloc MF: 5 rows", flush=True)
    a = _call(False, z2, rows2, widths[:6])
    b = _call(True, z2, rows2, widths[:6])
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_noalloc_refuses_beyond_scratch():
    z = np.zeros((4, 10), np.float32)
    rows = np.full(2048, 1 << 19, dtype=np.int64)   # 2048 rows x 256 chunks of partials > 8 MiB
    from sigpyproc3_amd import _lib

    w = np.array([1, 2], dtype=np.int32)
    rl = np.ascontiguousarray(rows)
    with pytest.raises(ValueError, match="scratch"):
        _lib.call("spp_boxcar_mf_noalloc", _lib.ptr(_lib.to_device(z)), 10, 2048, _lib.ptr(rl), _lib.ptr(w), 2,
                  None, None, None, _lib.stream_handle())
