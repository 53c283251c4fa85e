"""GPU parity of the one-pass normalise + boxcar MF (spp_normalise_boxcar_mf).

It must give spp_boxcar_mf(spp_running_normalise(x)) -- the two-pass GPU form, itself
pinned against the float64 oracle in test_gpu_parity.py and test_gpu_mf_odd.py -- with the
same (peak, width) and snr within rel 2e-6 (only the order of the float64 sum S differs),
and the oracle's answer directly on sampled rows. Row lengths cover: shorter than one
2048-sample segment, segment and 32768-sample chunk edges, boxes that cross the row end
(the circular wrap), good sizes far above n (z[0, Ng - n) counted twice in S), odd good
sizes (the FFT path), the 12- and 32-width kernels, odd windows and unaligned rows.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu

W9 = [1, 2, 4, 8, 16, 32, 64, 128, 256]
W20 = [1, 2, 3, 4, 6, 8, 11, 16, 22, 32, 45, 64, 90, 128, 181, 256, 362, 512, 724, 1024]


def _two_pass(x, row_len, window, widths):
    from sigpyproc3_amd.sweep import boxcar_mf, running_normalise

    return boxcar_mf(running_normalise(x, row_len, window), row_len, widths)


def _fused(x, row_len, window, widths):
    from sigpyproc3_amd.sweep import normalise_boxcar_mf

    return normalise_boxcar_mf(x, row_len, window, widths)


def _compare(a, b, label):
    snr_a, peak_a, w_a = a
    snr_b, peak_b, w_b = b
    np.testing.assert_array_equal(w_a, w_b, err_msg=f"{label}: width index")
    np.testing.assert_array_equal(peak_a, peak_b, err_msg=f"{label}: peak")
    np.testing.assert_allclose(snr_a, snr_b, rtol=2e-6, atol=1e-6, err_msg=f"{label}: snr")


def _rows(rng, row_len, ld, *, integer=False, pulses=()):
    x = np.zeros((len(row_len), ld), np.float32)
    for i, n in enumerate(row_len):
        if integer:
            x[i, :n] = rng.integers(900, 1100, n)
        else:
            x[i, :n] = rng.normal(100.0, 7.0, n)
    for i, t, w, a in pulses:
        n = row_len[i]
        for k in range(w):
            x[i, (t + k) % n] += a
    return x


@pytest.mark.parametrize("window", [4096, 1001, 64])
def test_fused_equals_two_pass_float_rows(window):
    rng = np.random.default_rng(window)
    row_len = np.array([256, 300, 2047, 2048, 2049, 4095, 4097, 32767, 32768, 34815, 34816, 65636, 100003,
                        70001])
    pulses = [(2, 1000, 16, 40.0), (6, 4090, 8, 60.0),        # box across the row end
              (11, 30, 32, 25.0), (13, 70000 - 5, 20, 30.0),  # in the doubly counted head; wrap
              (9, 32760, 12, 50.0)]                           # across a chunk edge
    x = _rows(rng, row_len, 100008, pulses=pulses)
    print(f"float rows, window {window}", flush=True)
    _compare(_fused(x, row_len, window, W9), _two_pass(x, row_len, window, W9), f"window {window}")


def test_fused_equals_two_pass_integer_plane_rows():
    rng = np.random.default_rng(3)
    row_len = np.array([524288, 524287, 509454, 512000, 300000])
    x = _rows(rng, row_len, 524288, integer=True, pulses=[(1, 200000, 64, 80.0), (2, 509440, 30, 90.0)])
    print("integer rows at plane lengths", flush=True)
    _compare(_fused(x, row_len, 4096, W9), _two_pass(x, row_len, 4096, W9), "integer rows")


def test_fused_many_widths_and_unaligned_rows():
    rng = np.random.default_rng(4)
    row_len = np.array([1024, 5000, 70001, 36001, 2049])
    x = _rows(rng, row_len, 70003, pulses=[(2, 40000, 300, 9.0), (3, 35990, 700, 6.0)])  # ld 70003: unaligned
    print("20 widths up to 1024, unaligned rows", flush=True)
    _compare(_fused(x, row_len, 4096, W20), _two_pass(x, row_len, 4096, W20), "20 widths")
    xa = np.zeros((x.shape[0], 70004), np.float32)
    xa[:, : x.shape[1]] = x
    print("20 widths, window 4096, aligned rows of ld 70004", flush=True)
    _compare(_fused(xa, row_len, 4096, W20), _two_pass(xa, row_len, 4096, W20), "aligned")


def test_fused_rows_with_odd_good_sizes():
    rng = np.random.default_rng(5)
    row_len = np.array([19500, 20000, 3100, 32769, 4096])  # good sizes 19683, 20000, 3125, 32805, 4096
    x = _rows(rng, row_len, 32769, pulses=[(0, 7000, 8, 30.0), (3, 32760, 16, 30.0)])
    print("rows with odd good sizes", flush=True)
    _compare(_fused(x, row_len, 1024, W9), _two_pass(x, row_len, 1024, W9), "odd good sizes")


def test_fused_refuses_row_equal_to_odd_good_size():
    x = np.ones((2, 1000), np.float32)
    with pytest.raises(ValueError, match="odd FFT good size"):
        _fused(x, np.array([1000, 729]), 64, [1, 2])


def test_fused_vs_fp64_oracle():
    rng = np.random.default_rng(6)
    row_len = np.array([40000, 65636, 2049, 9000])
    x = _rows(rng, row_len, 65636, pulses=[(0, 12345, 16, 20.0), (1, 65630, 24, 15.0), (3, 8995, 40, 10.0)])
    snr, peak, widx = _fused(x, row_len, 4096, W9)
    for i, n in enumerate(row_len):
        z = O.running_normalise(x[i, :n], 4096)
        conv = O.boxcar_conv_closed(z, W9).astype(np.float32)
        it, pk, v = O.mf_argmax(conv)
        assert (int(widx[i]), int(peak[i])) == (it, pk), f"row {i}"
        assert abs(float(snr[i]) - float(v)) <= 1e-5 * abs(float(v)) + 1e-5 * np.sqrt(W9[it]), (i, snr[i], v)


def test_sweep_result_matched_filter_uses_one_pass():
    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.sweep import SweepPlan, SweepResult

    rng = np.random.default_rng(7)
    nchans, nsamps = 64, 40000
    data = rng.integers(0, 256, (nsamps, nchans), dtype=np.uint8)
    data[20000:20008] += 30
    table = np.zeros((3, nchans), np.int32)
    table[1] = np.arange(nchans) // 4
    table[2] = np.arange(nchans) // 2
    plan = SweepPlan(table, nsamps, 8, 0)
    plane = plan.execute(_lib.to_device(data.reshape(-1)))
    res = SweepResult(plane, plan.row_len, np.arange(3.0), header=None, plan=plan, table=table)
    _compare(res.matched_filter(4096, W9), _two_pass(plane, plan.row_len, 4096, W9), "sweep rows")
