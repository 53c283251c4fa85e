"""GPU parity of the matched filter where the FFT good size Ng is odd.

kernels.convolve_templates (core/kernels.py:893-905) calls irfft with its default
length, Ng - 1 samples for an odd Ng: the result is then not the Ng-circular convolution,
and a series of exactly Ng samples (Ng odd) makes the reference raise. The GPU path is
compared with the float64 oracle (``O.convolve_templates_fft``, pinned by explicit DFT
sums in tests/test_oracle.py) at rel 1e-5 + abs 1e-5 on every conv value, and with exact
(template, peak) on the argmax.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu

# good sizes: 121 -> 125, 3100 -> 3125, 19500 -> 19683, 455000 -> 455625 (odd);
# 3001 -> 3072 (even)
ODD_N = [121, 3100, 19500, 455000]


def _bank(kind, nmax):
    from sigpyproc3_amd.core.filters import MatchedFilter, Template

    if kind == "boxcar":
        widths = [w for w in MatchedFilter.get_box_width_spacing(256, 2.0) if w <= nmax]
    else:
        widths = np.geomspace(1, 16, 5)
    return [getattr(Template, f"gen_{kind}")(w) for w in widths]


@pytest.mark.parametrize("n", ODD_N + [3001])
@pytest.mark.parametrize("kind", ["boxcar", "gaussian"])
def test_convolve_templates_vs_oracle(n, kind):
    from sigpyproc3_amd.core import kernels

    print(f"convolve_templates n={n} kind={kind}", flush=True)
    rng = np.random.default_rng(n)
    z = rng.normal(size=n).astype(np.float32)
    z[n // 3 : n // 3 + 8] += 3.0
    bank = _bank(kind, n)
    convs = kernels.convolve_templates(z, [t.data for t in bank], [t.ref_bin for t in bank])
    ref = O.convolve_templates_fft(z, [t.data for t in bank], [t.ref_bin for t in bank])
    assert convs.shape == ref.shape == (len(bank), n)
    np.testing.assert_allclose(convs, ref, rtol=1e-5, atol=1e-5)
    if O.good_size_real(n) % 2:
        # the Ng-circular form differs from the reference here: the test can tell them apart
        circ = np.array([O.template_conv_circular(z, t.data, t.ref_bin)[:n] for t in bank])
        assert np.abs(circ - ref).max() > 1e-2


@pytest.mark.parametrize("n", [243, 625, 729])
def test_convolve_templates_refuses_length_equal_to_odd_good_size(n):
    from sigpyproc3_amd.core import kernels
    from sigpyproc3_amd.core.filters import MatchedFilter

    z = np.random.default_rng(n).normal(size=n).astype(np.float32)
    with pytest.raises(ValueError, match="odd FFT good size"):
        kernels.convolve_templates(z, [np.ones(4, np.float32)], [0])
    with pytest.raises(ValueError, match="odd FFT good size"):
        MatchedFilter(z, temp_kind="boxcar", nbins_max=8)


@pytest.mark.parametrize("n", [19500, 455000])
def test_matched_filter_odd_good_size(n):
    from sigpyproc3_amd.core.filters import MatchedFilter

    print(f"MatchedFilter n={n}", flush=True)
    rng = np.random.default_rng(7 + n)
    x = rng.normal(10.0, 2.0, size=n).astype(np.float32)
    x[n // 2 : n // 2 + 12] += 4.0
    mf = MatchedFilter(x, temp_kind="boxcar", nbins_max=64, spacing_factor=2.0)
    zs, _, _ = O.estimate_zscore(x)
    ref = O.convolve_templates_fft(zs, [t.data for t in mf.temp_bank], [t.ref_bin for t in mf.temp_bank])
    np.testing.assert_allclose(mf.convs, ref, rtol=1e-5, atol=1e-5)
    it, pk, v = O.mf_argmax(ref.astype(np.float32))
    best = next(k for k, t in enumerate(mf.temp_bank) if t is mf.best_temp)
    assert (best, mf.peak_bin) == (it, pk)
    assert abs(mf.snr - v) <= 1e-5 * abs(v) + 1e-5
    assert abs(mf.peak_bin - n // 2) < 16


def test_boxcar_mf_rows_with_odd_and_even_good_sizes():
    from sigpyproc3_amd.sweep import boxcar_mf

    rows = np.array([121, 3100, 19500, 20000, 4097, 455000])
    rng = np.random.default_rng(11)
    ld = int(rows.max())
    z = rng.normal(size=(rows.size, ld)).astype(np.float32)
    for i, n in enumerate(rows):
        z[i, n // 4 : n // 4 + 6] += 2.5
    widths = [1, 2, 4, 8, 16, 32, 64]
    snr, peak, widx = boxcar_mf(z, rows, widths)
    for i, n in enumerate(rows):
        conv = O.boxcar_conv_closed(z[i, :n], widths).astype(np.float32)
        it, pk, v = O.mf_argmax(conv)
        assert (widx[i], peak[i]) == (it, pk), f"row {i} (n={n})"
        assert abs(snr[i] - v) <= 1e-5 * abs(v) + 1e-5 * np.sqrt(widths[it])


def test_boxcar_mf_refuses_row_equal_to_odd_good_size():
    from sigpyproc3_amd.sweep import boxcar_mf

    z = np.zeros((2, 1000), np.float32)
    with pytest.raises(ValueError, match="odd FFT good size"):
        boxcar_mf(z, np.array([1000, 729]), [1, 2])
