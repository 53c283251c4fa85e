"""Pin the CPU oracle against the reference's known answers and against mathematics.

Each function of ``oracle.spp_oracle`` is checked here against something other than its
own formula: the reference's golden vectors (tests/test_kernels.py:12-41 unpack KAT, the
quickstart notebook's dedisperse KAT), library routines (np.unpackbits, scipy's
pocketfft good size, numpy FFTs), and brute force on small inputs.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.fft

from oracle import c_oracle
from oracle import spp_oracle as O

# tests/test_kernels.py:12-41 (input [7, 23])
UNPACK_KAT = {
    (1, "big"): [0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0, 1, 0, 1, 1, 1],
    (1, "little"): [1, 1, 1, 0, 0, 0, 0, 0, 1, 1, 1, 0, 1, 0, 0, 0],
    (2, "big"): [0, 0, 1, 3, 0, 1, 1, 3],
    (2, "little"): [3, 1, 0, 0, 3, 1, 1, 0],
    (4, "big"): [0, 7, 1, 7],
    (4, "little"): [7, 0, 7, 1],
}


@pytest.mark.parametrize(("nbits", "order"), list(UNPACK_KAT))
def test_unpack_kat(nbits, order):
    out = O.unpack(np.array([7, 23], dtype=np.uint8), nbits, order)
    np.testing.assert_array_equal(out, np.array(UNPACK_KAT[(nbits, order)], dtype=np.uint8))


@pytest.mark.parametrize("order", ["big", "little"])
def test_unpack_1bit_is_unpackbits(order, rng):
    a = rng.integers(0, 256, 4096, dtype=np.uint8)
    np.testing.assert_array_equal(O.unpack(a, 1, order), np.unpackbits(a, bitorder=order))


@pytest.mark.parametrize("nbits", [1, 2, 4])
@pytest.mark.parametrize("order", ["big", "little"])
def test_pack_unpack_roundtrip(nbits, order, rng):
    a = rng.integers(0, 256, 2048, dtype=np.uint8)
    np.testing.assert_array_equal(O.pack(O.unpack(a, nbits, order), nbits, order), a)


def test_c_unpack_matches(rng):
    a = rng.integers(0, 256, 10000, dtype=np.uint8)
    for nbits in (1, 2, 4):
        for order in ("big", "little"):
            np.testing.assert_array_equal(c_oracle.unpack(a, nbits, order), O.unpack(a, nbits, order))


def test_delays_invariants():
    # tests/test_params.py:9-37: zero delay at the reference channel, int32, shapes
    freqs = np.arange(1000, dtype=np.float32) * (-1.0) + 1400
    d = O.compute_dmdelays(freqs, 100, 0.001, float(freqs[0]))
    assert d.dtype == np.int32 and d.shape == (1000,) and d[0] == 0
    dv = O.compute_dmdelays(freqs, np.linspace(0, 100, 50), 0.001, float(freqs[0]))
    assert dv.shape == (50, 1000)
    np.testing.assert_equal(dv[:, 0], 0)
    assert (np.diff(dv, axis=1) >= 0).all()  # delays grow towards low frequencies


@pytest.mark.parametrize(("dm", "maxdelay"), [(0, 0), (100, 0.8), (5000, 40.4)])
def test_delays_seconds_vs_reference_test(dm, maxdelay):
    # tests/test_header.py:35-41 on parkes_4bit (832 ch, fch1 4030, foff -4, tsamp 512 us)
    f = O.chan_freqs(832, -4.0, 4030.0)
    d = O.compute_dmdelays(f, dm, 0.000512, 4030.0, in_samples=False)
    np.testing.assert_allclose(d.max(), maxdelay, atol=0.1)


@pytest.mark.parametrize("ref", ["fch1", "fmax", "fmin", "fcenter"])
def test_delays_within_one_sample_of_float64(ref):
    f = O.chan_freqs(4096, -0.0732421875, 1550.0)
    rf = O.ref_frequency(ref, 1550.0, -0.0732421875, 4096)
    dms = np.arange(0, 2048, 7)
    d32 = O.compute_dmdelays(f, dms, 64e-6, rf)
    d64 = O.compute_dmdelays_f64(f, dms, 64e-6, rf)
    assert np.abs(d32 - d64).max() <= 1
    assert (d32 == d64).mean() > 0.99


def test_notebook_kat(datadir):
    # docs/tutorials/quickstart.ipynb:296-417: FilReader("tutorial.fil").dedisperse(30)
    hdr, raw = O.read_sigproc(datadir / "tutorial.fil")
    s = O.unpacked_stream(hdr, raw)
    d = O.compute_dmdelays(O.chan_freqs(hdr["nchans"], hdr["foff"], hdr["fch1"]), 30, hdr["tsamp"],
                           hdr["fch1"])
    assert int(d.max()) == 17
    tim = O.filterbank_dedisperse(s, hdr["nsamples"], hdr["nchans"], d)
    assert tim.size == 187503
    # the notebook prints the float32 sum to 7 significant digits
    assert f"{float(tim.sum()):.6e}" == "1.963699e+07"
    assert float(tim.sum()) == 19636992.0  # exact integer sum
    assert (tim.max(), tim.min(), np.median(tim)) == (121.0, 88.0, 105.0)


def test_read_plan_blocks_tile_outputs():
    # Outputs of successive gulps tile [0, nsamps - skip) exactly (base.py:413-420)
    for nsamps, gulp, skip in [(65536, 16384, 7418), (4096, 512, 200), (187520, 16384, 17),
                               (1000, 1000, 0), (999, 400, 399)]:
        if skip >= min(nsamps, gulp):
            with pytest.raises(ValueError):
                O.read_plan_blocks(nsamps, gulp, skip)
            continue
        g = min(nsamps, gulp)
        covered = []
        pos = 0
        for ii, nr, sk in O.read_plan_blocks(nsamps, gulp, skip):
            assert nr <= g
            covered.append((ii * (g - skip), ii * (g - skip) + nr - skip))
            pos += nr + sk
        assert pos == nsamps
        if g < 2 * skip:
            continue  # dedisperse always reads gulp >= 2 * max_delay (base.py:402)
        flat = np.concatenate([np.arange(a, b) for a, b in covered])
        np.testing.assert_array_equal(flat, np.arange(nsamps - skip))


def test_read_plan_blocks_reference_example():
    # 8-bit 1024 ch, 65536 samples, DM 300: 6 full reads and one of 11740 samples
    blocks = O.read_plan_blocks(65536, 16384, 7418)
    assert [b[1] for b in blocks] == [16384] * 6 + [11740]


def _small_stream(rng, nsamps, nchans, nbits):
    return rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)


def test_dedisperse_gulp_matches_brute_force(rng):
    nsamps, nchans = 3000, 37
    s = _small_stream(rng, nsamps, nchans, 8)
    d = np.sort(rng.integers(0, 200, nchans)).astype(np.int32)[::-1].copy()
    tim = O.filterbank_dedisperse(s, nsamps, nchans, d, gulp=500)
    x = s.reshape(nsamps, nchans).astype(np.float64)
    md = d.max()
    brute = np.array([sum(x[t + d[c], c] for c in range(nchans)) for t in range(nsamps - md)])
    np.testing.assert_array_equal(tim, brute.astype(np.float32))


def test_dedisperse_negative_delays_wrap(rng):
    # numba wraps a negative row index by the gulp length (kernels.py:361)
    nsamps, nchans = 64, 5
    s = _small_stream(rng, nsamps, nchans, 8)
    d = np.array([0, -3, -7, 2, -1], dtype=np.int32)
    out = np.zeros(nsamps, dtype=np.float32)
    O.dedisperse_kernel(s, out, d, nchans, nsamps, 0)
    x = s.reshape(nsamps, nchans).astype(np.float64)
    for t in range(nsamps - 7):
        assert out[t] == sum(x[(t + d[c]) % nsamps, c] for c in range(nchans))


def test_c_dedisperse_matches_numpy(rng):
    nsamps, nchans = 5000, 64
    s = _small_stream(rng, nsamps, nchans, 8)
    d = rng.integers(-50, 300, nchans).astype(np.int32)
    a = np.zeros(nsamps, np.float32)
    b = np.zeros(nsamps, np.float32)
    O.dedisperse_kernel(s, a, d, nchans, nsamps, 0)
    c_oracle.dedisperse(s, b, d, nchans, nsamps, 0)
    np.testing.assert_array_equal(a, b)


def test_subband_reduces_to_dedisperse(rng):
    nsamps, nchans = 2000, 32
    s = _small_stream(rng, nsamps, nchans, 4)
    d = np.sort(rng.integers(0, 100, nchans)).astype(np.int32)[::-1].copy()
    sub = O.filterbank_subband(s, nsamps, nchans, d, 1, gulp=700)
    tim = O.filterbank_dedisperse(s, nsamps, nchans, d, gulp=700)
    np.testing.assert_array_equal(sub, tim)


def test_subband_brute_and_c(rng):
    nsamps, nchans, nsub = 900, 24, 6
    s = _small_stream(rng, nsamps, nchans, 8)
    d = rng.integers(0, 40, nchans).astype(np.int32)
    cts = (np.arange(nchans) // (nchans // nsub)).astype(np.int32)
    md = int(d.max())
    a = np.zeros((nsamps - md) * nsub, np.float32)
    O.subband_kernel(s, a, d, cts, md, nchans, nsub, nsamps)
    x = s.reshape(nsamps, nchans).astype(np.float64)
    brute = np.zeros((nsamps - md, nsub))
    for c in range(nchans):
        brute[:, cts[c]] += x[d[c] : d[c] + nsamps - md, c]
    np.testing.assert_array_equal(a, brute.ravel().astype(np.float32))
    b = np.zeros_like(a)
    c_oracle.subband(s, b, d, cts, md, nchans, nsub, nsamps)
    np.testing.assert_array_equal(a, b)


def test_sweep_rows_are_dedisperse(rng):
    nsamps, nchans = 4000, 48
    s = _small_stream(rng, nsamps, nchans, 2)
    f = O.chan_freqs(nchans, -1.0, 1500.0)
    table = O.compute_dmdelays(f, np.arange(0, 400, 37), 1e-3, 1500.0)
    rows = O.sweep_plane(s, nchans, table)
    plane = c_oracle.sweep(s, nchans, table)
    for i, r in enumerate(rows):
        np.testing.assert_array_equal(r, O.filterbank_dedisperse(s, nsamps, nchans, table[i]))
        np.testing.assert_array_equal(plane[i, : r.size], r)


def test_good_size_matches_pocketfft():
    ns = list(range(1, 5000)) + [509454, 1018893, 2**19 - 14834, 1000003]
    for n in ns:
        assert O.good_size_real(n) == scipy.fft.next_fast_len(n, real=True)


def test_box_widths():
    np.testing.assert_array_equal(O.box_width_spacing(256, 2.0), [1, 2, 4, 8, 16, 32, 64, 128, 256])
    w = O.box_width_spacing(16, 1.2)
    assert w[0] == 1 and (np.diff(w) > 0).all() and w[-1] <= 16


def test_irfft_default_brute_matches_numpy(rng):
    # numpy's default irfft length is 2 * (m - 1); the imaginary parts of the first and last
    # bins do not enter (they are nonzero here on purpose)
    for m in (2, 3, 8, 63, 1563):
        x = rng.normal(size=m) + 1j * rng.normal(size=m)
        np.testing.assert_allclose(O.irfft_default_brute(x), np.fft.irfft(x), atol=1e-12, rtol=0)


def _explicit_reference_convs(z, temps, ref_bins):
    """kernels.convolve_templates from its definition with explicit DFT sums (no FFT)."""
    z = np.asarray(z, dtype=np.float64)
    n = z.size
    ng = O.good_size_real(n)
    zt = z[np.arange(ng) % n]
    k = np.arange(ng // 2 + 1)
    ph = np.exp(-2j * np.pi * np.outer(k, np.arange(ng)) / ng)
    d = ph @ zt
    out = []
    for h, rb in zip(temps, ref_bins):
        tp = np.zeros(ng)
        tp[: len(h)] = h
        tp = np.roll(np.roll(tp, -rb)[::-1], 1)
        tp = tp - tp.mean()
        norm = np.sqrt((tp**2).sum())
        tn = tp / norm if norm != 0 else tp
        y = O.irfft_default_brute(d * (ph @ tn))
        if y.size < n:
            raise ValueError("short")
        out.append(y[:n])
    return np.array(out)


@pytest.mark.parametrize("n", [13, 50, 121, 1000, 3100])
def test_convolve_templates_fft_vs_explicit_dft(n, rng):
    # odd good sizes: 13 -> 15, 121 -> 125, 3100 -> 3125; even: 50, 1000
    z = rng.normal(size=n)
    x = np.arange(-4, 5)
    temps = [np.ones(1), np.ones(4), np.exp(-0.5 * x**2 / 4.0)]
    refs = [0, 0, 4]
    ref = _explicit_reference_convs(z, temps, refs)
    np.testing.assert_allclose(O.convolve_templates_fft(z, temps, refs), ref, atol=1e-10, rtol=0)
    for i, (h, rb) in enumerate(zip(temps, refs)):
        np.testing.assert_allclose(O.template_conv_closed(z, h, rb), ref[i], atol=1e-10, rtol=0)
    np.testing.assert_allclose(O.boxcar_conv_closed(z, [1, 4]), ref[:2], atol=1e-10, rtol=0)
    if O.good_size_real(n) % 2:
        # the reference output is not the Ng-circular convolution at odd good sizes
        circ = O.boxcar_conv_circular(z, [4])[0, :n]
        assert np.abs(circ - ref[1]).max() > 0.05


@pytest.mark.parametrize("n", [1, 3, 5, 243, 625, 675, 729])
def test_convolve_templates_raises_at_odd_good_size_lengths(n, rng):
    # n is its own (odd) good size: irfft returns n - 1 samples and kernels.py:905 raises
    assert O.good_size_real(n) == n and n % 2 == 1
    z = rng.normal(size=n)
    with pytest.raises(ValueError):
        O.convolve_templates_fft(z, [np.ones(1)], [0])
    with pytest.raises(ValueError):
        O.boxcar_conv_closed(z, [1])
    with pytest.raises(ValueError):
        _explicit_reference_convs(z, [np.ones(1)], [0])


@pytest.mark.parametrize("n", [19500, 455000])
def test_closed_form_mapping_matches_fft_restatement_large(n, rng):
    # odd good sizes 19683 and 455625: closed form + irfft-length mapping vs the FFT restatement
    z = rng.normal(size=n)
    widths = [1, 8, 256]
    fft = O.convolve_templates_fft(z, [np.ones(w) for w in widths], [0] * 3)
    np.testing.assert_allclose(O.boxcar_conv_closed(z, widths), fft, atol=1e-9, rtol=0)


@pytest.mark.parametrize("n", [1000, 4093, 30001, 121, 3100, 19500])
def test_boxcar_closed_form_matches_fft(n, rng):
    z = rng.normal(size=n)
    widths = [w for w in (1, 2, 4, 8, 16, 256) if w <= n]  # a template fits the padded series
    fft = O.convolve_templates_fft(z, [np.ones(w) for w in widths], [0] * len(widths))
    closed = O.boxcar_conv_closed(z, widths)
    np.testing.assert_allclose(closed, fft, atol=1e-9, rtol=0)


def test_template_closed_form_matches_fft(rng):
    n = 2000
    z = rng.normal(size=n)
    x = np.arange(-9, 10)
    h = np.exp(-0.5 * x**2 / 9.0)
    fft = O.convolve_templates_fft(z, [h], [9])[0]
    closed = O.template_conv_closed(z, h, 9)
    np.testing.assert_allclose(closed, fft, atol=1e-9, rtol=0)


def test_boxcar_closed_form_brute_force(rng):
    n = 50
    z = rng.normal(size=n)
    ng = O.good_size_real(n)
    zt = z[np.arange(ng) % n]
    for w in (1, 3, 7):
        out = O.boxcar_conv_closed(z, [w])[0]
        for t in range(n):
            s = sum(zt[(t + j) % ng] for j in range(w))
            ref = (s - w * zt.sum() / ng) / np.sqrt(w * (1 - w / ng))
            assert abs(out[t] - ref) < 1e-12


def test_fft_float32_close_to_closed_form(rng):
    # the reference computes convs with float32 FFTs; the closed form is the exact value
    n = 509454 // 16
    z = rng.normal(size=n).astype(np.float32)
    widths = [1, 16, 256]
    f32 = O.convolve_templates_fft(z, [np.ones(w, np.float32) for w in widths], [0] * 3, dtype=np.float32)
    closed = O.boxcar_conv_closed(z, widths)
    assert np.abs(f32 - closed).max() < 2e-4


def test_zscore_reference_definitions(rng):
    x = rng.normal(5, 2, 1001).astype(np.float32)
    z, loc, scale = O.estimate_zscore(x, "median", "iqr")
    q1, q3 = np.percentile(x.astype(np.float64), [25, 75])
    assert loc[0] == np.median(x)
    np.testing.assert_allclose(scale[0], (q3 - q1) / 1.3489795003921634)
    np.testing.assert_allclose(z, (x - loc[0]) / scale[0], rtol=1e-6)


def test_symmetric_pad_matches_numpy(rng):
    x = rng.normal(size=7)
    for w in (1, 2, 3, 4, 6, 9, 20):
        p = O.symmetric_pad_window(x, w)
        L = w // 2
        n = x.size
        idx = np.arange(p.size) - L
        m = np.mod(idx, 2 * n)
        sym = np.where(m < n, m, 2 * n - 1 - m)
        np.testing.assert_array_equal(p, x[sym])


def test_running_filter_brute_force(rng):
    x = rng.normal(size=300)
    for w in (1, 4, 5, 64):
        got = O.running_filter_mean(x, w)
        p = O.symmetric_pad_window(x, w)
        ref = np.array([p[t : t + w].mean() for t in range(x.size)])
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_running_normalise_brute_force(rng):
    x = rng.integers(0, 1000, 500).astype(np.float32)
    w = 64
    z = O.running_normalise(x, w)
    p = O.symmetric_pad_window(x, w)
    for t in range(x.size):
        win = p[t : t + w]
        sd = win.std()
        ref = (x[t] - win.mean()) / (sd if sd > 0 else 1.0)
        assert abs(float(z[t]) - ref) <= 1e-5 * max(1.0, abs(ref))


# ---------------------------------------------------------------------------
# gulp pre-steps (kernels.py:255-345, 465-500)
# ---------------------------------------------------------------------------


def test_pack_spills_out_of_range_values():
    # kernels.pack2_8_big / pack4_8_little / pack1_8_big OR shifted values unmasked
    # (kernels.py:72-173); bits above 7 are dropped by the uint8 store.
    a = np.array([0, 5, 0, 0, 7, 0, 0, 1], dtype=np.uint8)
    assert O.pack(a, 2, "big").tolist() == [(5 << 4) & 0xFF, ((7 << 6) | 1) & 0xFF]
    b = np.array([17, 2], dtype=np.uint8)
    assert O.pack(b, 4, "little").tolist() == [(17 | (2 << 4)) & 0xFF]
    c = np.array([2, 0, 0, 0, 0, 0, 0, 3], dtype=np.uint8)
    assert O.pack(c, 1, "big").tolist() == [((2 << 7) | 3) & 0xFF]


def test_collapse_offsets_and_sums(rng):
    x = rng.integers(0, 256, (10, 3), dtype=np.uint8)
    # gulp 4 over 10 samples: gulps (0, 4), (1, 4), (2, 2); the last lands at 2 * 2 = 4
    expect = x.sum(axis=1).astype(np.float32)
    got = O.collapse(x.reshape(-1), 10, 3, gulp=4)
    exp = np.zeros(10, dtype=np.float32)
    exp[0:4], exp[4:8] = expect[0:4], expect[4:8]
    exp[4:6] = expect[8:10]
    np.testing.assert_array_equal(got, exp)
    # a dividing gulp gives the plain channel sum; mean divides by nchans in float32
    np.testing.assert_array_equal(O.collapse(x.reshape(-1), 10, 3, gulp=5), expect)
    np.testing.assert_array_equal(O.collapse(x.reshape(-1), 10, 3, gulp=5, method="mean"),
                                  expect / np.float32(3))


def test_bandpass_sequential_float32_and_exact_integers(rng):
    x = rng.normal(1000.0, 300.0, (257, 5)).astype(np.float32)
    got = O.bandpass(x.reshape(-1), 257, 5, gulp=100)
    for c in range(5):                       # brute force: scalar float32 adds in time order
        acc = np.float32(0.0)
        for t in range(257):
            acc = np.float32(acc + x[t, c])
        assert got[c] == acc
    u = rng.integers(0, 256, (3000, 7), dtype=np.uint8)
    exact = u.astype(np.int64).sum(axis=0)   # < 2^24: every order gives the integer
    np.testing.assert_array_equal(O.bandpass(u.reshape(-1), 3000, 7, gulp=700), exact.astype(np.float32))
    np.testing.assert_array_equal(O.bandpass(u.reshape(-1), 3000, 7, gulp=700, parallel=True),
                                  exact.astype(np.float32))
    np.testing.assert_array_equal(O.bandpass(u.reshape(-1), 3000, 7, gulp=700, method="mean"),
                                  exact.astype(np.float32) / np.float32(3000))


def test_bandpass_parallel_rounds_once_per_gulp():
    """extract_bpass_par adds a float64 gulp sum to the float32 bandpass in float64 and rounds
    once. One channel, gulps of 65 800 samples: gulp 1 sums to 1, gulp 2 to 2^24 + 1. Rounded
    once, 1 + 2^24 + 1 = 2^24 + 2 is representable; rounding the gulp sum to float32 first
    (2^24, a tie to even) and then adding 1 (another tie) would give 2^24."""
    g = 65800
    x = np.zeros(2 * g, dtype=np.uint8)
    x[0] = 1
    x[g : g + 65793] = 255                    # 65793 * 255 = 2^24 - 1
    x[g + 65793] = 2
    assert int(x[g:].astype(np.int64).sum()) == 2**24 + 1
    got = O.bandpass(x, 2 * g, 1, gulp=g, parallel=True)
    assert got[0] == np.float32(2**24 + 2)
    # the sequential float32 order: 1, then 2^24 - 1 more reaches 2^24 exactly, and + 2 -> 2^24 + 2
    seq = O.bandpass(x, 2 * g, 1, gulp=g)
    acc = np.float32(0)
    for v in x:                              # brute force, scalar float32 adds
        acc = np.float32(acc + np.float32(v))
    assert seq[0] == acc


def test_fma_exact_single_rounding():
    a = 1.0 + 2.0**-30
    # a*a = 1 + 2^-29 + 2^-60: two roundings lose the 2^-60 term, one rounding keeps it
    assert O.fma_exact(a, a, -1.0) == 2.0**-29 + 2.0**-60
    assert a * a - 1.0 == 2.0**-29


def test_remove_zerodm_identities(rng):
    nchans = 16
    x = rng.integers(0, 256, (40, nchans), dtype=np.uint8)
    # flat bandpass: w = 1/16 exactly, so fused and unfused agree and y = x - mean + b exactly
    b = np.full(nchans, 100.0, dtype=np.float32)
    y32 = O.remove_zerodm(x.reshape(-1), nchans, b, np.float32).reshape(-1, nchans)
    expect = x - x.astype(np.float64).mean(axis=1, keepdims=True) + 100.0
    np.testing.assert_array_equal(y32, expect.astype(np.float32))
    np.testing.assert_array_equal(y32, O.remove_zerodm(x.reshape(-1), nchans, b, np.float32,
                                                       fused=False).reshape(-1, nchans))
    # uint8 output: low 8 bits of the truncation toward zero
    u8 = O.remove_zerodm(x.reshape(-1), nchans, b, np.uint8).reshape(-1, nchans)
    np.testing.assert_array_equal(u8, (np.trunc(expect).astype(np.int64) & 0xFF).astype(np.uint8))
    # weighted bandpass: the weighted zero-DM is removed, sum_c (y - b) = 0 up to rounding
    b2 = rng.uniform(50, 150, nchans).astype(np.float32)
    y = O.remove_zerodm(x.reshape(-1), nchans, b2, np.float32).reshape(-1, nchans).astype(np.float64)
    np.testing.assert_allclose((y - b2).sum(axis=1), 0.0, atol=1e-3)


# ---------------------------------------------------------------------------
# in-memory blocks (kernels.py:1059-1193)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("shifts", [np.arange(5), -np.arange(5), np.array([2, -1, 0, 3, -2]),
                                    np.array([6, -7, 8, -9, 10])])
def test_roll_block_rows_and_inverse(shifts):
    # tests/test_kernels.py:472-491
    arr = np.arange(25).reshape(5, 5).astype(np.float32)
    rolled = O.roll_block(arr, shifts)
    for i, s in enumerate(shifts):
        np.testing.assert_array_equal(rolled[i], np.roll(arr[i], s))
    np.testing.assert_array_equal(O.roll_block(rolled, -shifts), arr)


@pytest.mark.parametrize("shifts", [np.arange(5), -np.arange(5), np.array([2, -1, 0, 3, -2])])
def test_roll_block_valid_is_unwrapped_slice(shifts):
    # tests/test_kernels.py:505-524, 526-535
    arr = np.arange(50).reshape(5, 10).astype(np.float32)
    lo, hi = max(0, shifts.max()), arr.shape[1] + min(0, shifts.min())
    np.testing.assert_array_equal(O.roll_block_valid(arr, shifts), O.roll_block(arr, shifts)[:, lo:hi])
    assert O.roll_block_valid(arr, np.array([-9, -5, -4, -2, -1])).shape == (5, 1)
    with pytest.raises(ValueError):
        O.roll_block_valid(np.zeros((10, 19), np.float32), np.arange(10) * 5)


def test_dmt_block_kats_and_order(rng):
    # tests/test_kernels.py:545-575: an impulse column sums to one in every output sample
    arr = np.zeros((10, 10), dtype=np.float32)
    arr[:, 5] = 1.0
    np.testing.assert_array_equal(O.dmt_block(arr, np.arange(10)[None]), np.ones((1, 10), np.float32))
    arr = np.zeros((10, 19), dtype=np.float32)
    arr[:, 9] = 1.0
    np.testing.assert_array_equal(O.dmt_block(arr, np.arange(10)[None], valid=True),
                                  np.ones((1, 10), np.float32))
    with pytest.raises(ValueError):                       # rows of different valid widths
        O.dmt_block(arr, np.stack([np.arange(10), np.zeros(10, int)]), valid=True)
    # float32 channel order from 0: brute force with scalar adds
    x = rng.normal(0, 1e4, (7, 30)).astype(np.float32)
    d = rng.integers(-40, 40, (3, 7))
    got = O.dmt_block(x, d)
    for i in range(3):
        for t in range(30):
            acc = np.float32(0.0)
            for c in range(7):
                acc = np.float32(acc + x[c, (t - d[i, c]) % 30])
            assert got[i, t] == acc


# ---------------------------------------------------------------------------
# fold (kernels.py:404-445)
# ---------------------------------------------------------------------------


def test_fold_conserves_sums_and_finds_the_pulse(rng):
    nchans, n, tsamp, period = 8, 6000, 1e-3, 0.025
    x = rng.integers(0, 10, (n, nchans)).astype(np.uint8)
    x[::25, :] += 100                                   # one pulse per period (25 samples)
    delays = np.zeros(nchans, dtype=np.int32)
    fold_ar = np.zeros(20 * 4 * 2, dtype=np.float32)
    count_ar = np.zeros(20 * 4 * 2, dtype=np.int32)
    O.fold_kernel(x.reshape(-1), fold_ar, count_ar, delays, 0, tsamp, period, 0.0, n, n, nchans, 20, 4, 2, 0)
    assert count_ar.sum() == n * nchans                  # every sample of every channel once
    assert fold_ar.sum() == float(x.astype(np.int64).sum())   # integer sums < 2^24: exact
    cube = (fold_ar / count_ar).reshape(4, 2, 20)
    prof = cube.sum(axis=(0, 1))
    # pulses at t = 25k land in phase bin int(20 * k + 0.5) % 20 = 0
    assert int(np.argmax(prof)) == 0 and prof[0] > 5 * np.median(prof)


def test_filterbank_fold_matches_gulp_free_fold(rng):
    """Gulps with skipback give the same cube as one pass (the schedule tiles time)."""
    nchans, n = 16, 9000
    x = rng.integers(0, 4, (n, nchans)).astype(np.uint8)
    d = np.sort(rng.integers(0, 300, nchans))[::-1].astype(np.int32)
    a = O.filterbank_fold(x.reshape(-1), n, nchans, 64e-6, d, 0.0123, nbins=16, nints=8, nbands=4, gulp=1000)
    b = O.filterbank_fold(x.reshape(-1), n, nchans, 64e-6, d, 0.0123, nbins=16, nints=8, nbands=4, gulp=n)
    np.testing.assert_array_equal(a, b)
