"""GPU parity: every HIP entry point against the CPU oracle, bit-exact for integer work.

All calls go through the C ABI (``libspp_hip.so``) via the package's drop-in API.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_oracle
from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu

UNPACK_KAT = {
    (1, "big"): [0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0, 1, 0, 1, 1, 1],
    (1, "little"): [1, 1, 1, 0, 0, 0, 0, 0, 1, 1, 1, 0, 1, 0, 0, 0],
    (2, "big"): [0, 0, 1, 3, 0, 1, 1, 3],
    (2, "little"): [3, 1, 0, 0, 3, 1, 1, 0],
    (4, "big"): [0, 7, 1, 7],
    (4, "little"): [7, 0, 7, 1],
}


def _torch():
    import torch

    return torch


@pytest.mark.parametrize(("nbits", "order"), list(UNPACK_KAT))
def test_unpack_kat(nbits, order):
    from sigpyproc3_amd.core import kernels

    expected = np.array(UNPACK_KAT[(nbits, order)], dtype=np.uint8)
    out = np.empty_like(expected)
    getattr(kernels, f"unpack{nbits}_8_{order}")(np.array([7, 23], dtype=np.uint8), out)
    np.testing.assert_array_equal(out, expected)


@pytest.mark.parametrize("nbits", [1, 2, 4])
@pytest.mark.parametrize("order", ["big", "little"])
@pytest.mark.parametrize(("size", "offset"), [(1 << 20, 0), (1001, 0), (4099, 3), (7, 1)])
def test_unpack_random_vs_oracle(nbits, order, size, offset, rng):
    from sigpyproc3_amd.io import bits

    torch = _torch()
    a = rng.integers(0, 256, size + offset, dtype=np.uint8)
    src = torch.from_numpy(a).cuda()[offset:]
    out = torch.empty(size * (8 // nbits) + 1, dtype=torch.uint8, device="cuda")[1:]
    bits.unpack(src, nbits, out, bitorder=order)
    np.testing.assert_array_equal(out.cpu().numpy(), O.unpack(a[offset:], nbits, order))
    host = bits.unpack(a[offset:], nbits, bitorder=order)
    np.testing.assert_array_equal(host, O.unpack(a[offset:], nbits, order))


def test_dedisperse_operator_u8_and_negative_delays(rng):
    from sigpyproc3_amd.core import kernels

    nsamps, nchans = 6000, 100
    s = rng.integers(0, 256, nsamps * nchans, dtype=np.uint8)
    for d in (np.sort(rng.integers(0, 700, nchans))[::-1].astype(np.int32),
              rng.integers(-300, 300, nchans).astype(np.int32)):
        a = np.full(nsamps + 50, -1, np.float32)
        b = a.copy()
        kernels.dedisperse(s, a, d, nchans, nsamps, 50)
        c_oracle.dedisperse(s, b, d, nchans, nsamps, 50)
        np.testing.assert_array_equal(a, b)


def test_dedisperse_operator_f32(rng):
    from sigpyproc3_amd.core import kernels

    nsamps, nchans = 3000, 40
    s = rng.normal(size=nsamps * nchans).astype(np.float32)
    d = rng.integers(-100, 100, nchans).astype(np.int32)
    a = np.zeros(nsamps, np.float32)
    b = np.zeros(nsamps, np.float32)
    kernels.dedisperse(s, a, d, nchans, nsamps, 0)
    c_oracle.dedisperse(s, b, d, nchans, nsamps, 0)
    # same float64 sum in channel order, rounded once to float32
    np.testing.assert_array_equal(a, b)


def test_notebook_kat_on_gpu(filfiles):
    # docs/tutorials/quickstart.ipynb:296-417
    from sigpyproc3_amd import FilReader

    tim = FilReader(filfiles["tutorial"]).dedisperse(30)
    assert tim.data.dtype == np.float32 and tim.header.nchans == 1 and tim.header.dm == 30
    assert tim.data.size == 187503
    assert float(tim.data.sum()) == 19636992.0
    assert f"{float(tim.data.sum()):.6e}" == "1.963699e+07"
    assert (tim.data.max(), tim.data.min(), np.median(tim.data)) == (121.0, 88.0, 105.0)


@pytest.mark.parametrize("key", ["1bit", "2bit", "4bit", "8bit", "tutorial"])
@pytest.mark.parametrize("ref", ["fch1", "fmin", "fcenter"])
def test_filterbank_dedisperse_vs_oracle(filfiles, key, ref):
    from sigpyproc3_amd import FilReader

    files = filfiles[key] if isinstance(filfiles[key], list) else [filfiles[key]]
    pieces = [O.read_sigproc(f) for f in files]
    hdr = pieces[0][0]
    stream = np.concatenate([O.unpacked_stream(h, r) for h, r in pieces])
    nsamples = sum(h["nsamples"] for h, _ in pieces)
    fil = FilReader(filfiles[key])
    rf = O.ref_frequency(ref, hdr["fch1"], hdr["foff"], hdr["nchans"])
    for dm in (0, 37.5, 250):
        d = O.compute_dmdelays(O.chan_freqs(hdr["nchans"], hdr["foff"], hdr["fch1"]), dm, hdr["tsamp"], rf)
        if 2 * int(np.abs(d).max()) >= nsamples:
            continue
        for gulp in (16384, 1000):
            tim = fil.dedisperse(dm, ref, gulp=gulp)
            ref_tim = O.filterbank_dedisperse(stream, nsamples, hdr["nchans"], d, gulp=gulp)
            np.testing.assert_array_equal(tim.data, ref_tim)


def test_filterbank_dedisperse_start_nsamps(filfiles):
    from sigpyproc3_amd import FilReader

    h, raw = O.read_sigproc(filfiles["4bit"])
    s = O.unpacked_stream(h, raw)
    fil = FilReader(filfiles["4bit"])
    d = fil.header.get_dmdelays(60)
    tim = fil.dedisperse(60, gulp=700, start=100)
    ref = O.filterbank_dedisperse(s, h["nsamples"], h["nchans"], d, gulp=700, start=100)
    np.testing.assert_array_equal(tim.data, ref)
    # A partial last gulp that does not end at the end of the stream reads a full buffer
    # and fails the byte-count check, as the reference's read_plan does
    # (io/fileio.py:271-282 fills the whole buffer; readers.py:224-230 raises).
    with pytest.raises(ValueError, match="Unexpected number of bytes"):
        fil.dedisperse(60, gulp=700, start=100, nsamps=3000)


def test_filterbank_subband_vs_oracle(filfiles, tmp_path):
    # tests/test_base.py:163-168 (header) + values against the oracle
    from sigpyproc3_amd import FilReader

    fil = FilReader(filfiles["4bit"])
    h, raw = O.read_sigproc(filfiles["4bit"])
    s = O.unpacked_stream(h, raw)
    for dm, nsub in ((0, 52), (120, 16), (120, 1)):
        out = fil.subband(dm=dm, nsub=nsub, outfile_name=str(tmp_path / f"s{dm}_{nsub}.fil"), gulp=900)
        new = FilReader(out)
        assert new.header.nchans == nsub and new.header.nbits == 32
        d = fil.header.get_dmdelays(dm)
        ref = O.filterbank_subband(s, h["nsamples"], h["nchans"], d, nsub, gulp=900)
        hdr2, raw2 = O.read_sigproc(out)
        got = np.frombuffer(raw2.tobytes(), dtype=np.float32)
        np.testing.assert_array_equal(got, ref)
        if dm == 0:
            assert new.header.nsamples == fil.header.nsamples


def test_subband_operator_vs_c(rng):
    from sigpyproc3_amd.core import kernels

    nsamps, nchans, nsub = 5000, 96, 12
    s = rng.integers(0, 16, nsamps * nchans, dtype=np.uint8)
    d = rng.integers(-40, 400, nchans).astype(np.int32)
    md = int(np.abs(d).max())
    cts = (np.arange(nchans) // (nchans // nsub)).astype(np.int32)
    a = np.full((nsamps - md) * nsub, 7.0, np.float32)
    b = np.zeros_like(a)
    kernels.subband(s, a, d, cts, md, nchans, nsub, nsamps)
    c_oracle.subband(s, b, d, cts, md, nchans, nsub, nsamps)
    np.testing.assert_array_equal(a, b)


def _sweep_case(rng, nsamps, nchans, nbits, dms, foff, tsamp, fch1=1500.0, variants=(None,)):
    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.sweep import SweepPlan

    order = O.DEFAULT_BITORDER.get(nbits, "big")
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    packed = samples if nbits == 8 else O.pack(samples, nbits, order)
    f = O.chan_freqs(nchans, foff, fch1)
    table = np.atleast_2d(O.compute_dmdelays(f, dms, tsamp, fch1)).astype(np.int32)
    plan = SweepPlan(table, nsamps, nbits, 1 if order == "big" else 0)
    dev = _lib.to_device(packed)
    ref = c_oracle.sweep(samples, nchans, table, ld=plan.max_len)
    if variants == "all":
        variants = range(SweepPlan.variant_count())
    for v in variants:
        if v is not None:
            plan.variant = v
        plane = plan.execute(dev).cpu().numpy()
        for i in range(table.shape[0]):
            n = int(plan.row_len[i])
            assert n == nsamps - table[i].max()
            np.testing.assert_array_equal(plane[i, :n], ref[i, :n], err_msg=f"row {i} variant {v}")
            assert not plane[i, n:].any()
    return plan


@pytest.mark.parametrize("nbits", [1, 2, 4, 8])
@pytest.mark.parametrize("nchans", [64, 336, 1000])
def test_sweep_vs_oracle(nbits, nchans, rng):
    _sweep_case(rng, 20000 + 333, nchans, nbits, np.arange(0, 61, 1.5), -300.0 / nchans, 256e-6,
                variants="all")


@pytest.mark.parametrize("seed", range(6))
def test_sweep_random_configs(seed):
    """Seeded random sweeps: field width, a ragged channel count, stream length, an irregular
    ascending DM list and the band, every block shape against the C oracle."""
    r = np.random.default_rng(1000 + seed)
    nbits = int(r.choice([1, 2, 4, 8]))
    step = 8 // nbits if nbits < 8 else 1
    nchans = int(r.integers(16, 700)) // step * step + step
    nsamps = int(r.integers(6000, 30000))
    dms = np.sort(r.uniform(0, 80, int(r.integers(3, 40))))
    bw = float(r.uniform(50, 400))
    print(f"seed {seed}: {nbits}-bit {nchans} ch {nsamps} samples, {dms.size} DMs, {bw:.0f} MHz", flush=True)
    _sweep_case(r, nsamps, nchans, nbits, dms, -bw / nchans, 128e-6, fch1=1500.0, variants="all")


def test_sweep_wide_spread_tiles(rng):
    # DM steps so large that tiles hold fewer than 16 rows; ragged last tile
    plan = _sweep_case(rng, 30000, 128, 2, np.arange(0, 2000, 90.0), -2.0, 64e-6, variants="all")
    assert plan.info.ndm_tiles > 2


def test_sweep_wide_accumulators(rng):
    # 4-bit with > 4368 channels and 8-bit with > 257 channels use 32-bit accumulators
    _sweep_case(rng, 4096 + 7, 4400, 4, np.array([0.0, 3.0, 7.0]), -0.05, 64e-6, variants="all")
    _sweep_case(rng, 5000, 1024, 8, np.arange(0, 20.0), -0.3, 64e-6, variants="all")
    # 2-bit above 21845 channels and 1-bit above 65535 channels overflow 16-bit lanes
    _sweep_case(rng, 3000, 22000, 2, np.array([0.0, 1.0, 2.0]), -0.01, 64e-6, variants="all")
    _sweep_case(rng, 2100, 65600, 1, np.array([0.0, 1.0]), -0.004, 64e-6, variants=(1, 3, 4, 5))


@pytest.mark.parametrize("nbits", [2, 8])
def test_sweep_execute_n_shorter_gulps(nbits, rng):
    """One plan (capacity nsamps) run on shorter prefixes, larger and smaller in turn, with
    the same plane buffer: every execution equals the oracle on that prefix."""
    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.sweep import SweepPlan

    nsamps, nchans = 24000, 256
    order = O.DEFAULT_BITORDER.get(nbits, "big")
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    packed = samples if nbits == 8 else O.pack(samples, nbits, order)
    f = O.chan_freqs(nchans, -1.0, 1500.0)
    table = np.atleast_2d(O.compute_dmdelays(f, np.arange(0, 40, 0.5), 64e-6, 1500.0)).astype(np.int32)
    plan = SweepPlan(table, nsamps, nbits, 1 if order == "big" else 0)
    dev = _lib.to_device(packed)
    plane = _lib.empty((plan.ndm * plan.max_len,), "float32")
    md = int(table.max())
    for n in (nsamps, md + 1, 9999, nsamps - 1, md + 3000, 2 * md + 17):
        ld = int(plan.row_len_for(n).max())
        plan.execute_n(dev, n, plane, ld=ld)
        got = plane[: plan.ndm * ld].reshape(plan.ndm, ld).cpu().numpy()
        ref = c_oracle.sweep(samples[: n * nchans], nchans, table, ld=ld)
        rl = plan.row_len_for(n)
        for i in range(plan.ndm):
            np.testing.assert_array_equal(got[i, : rl[i]], ref[i, : rl[i]], err_msg=f"n {n} row {i}")
            assert not got[i, rl[i]:].any()
    with pytest.raises(ValueError, match="exceed"):
        plan.execute_n(dev, md, plane)
    bigger = _lib.zeros(((nsamps + 64) * nchans * nbits // 8,), "uint8")
    with pytest.raises(ValueError, match="exceeds"):
        plan.execute_n(bigger, nsamps + 64, plane, ld=nsamps)


def test_sweep_rows_equal_filterbank_dedisperse(filfiles):
    from sigpyproc3_amd import FilReader

    fil = FilReader(filfiles["tutorial"])
    dms = np.arange(0, 61, 1.0)
    res = fil.dedisperse_sweep(dms)
    for i in (0, 17, 30, 60):
        ts = res.timeseries(i)
        ref = fil.dedisperse(dms[i])
        np.testing.assert_array_equal(ts.data, ref.data)
        assert ts.header.nsamples == ref.header.nsamples


def test_running_normalise_vs_oracle(rng):
    from sigpyproc3_amd.sweep import running_normalise

    nrows, ld = 5, 30000
    x = rng.integers(0, 5000, (nrows, ld)).astype(np.float32)
    row_len = np.array([30000, 29000, 4096, 5000, 12345])
    for w in (4096, 1001, 64):
        z = running_normalise(x, row_len, w)
        for i in range(nrows):
            n = row_len[i]
            ref = O.running_normalise(x[i, :n], w)
            np.testing.assert_allclose(z[i, :n], ref, rtol=2e-7, atol=1e-6)


def test_boxcar_mf_vs_oracle(rng):
    from sigpyproc3_amd.sweep import boxcar_mf

    nrows, ld = 6, 20000
    z = rng.normal(size=(nrows, ld)).astype(np.float32)
    z[2, 5000:5016] += 3.0
    z[4, 100:108] += 4.0
    row_len = np.array([20000, 19500, 17001, 9000, 2048, 4097])
    widths = [1, 2, 4, 8, 16, 32, 64, 128, 256]
    snr, peak, widx = boxcar_mf(z, row_len, widths)
    for i in range(nrows):
        n = row_len[i]
        conv = O.boxcar_conv_closed(z[i, :n], widths).astype(np.float32)
        it, pk, v = O.mf_argmax(conv)
        assert (widx[i], peak[i]) == (it, pk)
        assert abs(snr[i] - v) <= 1e-5 * abs(v) + 1e-5 * np.sqrt(widths[it])


@pytest.mark.parametrize("kind", ["boxcar", "gaussian", "lorentzian"])
def test_convolve_templates_vs_oracle(kind, rng):
    from sigpyproc3_amd.core import kernels
    from sigpyproc3_amd.core.filters import MatchedFilter, Template

    z = rng.normal(size=3001).astype(np.float32)
    if kind == "boxcar":
        widths = MatchedFilter.get_box_width_spacing(32, 1.5)
    else:
        widths = np.geomspace(1, 32, 9)
    bank = [getattr(Template, f"gen_{kind}")(w) for w in widths]
    convs = kernels.convolve_templates(z, [t.data for t in bank], [t.ref_bin for t in bank])
    for k, t in enumerate(bank):
        ref = O.template_conv_closed(z, t.data, t.ref_bin)
        np.testing.assert_allclose(convs[k], ref, rtol=1e-5, atol=1e-5)
        fft = O.convolve_templates_fft(z, [t.data], [t.ref_bin])[0]
        np.testing.assert_allclose(convs[k], fft, rtol=1e-5, atol=1e-5)


def _pulse_data():
    # tests/test_filters.py:10-25
    rng = np.random.default_rng(42)
    x = np.linspace(-10, 10, 1000, dtype=np.float32)
    return np.exp(-0.5 * (x**2)) + rng.normal(0, 0.1, 1000).astype(np.float32)


@pytest.mark.parametrize("temp_kind", ["boxcar", "gaussian", "lorentzian"])
def test_matched_filter_end_to_end(temp_kind):
    # tests/test_filters.py:163-168
    from sigpyproc3_amd.core.filters import MatchedFilter

    mf = MatchedFilter(_pulse_data(), temp_kind=temp_kind)
    assert mf.snr > 5
    assert 450 < mf.peak_bin < 550
    assert mf.convs.shape == (len(mf.temp_widths), 1000)
    assert mf.best_model.shape == (1000,)
    start, end = mf.on_pulse
    assert 0 <= start < end <= 1000


def test_matched_filter_detection_and_noise():
    # tests/test_filters.py:71-93
    from sigpyproc3_amd.core.filters import MatchedFilter

    mf = MatchedFilter(_pulse_data(), temp_kind="gaussian")
    assert abs(mf.peak_bin - 500) <= 5 and mf.snr > 10
    rng = np.random.default_rng(42)
    noise = rng.normal(loc=5, scale=2, size=1000).astype(np.float32)
    assert MatchedFilter(noise).snr < 6


# ---------------------------------------------------------------------------
# gulp pre-steps: collapse, bandpass, channel mask, zero-DM removal
# ---------------------------------------------------------------------------


def _fixture(filfiles, key):
    return filfiles[key] if not isinstance(filfiles[key], list) else filfiles[key]


def _stream(path):
    files = path if isinstance(path, list) else [path]
    pieces = [O.read_sigproc(f) for f in files]
    hdr = pieces[0][0]
    return hdr, np.concatenate([O.unpacked_stream(h, r) for h, r in pieces]), sum(h["nsamples"] for h, _ in pieces)


@pytest.mark.parametrize("key", ["1bit", "2bit", "4bit", "8bit", "tutorial"])
def test_collapse_vs_oracle(filfiles, key):
    from sigpyproc3_amd import FilReader

    path = _fixture(filfiles, key)
    hdr, s, n = _stream(path)
    fil = FilReader(path)
    for gulp, method in ((16384, "sum"), (1000, "sum"), (777, "mean")):
        tim = fil.collapse(method=method, gulp=gulp)
        np.testing.assert_array_equal(tim.data, O.collapse(s, n, hdr["nchans"], gulp=gulp, method=method))
        assert tim.header.nchans == 1 and tim.header.dm == 0


@pytest.mark.parametrize("key", ["1bit", "2bit", "4bit", "8bit", "tutorial"])
def test_bandpass_vs_oracle(filfiles, key):
    from sigpyproc3_amd import FilReader

    path = _fixture(filfiles, key)
    hdr, s, n = _stream(path)
    fil = FilReader(path)
    for parallel in (False, True):
        for gulp, method in ((16384, "sum"), (999, "mean")):
            bp = fil.bandpass(method=method, gulp=gulp, parallel=parallel)
            ref = O.bandpass(s, n, hdr["nchans"], gulp=gulp, parallel=parallel, method=method)
            np.testing.assert_array_equal(bp.data, ref)
            assert bp.header.nsamples == hdr["nchans"]


def test_bandpass_float_sequential_order(tmp_path, rng):
    """float32 data: the sequential float32 accumulation order is reproduced exactly."""
    from sigpyproc3_amd import FilReader
    from sigpyproc3_amd.synth import SynthConfig

    cfg = SynthConfig(nchans=40, nsamps=5000, nbits=32, fch1=1400.0, foff=-1.0, tsamp=1e-3, seed=0)
    x = rng.normal(100.0, 30.0, (cfg.nsamps, cfg.nchans)).astype(np.float32)
    path = tmp_path / "f32.fil"
    from sigpyproc3_amd.io import sigproc

    with open(path, "wb") as fp:
        fp.write(sigproc.encode_header(cfg.header_dict()))
        fp.write(x.tobytes())
    bp = FilReader(str(path)).bandpass(gulp=1300)
    np.testing.assert_array_equal(bp.data, O.bandpass(x.reshape(-1), cfg.nsamps, cfg.nchans, gulp=1300))


@pytest.mark.parametrize("key", ["1bit", "2bit", "4bit", "8bit", "tutorial"])
def test_apply_channel_mask_vs_oracle(filfiles, key, tmp_path, rng):
    from sigpyproc3_amd import FilReader

    path = _fixture(filfiles, key)
    hdr, s, n = _stream(path)
    fil = FilReader(path)
    mask = rng.random(hdr["nchans"]) < 0.25
    for value in (0, 3, 200):
        out = fil.apply_channel_mask(mask, mask_value=value, outfile_name=str(tmp_path / f"m{value}.fil"),
                                     gulp=1500)
        h2, raw2 = O.read_sigproc(out)
        v = np.float32(value).astype(fil.header.dtype)
        masked = O.mask_channels(s, hdr["nchans"], mask, v)
        expect = O.pack(masked, hdr["nbits"], O.DEFAULT_BITORDER[hdr["nbits"]]) if hdr["nbits"] < 8 else masked
        np.testing.assert_array_equal(raw2, expect)
        assert (h2["nchans"], h2["nbits"], h2["nsamples"]) == (hdr["nchans"], hdr["nbits"], n)


def _zerodm_kernel(s_packed, nbits, big, nchans, nsamps, bpass, out_dtype):
    from sigpyproc3_amd import _lib

    w = bpass / bpass.sum()
    dev = _lib.to_device(np.ascontiguousarray(s_packed))
    out = _lib.empty((nsamps * nchans,), "float32" if out_dtype == np.float32 else "uint8")
    w_d = _lib.to_device(w.astype(np.float32))       # keep the device copies alive
    b_d = _lib.to_device(bpass.astype(np.float32))
    _lib.call("spp_remove_zerodm", _lib.ptr(dev), nbits, big, nchans, nsamps, _lib.ptr(w_d),
              _lib.ptr(b_d), _lib.ptr(out), _lib.stream_handle())
    return out.cpu().numpy()


@pytest.mark.parametrize("nbits", [2, 8, 32])
def test_remove_zerodm_kernel_vs_exact_oracle(nbits, rng):
    """Small inputs against the fused-multiply-add oracle evaluated in exact arithmetic."""
    nchans, nsamps = 48, 300
    if nbits == 32:
        x = rng.normal(20.0, 5.0, (nsamps, nchans)).astype(np.float32)
        packed, out_dtype = x.view(np.uint8), np.float32
    else:
        x = rng.integers(0, 1 << nbits, (nsamps, nchans), dtype=np.uint8)
        packed = x if nbits == 8 else O.pack(x.reshape(-1), nbits, "big")
        out_dtype = np.uint8
    bpass = rng.uniform(1.0e3, 5.0e4, nchans).astype(np.float32)
    got = _zerodm_kernel(packed, nbits, 1, nchans, nsamps, bpass, out_dtype)
    np.testing.assert_array_equal(got, O.remove_zerodm(x.reshape(-1), nchans, bpass, out_dtype, fused=True))


@pytest.mark.parametrize("key", ["2bit", "8bit", "tutorial"])
def test_remove_zerodm_file_vs_oracle(filfiles, key, tmp_path, rng):
    from sigpyproc3_amd import FilReader

    path = _fixture(filfiles, key)
    hdr, s, n = _stream(path)
    fil = FilReader(path)
    nch = hdr["nchans"]
    out = fil.remove_zerodm(outfile_name=str(tmp_path / "z.fil"), gulp=2000)
    bpass = O.bandpass(s, n, nch)        # Filterbank.bandpass() defaults, float32 sums
    nbits = hdr["nbits"]
    first = path[0] if isinstance(path, list) else path
    raw = np.concatenate([O.read_sigproc(f)[1] for f in (path if isinstance(path, list) else [path])])
    vals = _zerodm_kernel(raw, nbits, 1 if nbits in (2, 4) else 0, nch, n, bpass, np.uint8)
    # the file holds the kernel's values packed the reference's way
    h2, raw2 = O.read_sigproc(out)
    expect = O.pack(vals, nbits, O.DEFAULT_BITORDER[nbits]) if nbits < 8 else vals
    np.testing.assert_array_equal(raw2, expect)
    assert (h2["nbits"], h2["nchans"]) == (nbits, nch) and first
    # values: exact at sampled positions, within one count of the unfused form everywhere
    x = s.reshape(-1, nch)
    idx = rng.integers(0, x.shape[0], 200)
    sub = O.remove_zerodm(x[idx].reshape(-1), nch, bpass, np.uint8, fused=True).reshape(-1, nch)
    # z (the zero-DM) of each sampled row is computed from its own row, so rows can be sampled
    np.testing.assert_array_equal(vals.reshape(-1, nch)[idx], sub)
    loose = O.remove_zerodm(s, nch, bpass, np.uint8, fused=False)
    diff = (vals.astype(np.int16) - loose.astype(np.int16)) % 256
    bad = np.minimum(diff, 256 - diff)
    assert bad.max() <= 1 and np.count_nonzero(bad) <= 1e-3 * bad.size


def test_gulp_pipeline_matches_single_buffer_and_stops_early(filfiles):
    """Double-buffered reads give the same results as the caller-allocator path, and a
    consumer that stops early ends the reader thread."""
    from sigpyproc3_amd import FilReader
    from sigpyproc3_amd.base import _GulpPipeline

    fil = FilReader(filfiles["tutorial"])
    gulp = 3000
    buf = bytearray(2 * gulp * fil.samp_stride)
    a = fil.dedisperse(30, gulp=gulp)
    b = fil.dedisperse(30, gulp=gulp, allocator=lambda n: buf[:n])
    np.testing.assert_array_equal(a.data, b.data)
    pipe = _GulpPipeline(fil, gulp, 0, None, 0)
    it = iter(pipe)
    next(it)
    it.close()
    assert not pipe.thread.is_alive()


# ---------------------------------------------------------------------------
# in-memory blocks: FilterbankBlock.dedisperse / dmt_transform, read_block
# ---------------------------------------------------------------------------


def test_read_block_vs_stream(filfiles):
    from sigpyproc3_amd import FilReader

    fil = FilReader(filfiles["4bit"])
    hdr, s, n = _stream(filfiles["4bit"])
    blk = fil.read_block(100, 1024)
    np.testing.assert_array_equal(blk.data, s.reshape(n, hdr["nchans"])[100:1124].T.astype(np.float32))
    assert blk.header.nsamples == 1024 and blk.nchans == hdr["nchans"]
    with pytest.raises(ValueError, match="out of range"):
        fil.read_block(n - 10, 100)


@pytest.mark.parametrize("dm", [0.0, 50.0, 200.0])
def test_block_dedisperse_vs_oracle(filfiles, dm, rng):
    from sigpyproc3_amd import FilReader
    from sigpyproc3_amd.block import FilterbankBlock

    fil = FilReader(filfiles["8bit"])
    hdr = fil.header
    data = rng.normal(0, 50, (hdr.nchans, hdr.nsamples)).astype(np.float32)
    block = FilterbankBlock(data, hdr)
    d = np.atleast_1d(hdr.get_dmdelays(dm)).astype(np.int64)
    out = block.dedisperse(dm)
    np.testing.assert_array_equal(out.data, O.roll_block(data, -d))
    assert out.dm == dm and out.data.shape == data.shape
    out_v = block.dedisperse(dm, only_valid_samples=True)
    np.testing.assert_array_equal(out_v.data, O.roll_block_valid(data, -d))
    assert out_v.header.nsamples == hdr.nsamples - int(d.max())


def test_block_dedisperse_direction(filfiles):
    # tests/test_blocks.py:100-124: an impulse at the middle stays at or before the middle
    from sigpyproc3_amd import FilReader
    from sigpyproc3_amd.block import FilterbankBlock

    fil = FilReader(filfiles["8bit"])
    data = np.zeros((fil.header.nchans, fil.header.nsamples), dtype=np.float32)
    block = FilterbankBlock(data, fil.header)
    block.data[:, block.nsamples // 2] = 1.0
    dd = block.dedisperse(50)
    np.testing.assert_array_equal(dd.data.mean(axis=1), block.data.mean(axis=1))
    assert dd.data[:, block.nsamples // 2 + 1 :].sum() == 0
    assert block.dedisperse(-50).data[:, : block.nsamples // 2].sum() == 0
    with pytest.raises(ValueError):
        fil.read_block(100, 1024).dedisperse(10000, only_valid_samples=True)


@pytest.mark.parametrize("dm", [0.0, 30.0, 120.0])
def test_dmt_transform_vs_oracle(filfiles, dm, rng):
    from sigpyproc3_amd import FilReader

    fil = FilReader(filfiles["8bit"])
    block = fil.read_block(0, 2048)
    for dmsteps in (16, 63):
        dmt = block.dmt_transform(dm, dmsteps=dmsteps)
        dm_arr = dm + np.linspace(-dm, dm, dmsteps)
        delays = np.atleast_2d(block.header.get_dmdelays(dm_arr)).astype(np.int64)
        np.testing.assert_array_equal(dmt.data, O.dmt_block(block.data, delays))
        np.testing.assert_array_equal(dmt.dms, dm_arr.astype(np.float32))
        assert dmt.header.nchans == 1
    # the valid-samples form keeps the block's nsamples in its header, so any DM range with
    # delays fails the length check (block.py:337-343), as in the reference
    if dm > 0:
        with pytest.raises(ValueError):
            block.dmt_transform(dm, dmsteps=8, only_valid_samples=True)
    else:
        v = block.dmt_transform(dm, dmsteps=8, only_valid_samples=True)
        np.testing.assert_array_equal(v.data, O.dmt_block(block.data, np.zeros((8, block.nchans), int), valid=True))


# ---------------------------------------------------------------------------
# fold
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("key", ["1bit", "2bit", "4bit", "8bit", "tutorial"])
def test_fold_vs_oracle(filfiles, key):
    from sigpyproc3_amd import FilReader

    path = _fixture(filfiles, key)
    hdr, s, n = _stream(path)
    fil = FilReader(path)
    for period, dm, accel, gulp in ((0.0123, 20.0, 0.0, 1500), (0.00714, 0.0, 3.5, 16384)):
        folded = fil.fold(period, dm, accel=accel, nbins=16, nints=8, nbands=8, gulp=gulp)
        d = np.atleast_1d(fil.header.get_dmdelays(dm))
        ref = O.filterbank_fold(s, n, hdr["nchans"], hdr["tsamp"], d, period, accel=accel, nbins=16,
                                nints=8, nbands=8, gulp=gulp)
        np.testing.assert_array_equal(folded.data, ref)
        assert (folded.nsubints, folded.nsubbands, folded.nbins) == (8, 8, 16)
        prof = folded.get_profile().data
        np.testing.assert_array_equal(prof, np.sum(ref, axis=(0, 1)))
    with pytest.raises(ValueError, match="too large"):
        fil.fold(0.01, 0.0, nbins=4096, nints=4096, nbands=64)


def test_gulp_loop_graph_replays_match_eager(rng):
    """The one-DM gulp loop over resident data captured as a HIP graph: three replays, with
    eager work and synchronisation between them, each equal bit for bit to the eager pass
    (and to the oracle)."""
    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.readers import plan_blocks

    nchans, nsamps, nbits = 256, 40000, 2
    samples = rng.integers(0, 4, nsamps * nchans, dtype=np.uint8)
    packed = O.pack(samples, nbits, "big")
    f = O.chan_freqs(nchans, -0.5, 1500.0)
    delays = O.compute_dmdelays(f, 120.0, 64e-6, 1500.0).astype(np.int32)
    md = int(delays.max())
    gulp = max(2 * md, 8192)
    _, blocks = plan_blocks(nsamps, gulp, md)
    dev = _lib.to_device(packed)
    d_dev = _lib.to_device(delays)
    tim = _lib.zeros((nsamps - md,), "float32")
    row_bytes = nchans * nbits // 8
    calls, pos = [], 0
    for ii, n, skip in blocks:
        calls.append((ii, n, pos))
        pos += n + skip

    def one_pass(st):
        for ii, n, p0 in calls:
            _lib.call("spp_dedisperse", _lib.ptr(dev[p0 * row_bytes:]), nbits, 1, nchans, n,
                      _lib.ptr(d_dev), md, _lib.ptr(tim), ii * (gulp - md), st)

    one_pass(_lib.stream_handle())
    torch.cuda.synchronize()
    eager = tim.cpu().numpy()
    ref = np.zeros(nsamps - md, np.float32)
    c_oracle.dedisperse(samples, ref, delays, nchans, nsamps, 0)
    np.testing.assert_array_equal(eager, ref)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        one_pass(_lib.stream_handle())
    for rep in range(3):
        tim.fill_(float(rep + 7))          # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tim.cpu().numpy(), eager, err_msg=f"replay {rep}")


@pytest.mark.parametrize("nbits", [1, 2, 4, 8])
@pytest.mark.parametrize("nchans", [256, 336])
@pytest.mark.parametrize(("nsamps", "misalign"), [(2600, 0), (2600, 1), (60000, 0)])
def test_packed_dedisperse_staging_paths(nbits, nchans, nsamps, misalign, rng):
    """spp_dedisperse on packed input: short streams split channels into narrow groups (the
    16-channel tile), long ones use the 64-channel tile; 336-channel 1/2-bit rows (42/84 B)
    and a stream starting one byte off a word boundary take the byte-unit staging."""
    from sigpyproc3_amd import _lib

    order = O.DEFAULT_BITORDER.get(nbits, "big")
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    packed = samples if nbits == 8 else O.pack(samples, nbits, order)
    buf = np.zeros(packed.size + 4, np.uint8)
    buf[misalign:misalign + packed.size] = packed
    f = O.chan_freqs(nchans, -1.0, 1500.0)
    delays = O.compute_dmdelays(f, 40.0 if nsamps < 10000 else 150.0, 64e-6, 1500.0).astype(np.int32)
    md = int(delays.max())
    dev = _lib.to_device(buf)
    d_dev = _lib.to_device(delays)
    out = _lib.zeros((nsamps - md,), "float32")
    _lib.call("spp_dedisperse", _lib.ptr(dev[misalign:]), nbits, 1 if order == "big" else 0, nchans, nsamps,
              _lib.ptr(d_dev), md, _lib.ptr(out), 0, _lib.stream_handle())
    ref = np.zeros(nsamps - md, np.float32)
    c_oracle.dedisperse(samples, ref, delays, nchans, nsamps, 0)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("nbits", [2, 8])
@pytest.mark.parametrize("pattern", ["dm0", "runs5", "runs4_off_grid"])
def test_packed_gulp_shared_delay_runs(nbits, pattern, rng):
    """Channels in runs of one delay. At DM 0 every 4-channel dword of a tile row is summed by
    one read and one byte sum; runs of 5 mix those with per-channel reads, and runs of 4 that
    start at channel 2 never line up with the dword grid. Dedisperse, and subband with
    6-channel bands (band changes off the 4-channel grid)."""
    from sigpyproc3_amd import _lib

    nchans, nsamps = 1024, 40000
    order = O.DEFAULT_BITORDER.get(nbits, "big")
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    packed = samples if nbits == 8 else O.pack(samples, nbits, order)
    ch = np.arange(nchans)
    if pattern == "dm0":
        delays = np.zeros(nchans, np.int32)
    elif pattern == "runs5":
        delays = ((ch // 5) * 2).astype(np.int32)
    else:
        delays = (((ch + 2) // 4) * 3).astype(np.int32)
    md = int(delays.max())
    big = 1 if order == "big" else 0
    dev = _lib.to_device(packed)
    d_dev = _lib.to_device(delays)
    out = _lib.zeros((nsamps - md,), "float32")
    _lib.call("spp_dedisperse", _lib.ptr(dev), nbits, big, nchans, nsamps, _lib.ptr(d_dev), md,
              _lib.ptr(out), 0, _lib.stream_handle())
    ref = np.zeros(nsamps - md, np.float32)
    c_oracle.dedisperse(samples, ref, delays, nchans, nsamps, 0)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    cts = (ch // 6).astype(np.int32)
    nsubs = int(cts.max()) + 1
    sub = _lib.zeros(((nsamps - md) * nsubs,), "float32")
    cts_dev = _lib.to_device(cts)
    _lib.call("spp_subband", _lib.ptr(dev), nbits, big, nchans, nsamps, _lib.ptr(d_dev), _lib.ptr(cts_dev),
              md, nsubs, _lib.ptr(sub), int(sub.numel()), _lib.stream_handle())
    ref_sub = np.zeros((nsamps - md) * nsubs, np.float32)
    c_oracle.subband(samples, ref_sub, delays, cts, md, nchans, nsubs, nsamps)
    np.testing.assert_array_equal(sub.cpu().numpy(), ref_sub)


@pytest.mark.parametrize(("nbits", "nchans"), [(1, 1024), (2, 1024), (4, 1024), (8, 1024), (2, 1000)])
@pytest.mark.parametrize("dm", [10.0, 200.0])
def test_packed_dedisperse_long_gulps_line_groups(nbits, nchans, dm, rng):
    """Gulps of 2^17 samples split channels into 64-channel groups. At DM 10 the delay spread
    across a 128-byte row line is small, so each block takes chunks one line apart (16, 8, 4 or
    2 groups per line for 1/2/4/8-bit data; 1000 channels end in a partial chunk and stage
    bytes, 250-byte rows). At DM 200 the 1/2/4-bit lines span too many delays and every group
    is one contiguous chunk; 8-bit lines (128 channels) stay interleaved, two chunks a block."""
    from sigpyproc3_amd import _lib

    nsamps = 1 << 17
    order = O.DEFAULT_BITORDER.get(nbits, "big")
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    packed = samples if nbits == 8 else O.pack(samples, nbits, order)
    f = O.chan_freqs(nchans, -300.0 / nchans, 1500.0)
    delays = O.compute_dmdelays(f, dm, 64e-6, 1500.0).astype(np.int32)
    md = int(delays.max())
    dev = _lib.to_device(packed)
    d_dev = _lib.to_device(delays)
    out = _lib.zeros((nsamps - md,), "float32")
    _lib.call("spp_dedisperse", _lib.ptr(dev), nbits, 1 if order == "big" else 0, nchans, nsamps,
              _lib.ptr(d_dev), md, _lib.ptr(out), 0, _lib.stream_handle())
    ref = np.zeros(nsamps - md, np.float32)
    c_oracle.dedisperse(samples, ref, delays, nchans, nsamps, 0)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    # Subband with 8 bands: chunks of whole 128-channel bands (125-channel bands for 1000
    # channels straddle the chunks).
    nsubs = 8
    cts = (np.arange(nchans) // (nchans // nsubs)).astype(np.int32)
    sub = _lib.zeros(((nsamps - md) * nsubs,), "float32")
    cts_dev = _lib.to_device(cts)
    _lib.call("spp_subband", _lib.ptr(dev), nbits, 1 if order == "big" else 0, nchans, nsamps, _lib.ptr(d_dev),
              _lib.ptr(cts_dev), md, nsubs, _lib.ptr(sub), int(sub.numel()), _lib.stream_handle())
    ref_sub = np.zeros((nsamps - md) * nsubs, np.float32)
    c_oracle.subband(samples, ref_sub, delays, cts, md, nchans, nsubs, nsamps)
    np.testing.assert_array_equal(sub.cpu().numpy(), ref_sub)


def test_running_normalise_long_float_rows_and_wide_windows(rng):
    """Rows spanning several normalisation blocks with non-integer data (the sliding window
    sums round), and windows longer than the row (the mirrored padding covers the window)."""
    from sigpyproc3_amd.sweep import running_normalise

    x = (rng.normal(size=(3, 100000)) * 3.0 + 50.0).astype(np.float32)
    row_len = np.array([100000, 70001, 3000])
    for w in (4096, 1001):
        z = running_normalise(x, row_len, w)
        for i in range(3):
            n = row_len[i]
            np.testing.assert_allclose(z[i, :n], O.running_normalise(x[i, :n], w), rtol=1e-6, atol=2e-6)
    y = rng.integers(0, 300, (2, 1500)).astype(np.float32)
    z = running_normalise(y, np.array([1500, 700]), 4096)
    for i, n in enumerate((1500, 700)):
        np.testing.assert_allclose(z[i, :n], O.running_normalise(y[i, :n], 4096), rtol=2e-7, atol=1e-6)


def test_boxcar_mf_many_widths(rng):
    from sigpyproc3_amd.sweep import boxcar_mf

    z = rng.normal(size=(3, 9000)).astype(np.float32)
    z[1, 4000:4100] += 1.5
    row_len = np.array([9000, 8500, 5000])
    widths = list(range(1, 21)) + [50, 100, 200, 400, 1000]
    snr, peak, widx = boxcar_mf(z, row_len, widths)
    for i in range(3):
        n = row_len[i]
        conv = O.boxcar_conv_closed(z[i, :n], widths).astype(np.float32)
        it, pk, v = O.mf_argmax(conv)
        assert (widx[i], peak[i]) == (it, pk)
        assert abs(snr[i] - v) <= 1e-5 * abs(v) + 1e-5 * np.sqrt(widths[it])
