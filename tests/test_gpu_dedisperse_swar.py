"""GPU parity of the word-parallel one-DM kernel (k_gulp_swar), the spp_dedisperse path for
rows of whole dwords, against the C oracle, bit for bit; and of spp_subband on irregular band
maps.

Cases beyond the operator and Filterbank tests in test_gpu_parity.py: delay patterns whose
spread inside one stage (16 bytes of a row) or inside one dword column exceeds what one pass
stages (random delays up to 3000 samples: passes of single channels), negative delays
(numba's wraparound of a negative row index), the 2-bit 4096-channel gulp at DM 0, 100 and
1000 (line-interleaved groups and one-chunk groups), channel counts that exercise the
16-bit counter widening, the compact series buffer of 4/8-bit launches with a stage wider than
one compact pass, and a stream whose packed rows are not whole dwords (the byte-staged kernel)
beside one that is.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_oracle
from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu


def _run(samples, nbits, nchans, nsamps, delays, *, offset=0):
    from sigpyproc3_amd import _lib

    order = O.DEFAULT_BITORDER.get(nbits, "big")
    packed = samples if nbits == 8 else O.pack(samples, nbits, order)
    buf = np.zeros(packed.size + 8, np.uint8)
    buf[offset : offset + packed.size] = packed
    md = int(np.abs(delays).max())
    dev = _lib.to_device(buf)
    d_dev = _lib.to_device(delays.astype(np.int32))
    out = _lib.zeros((nsamps - md,), "float32")
    _lib.call("spp_dedisperse", _lib.ptr(dev[offset:]), nbits, 1 if order == "big" else 0, nchans, nsamps,
              _lib.ptr(d_dev), md, _lib.ptr(out), 0, _lib.stream_handle())
    ref = np.zeros(nsamps - md, np.float32)
    c_oracle.dedisperse(samples, ref, delays.astype(np.int32), nchans, nsamps, 0)
    return out.cpu().numpy(), ref


@pytest.mark.parametrize("nbits", [1, 2, 4, 8])
def test_random_wide_delays(nbits):
    rng = np.random.default_rng(nbits)
    nchans, nsamps = 256, 12000
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    delays = rng.integers(0, 3000, nchans)
    print(f"{nbits}-bit, random delays up to {delays.max()}", flush=True)
    got, ref = _run(samples, nbits, nchans, nsamps, delays)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("nbits", [1, 2, 4, 8])
def test_negative_delays_wrap(nbits):
    rng = np.random.default_rng(10 + nbits)
    nchans, nsamps = 128, 9000
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    delays = rng.integers(-400, 400, nchans)
    print(f"{nbits}-bit, delays in [-400, 400)", flush=True)
    got, ref = _run(samples, nbits, nchans, nsamps, delays)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("dm", [0.0, 100.0, 1000.0])
def test_2bit_4096_channel_gulp(dm):
    rng = np.random.default_rng(int(dm))
    nchans, nsamps = 4096, 1 << 16
    samples = rng.integers(0, 4, nsamps * nchans, dtype=np.uint8)
    f = O.chan_freqs(nchans, -300.0 / nchans, 1550.0)
    delays = O.compute_dmdelays(f, dm, 64e-6, 1550.0).astype(np.int32)
    print(f"2-bit 4096 ch x 2^16, DM {dm}: max delay {delays.max()}", flush=True)
    got, ref = _run(samples, 2, nchans, nsamps, delays)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize(("nbits", "nchans"), [(2, 2048), (4, 2048), (8, 1024), (1, 4096)])
def test_counter_widening_many_channels(nbits, nchans):
    """All samples at their maximum over many channels in few blocks: every counter level
    carries its largest sums (a short stream keeps one channel group per time tile)."""
    nsamps = 3000
    samples = np.full(nsamps * nchans, (1 << nbits) - 1, dtype=np.uint8)
    samples[::7] = 0
    delays = (np.arange(nchans) // 64).astype(np.int32)
    print(f"{nbits}-bit, {nchans} ch at full scale", flush=True)
    got, ref = _run(samples, nbits, nchans, nsamps, delays)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("nbits", [4, 8])
def test_compact_form_with_one_wide_stage(nbits):
    """4/8-bit launches whose largest delay is small next to the channel count take the
    compact series buffer (fewer rows per pass); one stage here still spans 600 samples, more
    than a compact pass holds, and is split into passes."""
    rng = np.random.default_rng(30 + nbits)
    nchans, nsamps = 512, 9000
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    delays = rng.integers(0, 40, nchans)
    delays[64:96] = np.linspace(0, 600, 32).astype(np.int64)  # one wide stage of channels
    print(f"{nbits}-bit, compact form, one stage spanning 600 samples", flush=True)
    got, ref = _run(samples, nbits, nchans, nsamps, delays)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("nbits", [1, 2, 4, 8])
def test_unaligned_stream_and_odd_rows(nbits):
    rng = np.random.default_rng(20 + nbits)
    nsamps = 7000
    for nchans, offset in [(96, 1), (72 if nbits < 8 else 36, 0)]:  # misaligned start; rows not whole dwords
        if (nchans * nbits) % 8:
            continue
        samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
        delays = np.sort(rng.integers(0, 500, nchans))[::-1].copy()
        print(f"{nbits}-bit, {nchans} ch, offset {offset}", flush=True)
        got, ref = _run(samples, nbits, nchans, nsamps, delays, offset=offset)
        np.testing.assert_array_equal(got, ref)


def _subband(samples, nbits, nchans, nsamps, delays, cts):
    from sigpyproc3_amd import _lib

    order = O.DEFAULT_BITORDER.get(nbits, "big")
    packed = samples if nbits == 8 else O.pack(samples, nbits, order)
    md = int(delays.max())
    nsubs = int(cts.max()) + 1
    dev = _lib.to_device(packed)
    d_dev = _lib.to_device(delays.astype(np.int32))
    c_dev = _lib.to_device(cts.astype(np.int32))
    sub = _lib.zeros(((nsamps - md) * nsubs,), "float32")
    _lib.call("spp_subband", _lib.ptr(dev), nbits, 1 if order == "big" else 0, nchans, nsamps, _lib.ptr(d_dev),
              _lib.ptr(c_dev), md, nsubs, _lib.ptr(sub), int(sub.numel()), _lib.stream_handle())
    ref = np.zeros((nsamps - md) * nsubs, np.float32)
    c_oracle.subband(samples, ref, delays.astype(np.int32), cts.astype(np.int32), md, nchans, nsubs, nsamps)
    return sub.cpu().numpy(), ref


@pytest.mark.parametrize("nbits", [2, 4, 8])
@pytest.mark.parametrize("kind", ["uneven", "interleaved"])
def test_subband_band_maps(nbits, kind):
    """spp_subband with uneven contiguous bands whose edges fall anywhere (inside dwords and
    staged column groups), and with bands whose channels interleave (a band run ends every 37
    channels), against the C oracle."""
    rng = np.random.default_rng(50 + nbits + (0 if kind == "uneven" else 7))
    nchans, nsamps = 1024, 20000
    samples = rng.integers(0, 1 << nbits, nsamps * nchans, dtype=np.uint8)
    f = O.chan_freqs(nchans, -300.0 / nchans, 1500.0)
    delays = O.compute_dmdelays(f, 60.0, 64e-6, 1500.0).astype(np.int64)
    if kind == "uneven":
        edges = np.sort(rng.choice(np.arange(1, nchans), 11, replace=False))
        cts = np.searchsorted(edges, np.arange(nchans), side="right")
    else:
        cts = (np.arange(nchans) // 37) % 6
    print(f"{nbits}-bit subband, {kind} bands ({int(cts.max()) + 1})", flush=True)
    got, ref = _subband(samples, nbits, nchans, nsamps, delays, cts)
    np.testing.assert_array_equal(got, ref)


def test_subband_full_scale_counters():
    """2-bit samples at their maximum in wide bands: every band run's sum is the largest it
    can be."""
    nchans, nsamps = 2048, 6000
    samples = np.full(nsamps * nchans, 3, dtype=np.uint8)
    samples[::5] = 0
    delays = (np.arange(nchans) // 32).astype(np.int64)
    cts = np.arange(nchans) // 512
    got, ref = _subband(samples, 2, nchans, nsamps, delays, cts)
    np.testing.assert_array_equal(got, ref)
