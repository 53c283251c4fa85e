"""Single-DM Filterbank.dedisperse of an 8-bit 1024-channel 65 536-sample file at DM 300.

The file (fch1 1500 MHz, foff -0.390625 MHz, tsamp 64 us, samples round(N(128, 16))
clipped to [0, 255], a DM-300 boxcar of width 16 and +24 starting at sample 20 000) is
written as SIGPROC and read back through FilReader, as a user of the reference would
(base.py:360-424 -> readers.py:171-233). The output must equal the oracle's
Filterbank.dedisperse bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu


def test_single_dm_dedisperse_8bit_1024ch(tmp_path):
    from sigpyproc3_amd import FilReader, synth

    cfg = synth.cfg_single_dm()
    packed = synth.generate_packed(cfg, device="cpu")
    path = tmp_path / "single_dm_8bit_1024ch.fil"
    synth.write_sigproc(path, cfg, packed)
    print("single-DM 8-bit 1024 ch x 65536, DM 300", flush=True)

    fil = FilReader(str(path))
    assert (fil.header.nchans, fil.header.nsamples, fil.header.nbits) == (1024, 65536, 8)
    tim = fil.dedisperse(300.0)

    delays = np.atleast_2d(O.compute_dmdelays(O.chan_freqs(1024, -0.390625, 1500.0), [300.0], 64e-6, 1500.0))[0]
    md = int(np.abs(delays).max())
    assert md == 7418
    assert tim.data.size == tim.header.nsamples == 65536 - md == 58118
    # the reference's read schedule for this file: six 16384-sample gulps and one of 11740
    blocks = O.read_plan_blocks(65536, max(2 * md, 16384), md)
    assert [n for _, n, _ in blocks] == [16384] * 6 + [11740]

    stream = np.frombuffer(packed.numpy().tobytes(), dtype=np.uint8).reshape(65536, 1024)
    ref = O.filterbank_dedisperse(stream, 65536, 1024, delays.astype(np.int32))
    assert ref.dtype == np.float32 and tim.data.dtype == np.float32
    np.testing.assert_array_equal(tim.data, ref)

    # the injected pulse: 16 samples of +24 in every channel from t = 20000
    peak = int(np.argmax(tim.data))
    assert 20000 <= peak < 20016
    assert tim.data[20000:20016].min() > np.median(tim.data) + 10 * 1024


def test_resident_stream_one_launch_equals_gulp_path(tmp_path):
    """The same file's stream resident in HBM, dedispersed in one launch, equals the gulp
    path; with mixed-sign delays (reference frequency mid-band) the resident path keeps the
    reference's gulp schedule and still matches."""
    import torch

    from sigpyproc3_amd import FilReader, synth
    from sigpyproc3_amd.sweep import dedisperse_resident, load_packed_stream

    cfg = synth.cfg_single_dm()
    packed = synth.generate_packed(cfg, device="cpu")
    path = tmp_path / "single_dm_8bit_1024ch.fil"
    synth.write_sigproc(path, cfg, packed)
    fil = FilReader(str(path))
    dev = load_packed_stream(fil, 0, fil.header.nsamples)
    for ref in ("fch1", "fcenter"):
        print(f"resident one-launch vs gulp path, ref_freq={ref}", flush=True)
        a = fil.dedisperse(300.0, ref_freq=ref)
        b = dedisperse_resident(dev, fil.header, 300.0, ref_freq=ref)
        torch.cuda.synchronize()
        assert a.data.shape == b.data.shape
        np.testing.assert_array_equal(a.data, b.data)
        assert b.header.nsamples == a.header.nsamples
