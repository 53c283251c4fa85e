"""Host-side logic and the C-ABI library surface (CPU only: no kernel is launched)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import spp_oracle as O
from sigpyproc3_amd import FilReader, _lib, params
from sigpyproc3_amd.core.filters import MatchedFilter, Template
from sigpyproc3_amd.core.kernels import good_size_real
from sigpyproc3_amd.io import bits, sigproc
from sigpyproc3_amd.readers import plan_blocks

ROOT = Path(__file__).resolve().parent.parent


def _declared_symbols() -> list[str]:
    text = (ROOT / "include" / "spp_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared_symbols()
    assert len(names) >= 17
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)
    assert lib.spp_version() == 1


def test_sweep_info_struct_layout():
    text = (ROOT / "include" / "spp_hip.h").read_text()
    body = text[text.index("typedef struct {") : text.index("} spp_sweep_info;")]
    fields = re.findall(r"(int64_t|int32_t)\s+(\w+);", body)
    assert [f[1] for f in fields] == [f[0] for f in _lib.SweepInfo._fields_]
    for (ctype, _), (_, pytype) in zip(fields, _lib.SweepInfo._fields_):
        assert ctypes.sizeof(pytype) == (8 if ctype == "int64_t" else 4)


def test_invalid_arguments_map_to_valueerror():
    lib = _lib.load()
    rc = lib.spp_unpack(None, 10, None, 3, 1, None)
    assert rc == _lib.SPP_EINVAL
    assert b"nbits" in lib.spp_last_error()
    with pytest.raises(ValueError, match="nbits"):
        _lib.check(rc)
    plan = ctypes.c_void_p()
    neg = np.array([[0, -1]], dtype=np.int32)
    with pytest.raises(ValueError, match="non-negative"):
        _lib.check(lib.spp_sweep_plan_create(neg.ctypes.data, 1, 2, 100, 8, 0, ctypes.byref(plan)))
    big = np.array([[0, 200]], dtype=np.int32)
    with pytest.raises(ValueError, match="max delay"):
        _lib.check(lib.spp_sweep_plan_create(big.ctypes.data, 1, 2, 100, 8, 0, ctypes.byref(plan)))


def test_sweep_plan_tiles_and_lengths():
    lib = _lib.load()
    f = O.chan_freqs(4096, -0.0732421875, 1550.0)
    table = O.compute_dmdelays(f, np.arange(0, 2048), 64e-6, 1550.0).astype(np.int32)
    plan = ctypes.c_void_p()
    _lib.check(lib.spp_sweep_plan_create(table.ctypes.data, 2048, 4096, 1 << 20, 2, 1, ctypes.byref(plan)))
    info = _lib.SweepInfo()
    _lib.check(lib.spp_sweep_plan_info(plan, ctypes.byref(info)))
    rl = np.zeros(2048, dtype=np.int64)
    _lib.check(lib.spp_sweep_plan_row_lengths(plan, rl.ctypes.data))
    lib.spp_sweep_plan_destroy(plan)
    np.testing.assert_array_equal(rl, (1 << 20) - table.max(axis=1))
    assert info.max_len == (1 << 20)  # DM 0
    assert info.min_len == (1 << 20) - 29683  # max delay at DM 2047 (2-bit, 4096 ch, 2^20 samples)
    assert info.total_samples == rl.sum() == 2117087959
    assert info.ndm_tiles == 128 and info.max_spread <= 252


def test_header_matches_oracle(filfiles):
    for key in ("1bit", "2bit", "4bit", "tutorial"):
        fil = FilReader(filfiles[key])
        hdr, _ = O.read_sigproc(filfiles[key])
        assert fil.header.nchans == hdr["nchans"]
        assert fil.header.nbits == hdr["nbits"]
        assert fil.header.nsamples == hdr["nsamples"]
        assert fil.header.fch1 == hdr["fch1"] and fil.header.foff == hdr["foff"]
        assert fil.header.tsamp == hdr["tsamp"]
    multi = FilReader(filfiles["8bit"])
    assert multi.header.nsamples == 8192
    assert multi.header.telescope == "Parkes"


@pytest.mark.parametrize("ref", ["fch1", "fmax", "fmin", "fcenter", 4000.0])
def test_delays_match_oracle(filfiles, ref):
    fil = FilReader(filfiles["4bit"])
    h = fil.header
    rf = O.ref_frequency(ref, h.fch1, h.foff, h.nchans)
    for dm in (0, 100, 5000, np.arange(0, 300, 13)):
        np.testing.assert_array_equal(
            h.get_dmdelays(dm, ref), O.compute_dmdelays(O.chan_freqs(h.nchans, h.foff, h.fch1), dm, h.tsamp, rf))


def test_get_dmdelays_bad_ref(filfiles):
    h = FilReader(filfiles["4bit"]).header
    with pytest.raises(ValueError):
        h.get_dmdelays(100, "top")
    with pytest.raises(TypeError):
        h.get_dmdelays(100, [1.0])


def test_plan_blocks_match_oracle():
    for args in [(65536, 16384, 7418), (4096, 512, 200), (187520, 16384, 17), (1000, 1000, 0)]:
        g, blocks = plan_blocks(*args)
        assert blocks == O.read_plan_blocks(*args)
    with pytest.raises(ValueError):
        plan_blocks(1024, 512, 1024)


def test_read_plan_8bit_matches_stream(filfiles):
    # tests/test_readers.py:63-92: sizes, multi-file stream, skipback
    fil = FilReader(filfiles["8bit"])
    streams = [O.read_sigproc(f) for f in filfiles["8bit"]]
    full = np.concatenate([O.unpacked_stream(h, r) for h, r in streams]).reshape(-1, 832)
    pos = 0
    for nr, ii, data in fil.read_plan(gulp=512):
        assert data.size == 512 * 832
        np.testing.assert_array_equal(data.reshape(nr, 832), full[pos : pos + nr])
        pos += nr
    pos = 0
    for nr, ii, data in fil.read_plan(gulp=400, skipback=200):
        np.testing.assert_array_equal(data.reshape(nr, 832), full[pos : pos + nr])
        pos += nr - (200 if nr == 400 else 0)


def test_read_plan_packed_bytes(filfiles):
    fil = FilReader(filfiles["2bit"])
    hdr, raw = O.read_sigproc(filfiles["2bit"])
    rb = hdr["nchans"] * 2 // 8
    pos = 0
    for nr, _ii, data in fil.read_plan_packed(gulp=1000, skipback=100):
        np.testing.assert_array_equal(data, raw[pos * rb : (pos + nr) * rb])
        pos += nr - (100 if nr == 1000 else 0)


def test_parallel_packed_reads_match_sequential(filfiles):
    """creadinto_packed_parallel (pieces read by a thread pool at explicit offsets) returns the
    bytes of the sequential read and leaves the stream at the same position; a request that
    crosses into the next file of a multi-file stream, or past its end, takes the sequential
    path. Small pieces so that the fixtures split into many."""
    from concurrent.futures import ThreadPoolExecutor

    pool = ThreadPoolExecutor(4)
    for key in ("2bit", "8bit"):
        seq, par = FilReader(filfiles[key]), FilReader(filfiles[key])
        total = seq.header.nsamples * seq.samp_stride
        cases = [(0, 100_000), (1234, 333_333), (total - 50_000, 50_000), (total - 40_000, 80_000)]
        if key == "8bit":  # two files of 4096 samples: a read across the boundary
            cases.append((4096 * seq.samp_stride - 10_000, 100_000))
        for start, n in cases:
            a = np.zeros(n, np.uint8)
            b = np.zeros(n, np.uint8)
            seq._file.seek(start)
            par._file.seek(start)
            na = seq._file.creadinto_packed(a)
            nb = par._file.creadinto_packed_parallel(b, pool, chunk=4096)
            assert na == nb
            np.testing.assert_array_equal(a[:na], b[:nb])
            assert seq._file.cur_data_pos_stream == par._file.cur_data_pos_stream
    pool.shutdown()


def test_read_plan_truncated_file(filfiles, tmp_path):
    # tests/test_readers.py:80-88: a short read raises ValueError
    h, raw = O.read_sigproc(filfiles["8bit"][0])
    hdr = sigproc.encode_header({k: v for k, v in h.items() if k in sigproc.SIGPROC_SCHEMA}
                                | {"nsamples": 4096}, allow_nsamples_overwrite=True)
    p = tmp_path / "trunc.fil"
    p.write_bytes(hdr + raw[: 4096 * 832 - 100].tobytes())
    fil = FilReader(str(p))
    assert fil.header.nsamples == 4095
    fil._header = fil.header.new_header({"nsamples": 4096})
    with pytest.raises(ValueError, match="Unexpected number of bytes"):
        for _ in fil.read_plan(gulp=512):
            pass


def test_sigproc_roundtrip(tmp_path):
    hdr = {"telescope_id": 4, "machine_id": 0, "data_type": 1, "source_name": "J0534+2200",
           "src_raj": 53431.9, "src_dej": 220052.0, "tstart": 58543.3, "tsamp": 0.000512, "nbits": 8,
           "fch1": 4030.0, "foff": -4.0, "nchans": 16, "nifs": 1, "signed": 0}
    p = tmp_path / "x.fil"
    p.write_bytes(sigproc.encode_header(hdr) + bytes(16 * 10))
    back = sigproc.parse_header(p)
    for k, v in hdr.items():
        assert back[k] == v
    assert back["nsamples"] == 10


def test_prep_outfile_header(filfiles, tmp_path):
    fil = FilReader(filfiles["4bit"])
    out = fil.header.prep_outfile(str(tmp_path / "o.fil"), updates={"nchans": 52, "nbits": 32}, nbits=32)
    out.cwrite(np.zeros(52 * 3, dtype=np.float32))
    out.close()
    h2 = sigproc.parse_header(tmp_path / "o.fil")
    assert h2["nchans"] == 52 and h2["nbits"] == 32 and h2["nsamples"] == 3


def test_bits_validation():
    # tests/test_bits.py:89-108
    with pytest.raises(ValueError):
        bits.unpack(np.arange(255, dtype=np.uint8), nbits=10)
    with pytest.raises(ValueError):
        bits.unpack(np.arange(255, dtype=np.float32), nbits=4)
    with pytest.raises(ValueError):
        bits.unpack(np.arange(255, dtype=np.uint8), nbits=4, bitorder="invalid")
    with pytest.raises(ValueError):
        bits.unpack(np.arange(255, dtype=np.uint8), nbits=2, unpacked=np.zeros(255, np.uint8))
    out = bits.unpack(np.empty((0,), dtype=np.uint8), nbits=2)
    assert out.size == 0


@pytest.mark.parametrize("nbits", [1, 2, 4])
@pytest.mark.parametrize("order", ["big", "little"])
def test_bits_pack_matches_oracle(nbits, order, rng):
    a = rng.integers(0, 1 << nbits, 4096, dtype=np.uint8)
    np.testing.assert_array_equal(bits.pack(a, nbits, bitorder=order), O.pack(a, nbits, order))


def test_bitsinfo():
    # tests/test_bits.py:112-155
    b = bits.BitsInfo(4)
    assert (b.itemsize, b.unpack, b.bitfact, b.digi_max, b.digi_mean) == (1, True, 2, 15, 7.5)
    assert bits.BitsInfo(1).bitorder == "little" and bits.BitsInfo(2).bitorder == "big"
    arr = np.array([-15.5, -10.5, -5.5, -0.5, 0.0, 0.5, 5.5, 10.5, 15.5], dtype=np.float32)
    np.testing.assert_array_equal(bits.BitsInfo(8).quantize(arr), [0, 0, 11, 117, 128, 138, 244, 255, 255])
    np.testing.assert_array_equal(bits.BitsInfo(2).quantize(arr), [0, 0, 0, 1, 2, 2, 3, 3, 3])


def test_params_match_oracle():
    f = O.chan_freqs(1024, -0.390625, 1500.0)
    np.testing.assert_array_equal(params.compute_dmdelays(f, 300, 64e-6, 1500.0),
                                  O.compute_dmdelays(f, 300, 64e-6, 1500.0))
    t = params.compute_dmdelays(f, np.arange(10), 64e-6, 1500.0)
    assert len(params.table_digest(t)) == 40


def test_good_size_and_widths():
    for n in (1, 7, 1000, 509454, 1018893):
        assert good_size_real(n) == O.good_size_real(n)
    np.testing.assert_array_equal(MatchedFilter.get_box_width_spacing(256, 2.0), O.box_width_spacing(256, 2.0))


def test_templates():
    # tests/test_filters.py:95-155
    t = Template.gen_boxcar(5)
    assert t.data.size == 5 and t.kind == "boxcar" and t.get_on_pulse(10, 20) == (10, 15)
    g = Template.gen_gaussian(5.0, 4.0)
    assert g.data.size == int(np.ceil(4.0 * 5.0 / 2.355) * 2 + 1)
    assert g.get_on_pulse(10, 20) == (5, 15)
    lo = Template.gen_lorentzian(5.0, 4.0)
    assert lo.data.size == int(np.ceil(4.0 * 5.0 / 2) * 2 + 1)
    with pytest.raises(ValueError):
        Template.gen_boxcar(-1)
    with pytest.raises(ValueError):
        Template(np.zeros(10), 5, 20)


@pytest.mark.parametrize("key", ["1bit", "2bit", "4bit", "8bit"])
@pytest.mark.parametrize("value", [0, 1, 5, 300.7])
def test_mask_tables_equal_unpack_mask_pack(filfiles, key, value, rng):
    """(b & keep) | set on the file's bytes equals kernels.mask_channels on the unpacked
    samples followed by the reference's pack (base.py:530-583)."""
    from sigpyproc3_amd import FilReader

    path = filfiles[key][0] if isinstance(filfiles[key], list) else filfiles[key]
    fil = FilReader(path)
    hdr = fil.header
    h, raw = O.read_sigproc(path)
    mask = rng.random(hdr.nchans) < 0.3
    v = np.float32(value).astype(hdr.dtype)
    keep, setb = fil._mask_tables(mask, v)
    rows = raw[: 64 * fil.samp_stride].reshape(64, fil.samp_stride)
    got = (rows & keep[None, :]) | setb[None, :]
    x = O.unpacked_stream(h, raw)[: 64 * hdr.nchans]
    masked = O.mask_channels(x, hdr.nchans, mask, v)
    if hdr.nbits < 8:
        order = O.DEFAULT_BITORDER[hdr.nbits]
        expect = O.pack(masked, hdr.nbits, order)
    else:
        expect = masked.astype(np.uint8)
    np.testing.assert_array_equal(got.reshape(-1), expect)


@pytest.mark.parametrize("skipback", [0, 300])
def test_read_plan_packed_into_alternates_buffers(filfiles, skipback):
    """Same schedule and bytes as read_plan_packed, gulp k read into buffer k % 2."""
    from sigpyproc3_amd import FilReader

    fil = FilReader(filfiles["4bit"])
    gulp = 1000
    ref = [(n, ii, raw.copy()) for n, ii, raw in fil.read_plan_packed(gulp=gulp, skipback=skipback)]
    bufs = [bytearray(gulp * fil.samp_stride), bytearray(gulp * fil.samp_stride)]
    got = []
    for k, (n, ii, i, nb) in enumerate(fil.read_plan_packed_into(bufs, gulp=gulp, skipback=skipback)):
        assert i == k % 2
        got.append((n, ii, np.frombuffer(bufs[i], dtype=np.uint8)[:nb].copy()))
    assert len(got) == len(ref)
    for (n1, i1, r1), (n2, i2, r2) in zip(ref, got):
        assert (n1, i1) == (n2, i2)
        np.testing.assert_array_equal(r1, r2)


def test_serve_admission_is_fifo_and_bounded():
    """The server's admission policy: arrival order, no overtaking, never over budget."""
    import collections
    import types

    from sigpyproc3_amd.serve import Arena, admit_in_order

    a = Arena(100 * 256)
    q = collections.deque(types.SimpleNamespace(id=i, need=n * 256) for i, n in enumerate([40, 50, 30, 5, 90]))
    first = admit_in_order(q, a)
    assert [r.id for r in first] == [0, 1] and a.used == 90 * 256
    assert admit_in_order(q, a) == []          # 30 does not fit; 5 behind it may not overtake
    a.free(first[0].offset, 40 * 256)
    got = admit_in_order(q, a)
    assert [r.id for r in got] == [2, 3] and a.used == 85 * 256
    for r in [first[1]] + got:
        a.free(r.offset, r.need)
    assert a.used == 0 and a.free_list == [[0, 100 * 256]]   # freed ranges coalesce
    (last,) = admit_in_order(q, a)
    assert last.id == 4
    a.free(last.offset, last.need)
    with pytest.raises(RuntimeError):
        a.free(last.offset, 256)               # already free


def test_serve_admission_one_gulp_per_stream():
    import collections
    import types

    from sigpyproc3_amd.serve import Arena, admit_in_order

    a = Arena(1000 * 256)
    q = collections.deque(types.SimpleNamespace(id=i, need=10 * 256, handle=h)
                          for i, h in enumerate(["s0", "s1", "s0", "s2", "s1"]))
    busy = set()
    got = admit_in_order(q, a, busy)
    assert [r.id for r in got] == [0, 1, 3] and busy == {"s0", "s1", "s2"}
    assert [r.id for r in q] == [2, 4]          # their streams' earlier gulps are on the device
    busy.discard("s0")
    assert [r.id for r in admit_in_order(q, a, busy)] == [2]
    assert [r.id for r in q] == [4]


def test_serve_arena_fragmentation():
    from sigpyproc3_amd.serve import Arena

    a = Arena(10 * 256)
    offs = [a.alloc(256) for _ in range(10)]
    assert offs == [256 * i for i in range(10)] and a.alloc(1) is None
    for o in offs[::2]:
        a.free(o, 256)
    assert a.used == 5 * 256 and a.alloc(512) is None   # free space is fragmented
    a.free(offs[1], 256)
    assert a.alloc(512) == 0                              # 0 and 1 coalesced (2 is free too)
    assert a.peak == 10 * 256
