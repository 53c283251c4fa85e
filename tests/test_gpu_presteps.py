"""GPU parity of the gulp pre-steps and the fold through the C ABI: collapse
(kernels.extract_tim), bandpass (kernels.extract_bpass / extract_bpass_par) and fold
(kernels.fold), against brute-force float32/float64 loops in the reference's order.

The integer paths sum exactly and fall back to the reference's float32 order when a running
sum could leave the exact-integer range (|sum| > 2^24) or starts at a non-integer; the cases
below drive every lane-unit width (16/8/4/1-byte loads, misaligned starts), ragged channel
counts and both sides of that fallback.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu


def _dev(a: np.ndarray, offset: int = 0):
    """Device copy of ``a`` starting ``offset`` bytes into a fresh allocation (misaligned loads)."""
    from sigpyproc3_amd import _lib

    raw = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    buf = _lib.zeros((raw.size + offset + 16,), "uint8")
    buf[offset : offset + raw.size].copy_(_lib.torch().from_numpy(raw))
    return buf, buf[offset:]


def _packed(x: np.ndarray, nbits: int, order: str) -> np.ndarray:
    return x.reshape(-1).astype(np.uint8) if nbits == 8 else O.pack(x.reshape(-1), nbits, order)


def _seq_f32(out0: np.ndarray, rows: np.ndarray) -> np.ndarray:
    """float32 adds in time order per channel (kernels.extract_bpass)."""
    acc = out0.astype(np.float32).copy()
    for r in rows.astype(np.float32):
        acc = (acc + r).astype(np.float32)
    return acc


@pytest.mark.parametrize("nbits", [1, 2, 4, 8])
@pytest.mark.parametrize("order", ["big", "little"])
def test_collapse_widths_alignment(nbits, order, rng):
    from sigpyproc3_amd import _lib

    _lib.require_gpu()
    st = _lib.stream_handle()
    for nchans, nsamps, offset in ((4096, 777, 0), (1032, 513, 4), (24, 300, 1), (8 * 8 // nbits, 5, 3)):
        x = rng.integers(0, 1 << nbits, (nsamps, nchans), dtype=np.uint8)
        _keep, src = _dev(_packed(x, nbits, order), offset)
        out = _lib.zeros((nsamps + 10,), "float32")
        _lib.call("spp_collapse", _lib.ptr(src), nbits, 1 if order == "big" else 0, nchans, nsamps,
                  _lib.ptr(out), 7, st)
        got = out.cpu().numpy()
        np.testing.assert_array_equal(got[7 : 7 + nsamps], x.astype(np.int64).sum(axis=1).astype(np.float32))
        assert not got[:7].any() and not got[7 + nsamps :].any()


def test_collapse_float32_channel_order(rng):
    from sigpyproc3_amd import _lib

    x = rng.normal(0, 1e3, (300, 77)).astype(np.float32)
    x[:, 5] = 1e30                           # a float64 sum in channel order, rounded once
    x[:, 6] = -1e30
    _keep, src = _dev(x)
    out = _lib.zeros((300,), "float32")
    _lib.call("spp_collapse", _lib.ptr(src), 32, 0, 77, 300, _lib.ptr(out), 0, _lib.stream_handle())
    exp = np.zeros(300, dtype=np.float64)
    for c in range(77):
        exp = exp + x[:, c].astype(np.float64)
    np.testing.assert_array_equal(out.cpu().numpy(), exp.astype(np.float32))


def _bandpass(src, nbits, big, nchans, nsamps, out_init, parallel):
    from sigpyproc3_amd import _lib

    need = _lib.workspace_size("spp_bandpass_workspace_size", nbits, nchans, nsamps)
    ws = _lib.workspace(need)
    out = _lib.to_device(out_init.astype(np.float32))
    _lib.call("spp_bandpass", _lib.ptr(src), nbits, big, nchans, nsamps, _lib.ptr(out), parallel,
              _lib.ptr(ws), ws.numel(), _lib.stream_handle())
    return out.cpu().numpy()


@pytest.mark.parametrize("nbits", [1, 2, 4, 8])
def test_bandpass_unit_widths_and_start_values(nbits, rng):
    """Every load width (aligned 16/8/4 bytes, misaligned bytes), ragged channel counts, and
    start values that keep the exact path (integers), cross 2^24 inside the gulp (exact prefix,
    then the float32 order), lie beyond 2^24, or are not integers (float32 order throughout)."""
    from sigpyproc3_amd import _lib

    _lib.require_gpu()
    for order in ("big", "little"):
        big = 1 if order == "big" else 0
        for nchans, nsamps, offset in ((4096, 3000, 0), (1032, 1777, 8), (200, 2500, 4), (24, 999, 1)):
            if (nchans * nbits) % 8:
                continue
            x = rng.integers(0, 1 << nbits, (nsamps, nchans), dtype=np.uint8)
            _keep, src = _dev(_packed(x, nbits, order), offset)
            maxsum = float(((1 << nbits) - 1) * nsamps)
            starts = {
                "zero": np.zeros(nchans),
                "ints": rng.integers(-1000, 1000, nchans).astype(np.float64),
                "cross": np.full(nchans, 2.0**24 - maxsum / 3.0).round(),
                "beyond": np.full(nchans, 2.0**24 + 2.0),
                "frac": rng.integers(0, 1000, nchans) + 0.5,
            }
            for name, o in starts.items():
                o32 = o.astype(np.float32)
                got = _bandpass(src, nbits, big, nchans, nsamps, o32, 0)
                np.testing.assert_array_equal(got, _seq_f32(o32, x), err_msg=f"seq {order} {nchans} {name}")
                got = _bandpass(src, nbits, big, nchans, nsamps, o32, 1)
                exp = (o32.astype(np.float64) + x.astype(np.float64).sum(axis=0)).astype(np.float32)
                np.testing.assert_array_equal(got, exp, err_msg=f"par {order} {nchans} {name}")


def test_bandpass_full_gulp_exact_and_parallel_single_rounding(rng):
    """2-bit, 4096 channels, 2^16 samples: the exact integer sums; and the 8-bit parallel
    form rounding float64(out) + gulp sum once (2^24 + 2, not 2^24)."""
    from sigpyproc3_amd import _lib

    _lib.require_gpu()
    x = rng.integers(0, 4, (1 << 16, 4096), dtype=np.uint8)
    _keep, src = _dev(O.pack(x.reshape(-1), 2, "big"))
    exact = x.astype(np.int64).sum(axis=0).astype(np.float32)
    np.testing.assert_array_equal(_bandpass(src, 2, 1, 4096, 1 << 16, np.zeros(4096), 0), exact)
    np.testing.assert_array_equal(_bandpass(src, 2, 1, 4096, 1 << 16, np.zeros(4096), 1), exact)
    g = np.zeros(65800, dtype=np.uint8)
    g[:65793] = 255
    g[65793] = 2                              # gulp sum 2^24 + 1
    _keep, src = _dev(g)
    got = _bandpass(src, 8, 0, 1, 65800, np.ones(1), 1)
    assert got[0] == np.float32(2**24 + 2)


def _fold_ref(x, delays, samples, key_start, key_pos, band_start, nbins, fold0, count0):
    """Each (key, band) cell continues its float32 sum over its samples in time order and the
    band's channels in channel order (kernels.fold's order for one cell)."""
    fold = fold0.astype(np.float32).copy()
    count = count0.copy()
    nsamps = x.shape[0]
    for k in range(len(key_pos)):
        ss = samples[key_start[k] : key_start[k + 1]]
        for b in range(len(band_start) - 1):
            c = np.arange(band_start[b], band_start[b + 1])
            rows = ss[:, None] + delays[c][None, :]
            rows = np.where(rows < 0, rows + nsamps, rows)
            vals = x[rows, c[None, :]].astype(np.float32).reshape(-1)
            pos = key_pos[k] + b * nbins
            seq = np.concatenate([[fold[pos]], vals]).astype(np.float32)
            fold[pos] = np.add.accumulate(seq, dtype=np.float32)[-1]
            count[pos] += vals.size
    return fold, count


@pytest.mark.parametrize("nbits", [2, 8])
def test_fold_exact_and_ordered_cells(nbits, rng):
    """Cells whose float32 sums stay exact integers, cells that pass 2^24 (8-bit, one band of
    every channel, thousands of samples per cell), and cells that start at non-integers."""
    from sigpyproc3_amd import _lib

    _lib.require_gpu()
    order = "big"
    nchans, nsamps, nbins = 64, 12000, 4
    x = rng.integers(0, 1 << nbits, (nsamps, nchans), dtype=np.uint8)
    if nbits == 8:
        x[:, :] = rng.integers(200, 256, (nsamps, nchans), dtype=np.uint8)
    delays = rng.integers(-30, 60, nchans).astype(np.int32)
    nvalid = nsamps - int(delays.max())
    for nbands in (1, 5):
        band_start = _band_starts(nchans, nbands)
        key = rng.integers(0, 2 * nbins, nvalid)              # 2 subints x nbins keys
        order_idx = np.argsort(key, kind="stable").astype(np.int32)
        keys, first = np.unique(key[order_idx], return_index=True)
        key_start = np.append(first, nvalid).astype(np.int32)
        key_pos = ((keys // nbins) * nbins * nbands + keys % nbins).astype(np.int32)
        ncell = 2 * nbands * nbins
        fold0 = rng.integers(0, 100, ncell).astype(np.float32)
        fold0[::3] += 0.25                                    # non-integer starts: ordered path
        count0 = rng.integers(0, 10, ncell).astype(np.int32)
        _keep, src = _dev(_packed(x, nbits, order))
        dd = [_lib.to_device(a) for a in (delays, order_idx, key_start, key_pos, band_start)]
        fd, cd = _lib.to_device(fold0), _lib.to_device(count0)
        need = _lib.workspace_size("spp_fold_workspace_size", nbits, nchans, nsamps, nbands)
        ws = _lib.workspace(need)
        _lib.call("spp_fold", _lib.ptr(src), nbits, 1, nchans, nsamps, _lib.ptr(dd[0]), _lib.ptr(dd[1]),
                  nvalid, _lib.ptr(dd[2]), _lib.ptr(dd[3]), len(keys), _lib.ptr(dd[4]), nbands, nbins,
                  _lib.ptr(fd), _lib.ptr(cd), _lib.ptr(ws), ws.numel(), _lib.stream_handle())
        ef, ec = _fold_ref(x, delays, order_idx, key_start, key_pos, band_start, nbins, fold0, count0)
        np.testing.assert_array_equal(fd.cpu().numpy(), ef, err_msg=f"nbands {nbands}")
        np.testing.assert_array_equal(cd.cpu().numpy(), ec)
        if nbits == 8 and nbands == 1:
            assert (ef > 2**24).any()                          # the ordered fallback ran


def _band_starts(nchans: int, nbands: int) -> np.ndarray:
    from sigpyproc3_amd.foldedcube import band_starts

    return band_starts(nchans, nbands)


def _zerodm(src, nbits, big, nchans, nsamps, bpass, out_offset=0):
    from sigpyproc3_amd import _lib

    w_d = _lib.to_device((bpass / bpass.sum()).astype(np.float32))
    b_d = _lib.to_device(bpass.astype(np.float32))
    esz = 4 if nbits == 32 else 1
    out = _lib.zeros((nsamps * nchans * esz + out_offset,), "uint8")
    view = out[out_offset:]
    _lib.call("spp_remove_zerodm", _lib.ptr(src), nbits, big, nchans, nsamps, _lib.ptr(w_d), _lib.ptr(b_d),
              _lib.ptr(view), _lib.stream_handle())
    res = view.cpu().numpy()
    return res.view(np.float32) if nbits == 32 else res


@pytest.mark.parametrize("nbits", [1, 2, 4, 8])
@pytest.mark.parametrize("order", ["big", "little"])
def test_remove_zerodm_register_path(nbits, order, rng):
    """The dword path (weights in registers, 1-8 dwords per thread) against the exact fused
    oracle on small rows, and against the per-row byte kernel (misaligned output) at widths
    up to 8192 channels."""
    big = 1 if order == "big" else 0
    nchans, nsamps = 64 * 8 // nbits, 40
    x = rng.integers(0, 1 << nbits, (nsamps, nchans), dtype=np.uint8)
    bpass = rng.uniform(1.0e3, 5.0e4, nchans).astype(np.float32)
    _keep, src = _dev(_packed(x, nbits, order))
    exp = O.remove_zerodm(x.reshape(-1), nchans, bpass, np.uint8, fused=True)
    np.testing.assert_array_equal(_zerodm(src, nbits, big, nchans, nsamps, bpass), exp)
    for nchans in (1024, 4096, 8192 if nbits == 8 else 2048, 1040):
        nsamps = 50
        x = rng.integers(0, 1 << nbits, (nsamps, nchans), dtype=np.uint8)
        bpass = rng.uniform(1.0e3, 5.0e4, nchans).astype(np.float32)
        _keep, src = _dev(_packed(x, nbits, order))
        fast = _zerodm(src, nbits, big, nchans, nsamps, bpass)
        slow = _zerodm(src, nbits, big, nchans, nsamps, bpass, out_offset=1)  # byte kernel
        np.testing.assert_array_equal(fast, slow, err_msg=f"{nchans}")
        dword = _zerodm(src, nbits, big, nchans, nsamps, bpass, out_offset=4)  # dword stores only
        np.testing.assert_array_equal(dword, slow, err_msg=f"{nchans} dword stores")
        # z is exact and y = x - z*w + b lands within a rounding of the unfused form
        ref = O.remove_zerodm(x.reshape(-1), nchans, bpass, np.uint8, fused=False)
        assert np.mean(fast == ref) > 0.99


@pytest.mark.parametrize("row_bytes, offset", [(4096, 0), (1028, 4), (77, 0), (512, 3)])
def test_mask_bytes_unit_widths(row_bytes, offset, rng):
    from sigpyproc3_amd import _lib

    nrows = 333
    data = rng.integers(0, 256, nrows * row_bytes, dtype=np.uint8)
    keep = rng.integers(0, 256, row_bytes, dtype=np.uint8)
    setb = rng.integers(0, 256, row_bytes, dtype=np.uint8) & ~keep
    _kd, dd = _dev(data, offset)
    _kk, kd = _dev(keep, offset)
    _ks, sd = _dev(setb, offset)
    _lib.call("spp_mask_bytes", _lib.ptr(dd), data.size, row_bytes, _lib.ptr(kd), _lib.ptr(sd),
              _lib.stream_handle())
    got = dd[: data.size].cpu().numpy()
    exp = ((data.reshape(nrows, row_bytes) & keep) | setb).reshape(-1)
    np.testing.assert_array_equal(got, exp)


def test_remove_zerodm_large_weights_clamped(rng):
    """Weights large enough that y leaves the int32 range: the clamped form (the low 8 bits of
    the saturated value, as the per-row byte kernel computes) is kept for those threads."""
    nchans, nsamps = 256, 20
    x = rng.integers(0, 256, (nsamps, nchans), dtype=np.uint8)
    bpass = rng.uniform(1.0e3, 5.0e4, nchans).astype(np.float32)
    bpass[::7] = 3.0e9                        # b alone exceeds 2^31
    _keep, src = _dev(x)
    fast = _zerodm(src, 8, 0, nchans, nsamps, bpass)
    slow = _zerodm(src, 8, 0, nchans, nsamps, bpass, out_offset=1)
    np.testing.assert_array_equal(fast, slow)


@pytest.mark.parametrize("ncols, out_cols, aligned", [(16384, 16384, True), (1000, 1000, True), (1001, 1001, True),
                                                       (5000, 3000, True), (4096, 4096, False)])
def test_roll_rows_vector_and_scalar(ncols, out_cols, aligned, rng):
    """Rows rolled by arbitrary offsets (wrapping inside a 16-byte group included) through the
    float4 path and the scalar path, against numpy."""
    from sigpyproc3_amd import _lib

    nrows = 37
    x = rng.normal(size=(nrows, ncols)).astype(np.float32)
    if out_cols == ncols:
        off = rng.integers(0, ncols, nrows).astype(np.int64)
        off[:4] = [0, 1, ncols - 1, ncols - 3]
    else:
        off = rng.integers(0, ncols - out_cols + 1, nrows).astype(np.int64)
    _kx, xd = _dev(x)
    od = _lib.to_device(off)
    buf = _lib.zeros((nrows * out_cols + 4,), "float32")
    out = buf if aligned else buf[1:]
    _lib.call("spp_roll_rows", _lib.ptr(xd), nrows, ncols, _lib.ptr(off), _lib.ptr(od), out_cols, _lib.ptr(out),
              _lib.stream_handle())
    got = out[: nrows * out_cols].cpu().numpy().reshape(nrows, out_cols)
    exp = np.stack([np.roll(x[r], -off[r])[:out_cols] for r in range(nrows)])
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("ncols, out_cols, ndm, kind", [(4096, 4096, 37, "smooth"), (3000, 3000, 20, "random"),
                                                         (5000, 2000, 13, "smooth"), (300, 300, 5, "random")])
def test_dmt_tiles_vs_sequential_sum(ncols, out_cols, ndm, kind, rng):
    """The tiled DM-time transform (DM rows grouped by offset spread, channel windows in LDS)
    against a float32 sum in channel order: smooth offsets (groups of up to 16 rows, ragged
    last group) and random offsets (one row per block), circular and valid-columns forms."""
    from sigpyproc3_amd import _lib

    nchans = 67
    x = rng.normal(size=(nchans, ncols)).astype(np.float32)
    lim = ncols - out_cols + 1 if out_cols < ncols else ncols
    if kind == "smooth":
        base = np.linspace(0, min(lim - 1, 3 * ndm), ndm)[:, None] * np.linspace(0, 1, nchans)[None, :]
        off = np.floor(base).astype(np.int64) % lim
    else:
        off = rng.integers(0, lim, (ndm, nchans)).astype(np.int64)
    _kx, xd = _dev(x)
    od = _lib.to_device(off.reshape(-1))
    out = _lib.zeros((ndm * out_cols,), "float32")
    offh = np.ascontiguousarray(off.reshape(-1))
    _lib.call("spp_dmt", _lib.ptr(xd), nchans, ncols, _lib.ptr(offh), _lib.ptr(od), ndm, out_cols, _lib.ptr(out),
              _lib.stream_handle())
    got = out.cpu().numpy().reshape(ndm, out_cols)
    t = np.arange(out_cols)
    exp = np.zeros((ndm, out_cols), dtype=np.float32)
    for c in range(nchans):
        exp = (exp + x[c][(t[None, :] + off[:, c : c + 1]) % ncols]).astype(np.float32)
    np.testing.assert_array_equal(got, exp)
