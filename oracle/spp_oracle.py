"""numpy restatement of the reference filterbank hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
import this module. It is the checker; the product (``sigpyproc3_amd``) never
imports it.

Every function cites the reference file:line it restates (paths relative to the
sigpyproc source tree ``src/sigpyproc``). Pinning (what each function is checked
against, beyond its own formula) -- see ``tests/test_oracle.py``:

=========================  ==========================================================
function                   pinned by
=========================  ==========================================================
unpack / pack              reference KAT on [7, 23] for all six variants
                           (tests/test_kernels.py:12-41); np.unpackbits for 1-bit
                           (tests/test_bits.py:14-36); pack/unpack round trip
chan_freqs, ref_frequency  header.py:197-256 definitions; fixture headers
compute_dmdelays           notebook KAT (tutorial.fil, DM 30: max delay 17);
                           delays[:,0]==0 at ref=fch1 (tests/test_params.py:9-37);
                           max delay in seconds +-0.1 s (tests/test_header.py:35-41);
                           float64 closed form differs by at most 1 sample
read_plan_blocks           closed-form tiling invariants (outputs tile [0, N-md))
dedisperse_kernel /        brute-force per-sample loop on tiny inputs; gulp
filterbank_dedisperse      independence for non-negative delays; notebook KAT
                           (sum 19636992, max 121, min 88, median 105)
subband_kernel             nsub == 1 reduces to dedisperse; brute force
sweep_plane                row i == filterbank_dedisperse(delays[i])
good_size_real             scipy.fft.next_fast_len(n, real=True) (pocketfft)
convolve_templates_fft     explicit DFT sums (no FFT) of the padded series and template
                           with irfft_default_brute, at odd and even good sizes; raises
                           for n == Ng odd (n = 243, 625, ...)
irfft_default_brute        numpy irfft's default-length output on random spectra,
                           including nonzero imaginary parts in the first/last bins
circular_to_reference      convolve_templates_fft (fp64) at odd and even good sizes
boxcar_conv_closed,        convolve_templates_fft (fp64); brute-force circular sums
template_conv_closed       (boxcar_conv_circular / template_conv_circular)
estimate_zscore            textbook median / IQR via np.percentile (stats.py:480-502)
running_filter_mean        brute-force window mean over the symmetric-padded series
running_normalise          PARITY UNPINNED by the reference: cfg "running mean/std"
                           normalisation has no reference function; it is defined
                           here and pinned only by brute force on small inputs
=========================  ==========================================================
"""

from __future__ import annotations

import math
import struct

import numpy as np

DM_CONSTANT_LK = 4.148808e3  # params.py:17

# ---------------------------------------------------------------------------
# bits: unpack / pack (core/kernels.py:20-173, io/bits.py:19-70, 175-182)
# ---------------------------------------------------------------------------

DEFAULT_BITORDER = {1: "little", 2: "big", 4: "big"}  # io/bits.py:175-182


def unpack(array: np.ndarray, nbits: int, bitorder: str = "big") -> np.ndarray:
    """kernels.unpack{1,2,4}_8_{big,little} (kernels.py:20-69).

    big: the first sample of a byte sits in the most significant bits.
    """
    array = np.asarray(array, dtype=np.uint8)
    if nbits not in (1, 2, 4):
        raise ValueError(nbits)
    nper = 8 // nbits
    mask = (1 << nbits) - 1
    out = np.empty((array.size, nper), dtype=np.uint8)
    for k in range(nper):
        shift = (8 - (k + 1) * nbits) if bitorder[0] == "b" else k * nbits
        out[:, k] = (array >> shift) & mask
    return out.reshape(-1)


def pack(array: np.ndarray, nbits: int, bitorder: str = "big") -> np.ndarray:
    """kernels.pack{1,2,4}_8_{big,little} (kernels.py:72-173): inverse of unpack. Values are
    shifted and OR-ed without masking, so out-of-range samples spill into higher fields."""
    array = np.asarray(array, dtype=np.uint8)
    nper = 8 // nbits
    a = array.reshape(-1, nper).astype(np.uint32)
    out = np.zeros(a.shape[0], dtype=np.uint32)
    for k in range(nper):
        shift = (8 - (k + 1) * nbits) if bitorder[0] == "b" else k * nbits
        out |= a[:, k] << shift
    return out.astype(np.uint8)


# ---------------------------------------------------------------------------
# header / params: channel frequencies and delay table
# ---------------------------------------------------------------------------


def chan_freqs(nchans: int, foff: float, fch1: float) -> np.ndarray:
    """Header.chan_freqs (header.py:228-237): float32 arange * foff + fch1."""
    return np.arange(nchans, dtype=np.float32) * foff + fch1


def ref_frequency(ref: str | float, fch1: float, foff: float, nchans: int) -> float:
    """Reference frequency resolution of Header.get_dmdelays (header.py:369-379).

    ftop = fch1 - foff/2 (197-206); fcenter = ftop + foff*nchans/2 (219-226);
    fmax/fmin = max/min of the float32 chan_freqs (239-256).
    """
    if isinstance(ref, str):
        if ref == "fch1":
            return float(fch1)
        if ref == "fmax":
            return float(chan_freqs(nchans, foff, fch1).max())
        if ref == "fmin":
            return float(chan_freqs(nchans, foff, fch1).min())
        if ref == "fcenter":
            ftop = fch1 - 0.5 * foff
            return float(ftop + 0.5 * foff * nchans)
        raise ValueError(ref)
    return float(ref)


def compute_dmdelays(freqs, dm, tsamp: float, ref_freq: float, *, in_samples=True):
    """params.compute_dmdelays (params.py:26-62), numpy-2 (NEP 50) float32 semantics.

    The arithmetic is float32 throughout: Python-float constants are weak scalars.
    The rounding is numpy's round-half-to-even on float32, then int32.
    """
    freqs = np.atleast_1d(freqs).astype(np.float32)
    dm = np.atleast_1d(dm)[:, np.newaxis].astype(np.float32)
    delays = dm * DM_CONSTANT_LK * ((freqs**-2) - (ref_freq**-2))
    if in_samples:
        delays = (delays / tsamp).round().astype(np.int32)
    return delays.squeeze()


def compute_dmdelays_f64(freqs, dm, tsamp: float, ref_freq: float) -> np.ndarray:
    """Float64 closed form of the same delay (used only to bound the float32 table)."""
    f = np.atleast_1d(freqs).astype(np.float64)
    d = np.atleast_1d(dm).astype(np.float64)[:, None]
    return np.rint(d * DM_CONSTANT_LK * (f**-2 - ref_freq**-2) / tsamp).squeeze()


# ---------------------------------------------------------------------------
# read_plan schedule (readers.py:171-233)
# ---------------------------------------------------------------------------


def read_plan_blocks(nsamps: int, gulp: int, skipback: int):
    """FilReader.read_plan block list (readers.py:184-216).

    Returns a list of ``(ii, nsamps_read, skip)`` exactly as the reference iterates.
    """
    gulp = min(nsamps, gulp)
    skipback = abs(skipback)
    if skipback >= gulp:
        raise ValueError(f"readsamps ({gulp}) must be > skipback ({skipback})")
    nreads, lastread = divmod(nsamps, (gulp - skipback))
    if lastread < skipback:
        nreads -= 1
        lastread = nsamps - (nreads * (gulp - skipback))
    blocks = [(ii, gulp, -skipback) for ii in range(nreads)]
    if lastread != 0:
        blocks.append((nreads, lastread, 0))
    return blocks


# ---------------------------------------------------------------------------
# dedisperse (kernels.py:340-362, base.py:360-424)
# ---------------------------------------------------------------------------


def dedisperse_kernel(in_arr, out_arr, delays, nchans: int, nsamps: int, out_start: int):
    """kernels.dedisperse (kernels.py:348-362) on one gulp.

    ``out[out_start + t] = f32(sum_c f64(in[t + d_c, c]))`` for ``t < nsamps - max|d|``.
    numba wraps a negative row index (``t + d_c < 0``) to ``t + d_c + nsamps``.
    """
    delays = np.asarray(delays, dtype=np.int64)
    max_delay = int(np.max(np.abs(delays)))
    nout = nsamps - max_delay
    if nout <= 0:
        return
    in2d = np.asarray(in_arr)[: nsamps * nchans].reshape(nsamps, nchans)
    chans = np.arange(nchans)
    step = max(1, (1 << 22) // max(nchans, 1))
    for a in range(0, nout, step):
        b = min(nout, a + step)
        rows = np.arange(a, b)[:, None] + delays[None, :]
        rows = np.where(rows < 0, rows + nsamps, rows)
        vals = in2d[rows, chans[None, :]].astype(np.float64)
        out_arr[out_start + a : out_start + b] = vals.sum(axis=1).astype(np.float32)


def filterbank_dedisperse(stream, nsamples: int, nchans: int, delays, gulp=16384,
                          start=0, nsamps=None):
    """Filterbank.dedisperse (base.py:360-424) over an unpacked [N][nchans] stream.

    ``stream`` is the whole unpacked data stream (what read_plan reads, file order).
    Returns the float32 time series of length ``nsamples - max|d|``.
    """
    delays = np.asarray(delays, dtype=np.int32)
    max_delay = int(np.max(np.abs(delays)))
    gulp = max(2 * max_delay, gulp)
    tim_len = nsamples - max_delay
    tim_ar = np.zeros(tim_len, dtype=np.float32)
    if nsamps is None:
        nsamps = nsamples - start
    stream = np.asarray(stream).reshape(-1, nchans)
    pos = start
    for ii, nsamps_r, skip in read_plan_blocks(nsamps, gulp, max_delay):
        data = stream[pos : pos + nsamps_r].reshape(-1)
        if data.size != nsamps_r * nchans:
            raise ValueError("short read")
        dedisperse_kernel(data, tim_ar, delays, nchans, nsamps_r, ii * (gulp - max_delay))
        pos += nsamps_r + skip
    return tim_ar


def dedisperse_brute(stream, nchans: int, delays):
    """Direct ``out[t] = sum_c x[t + d_c][c]`` for non-negative delays (F7 closed form)."""
    x = np.asarray(stream).reshape(-1, nchans)
    d = np.asarray(delays, dtype=np.int64)
    n = x.shape[0] - int(np.max(np.abs(d)))
    out = np.zeros(max(n, 0), dtype=np.float64)
    for c in range(nchans):
        out += x[d[c] : d[c] + n, c]
    return out.astype(np.float32)


def sweep_plane(stream, nchans: int, table):
    """DM sweep: row i is ``Filterbank.dedisperse`` at DM i (length N - max|d_i|).

    Non-negative delays only (the gulp schedule is then irrelevant, F7).
    Returns (list of float32 rows).
    """
    table = np.atleast_2d(np.asarray(table, dtype=np.int64))
    if (table < 0).any():
        raise ValueError("sweep needs non-negative delays")
    x = np.asarray(stream).reshape(-1, nchans)
    return [dedisperse_brute(x, nchans, row) for row in table]


# ---------------------------------------------------------------------------
# subband (kernels.py:379-401, base.py:1030-1107)
# ---------------------------------------------------------------------------


def subband_kernel(inarray, outarray, delays, chan_to_sub, maxdelay, nchans, nsubs, nsamps):
    """kernels.subband (kernels.py:387-401), with ``outarray`` zeroed by the caller.

    ``out[t*nsubs + chan_to_sub[c]] += in[nchans*(t + d_c) + c]`` for t < nsamps-maxdelay.
    A negative flat index wraps (numba) by the array size.
    """
    n = nsamps - maxdelay
    if n <= 0:
        return
    inarr = np.asarray(inarray)
    size = inarr.size
    d = np.asarray(delays, dtype=np.int64)
    cts = np.asarray(chan_to_sub, dtype=np.int64)
    t = np.arange(n, dtype=np.int64)
    for c in range(nchans):
        idx = nchans * (t + d[c]) + c
        idx = np.where(idx < 0, idx + size, idx)
        oidx = nsubs * t + cts[c]
        ok = oidx < outarray.size
        np.add.at(outarray, oidx[ok], inarr[idx[ok]].astype(np.float32))


def filterbank_subband(stream, nsamples, nchans, delays, nsub, gulp=16384, start=0,
                       nsamps=None):
    """Filterbank.subband data (base.py:1071-1107): the float32 samples written.

    The reference accumulates into ``np.empty`` (never zeroed); here the output
    buffer is zeroed per gulp, which is the kernel semantics the reference comment
    asks for (base.py:1073).
    """
    delays = np.asarray(delays, dtype=np.int32)
    subfactor = nchans // nsub
    max_delay = int(np.max(np.abs(delays)))
    gulp = max(2 * max_delay, gulp)
    chan_to_sub = np.arange(nchans, dtype=np.int32) // subfactor
    if nsamps is None:
        nsamps = nsamples - start
    stream = np.asarray(stream).reshape(-1, nchans)
    pos = start
    pieces = []
    for _ii, nsamps_r, skip in read_plan_blocks(nsamps, gulp, max_delay):
        data = stream[pos : pos + nsamps_r].reshape(-1)
        out_ar = np.zeros((gulp - max_delay) * nsub, dtype=np.float32)
        subband_kernel(data, out_ar, delays, chan_to_sub, max_delay, nchans, nsub, nsamps_r)
        pieces.append(out_ar[: (nsamps_r - max_delay) * nsub].copy())
        pos += nsamps_r + skip
    return np.concatenate(pieces)


# ---------------------------------------------------------------------------
# gulp pre-steps: collapse, bandpass, channel mask, zero-DM (kernels.py:255-345, 465-500)
# ---------------------------------------------------------------------------


def collapse(stream, nsamples: int, nchans: int, gulp=16384, start=0, nsamps=None, method="sum"):
    """Filterbank.collapse (base.py:256-306): kernels.extract_tim (kernels.py:255-274) per
    gulp, written at ``ii * nsamps_r`` as the reference does, then ``/= nchans`` for mean."""
    tim_len = (nsamples - start) if nsamps is None else nsamps
    tim = np.zeros(tim_len, dtype=np.float32)
    x = np.asarray(stream).reshape(-1, nchans)
    pos = start
    for ii, nsamps_r, skip in read_plan_blocks(tim_len, gulp, 0):
        rows = x[pos : pos + nsamps_r].astype(np.float64)
        tim[ii * nsamps_r : ii * nsamps_r + nsamps_r] = rows.sum(axis=1).astype(np.float32)
        pos += nsamps_r + skip
    if method == "mean":
        tim /= nchans
    return tim


def bandpass(stream, nsamples: int, nchans: int, gulp=16384, start=0, nsamps=None,
             parallel=False, method="sum"):
    """Filterbank.bandpass (base.py:308-358).

    parallel=False: kernels.extract_bpass (kernels.py:277-291), ``out[c] += x[t][c]`` in
    float32 in time order. parallel=True: kernels.extract_bpass_par (kernels.py:294-310),
    ``out_arr[ichan] += summ`` with a float64 ``summ``: numba adds float32 + float64 in
    float64 and rounds once on the float32 store, per gulp.
    """
    n = (nsamples - start) if nsamps is None else nsamps
    out = np.zeros(nchans, dtype=np.float32)
    x = np.asarray(stream).reshape(-1, nchans)
    pos = start
    num = 0
    for _ii, nsamps_r, skip in read_plan_blocks(n, gulp, 0):
        rows = x[pos : pos + nsamps_r]
        if parallel:
            out = (out.astype(np.float64) + rows.astype(np.float64).sum(axis=0)).astype(np.float32)
        else:
            for t in range(nsamps_r):
                out += rows[t].astype(np.float32)
        num += nsamps_r
        pos += nsamps_r + skip
    if method == "mean":
        out /= num
    return out


def mask_channels(stream, nchans: int, chan_mask, mask_value):
    """kernels.mask_channels (kernels.py:321-345): masked channels set to ``mask_value``
    (already cast to the data dtype), on a copy of the unpacked stream."""
    x = np.array(stream).reshape(-1, nchans)
    x[:, np.asarray(chan_mask, dtype=bool)[:nchans]] = mask_value
    return x.reshape(-1)


def fma_exact(a: float, b: float, c: float) -> float:
    """a*b + c rounded once to float64 (exact rational arithmetic)."""
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))


def remove_zerodm(stream, nchans: int, bpass, out_dtype, fused=True):
    """kernels.remove_zerodm (kernels.py:469-500) on the unpacked stream.

    ``z = sum_c x[t][c]`` (float64), ``y = x - z*w + b`` with ``w = bpass / bpass.sum()``
    (float32, numpy), evaluated in float64. fused=True forms ``x - z*w`` with one rounding
    (the contraction fastmath allows; exact, slow: small inputs only), fused=False with two.
    uint8 output keeps the low 8 bits of the truncated value (the float64 -> uint8 store on
    x86-64); float32 output rounds once.
    """
    bpass = np.asarray(bpass, dtype=np.float32)
    w = bpass / bpass.sum()
    x = np.asarray(stream).reshape(-1, nchans).astype(np.float64)
    z = x.sum(axis=1)
    w64, b64 = w.astype(np.float64), bpass.astype(np.float64)
    if fused:
        y = np.empty_like(x)
        for t in range(x.shape[0]):
            for c in range(nchans):
                y[t, c] = fma_exact(-z[t], float(w64[c]), float(x[t, c])) + b64[c]
    else:
        y = (x - z[:, None] * w64[None, :]) + b64[None, :]
    if np.dtype(out_dtype) == np.float32:
        return y.astype(np.float32).reshape(-1)
    return (np.trunc(y).astype(np.int64) & 0xFF).astype(np.uint8).reshape(-1)


# ---------------------------------------------------------------------------
# fold (kernels.py:404-445, base.py:1109-1207)
# ---------------------------------------------------------------------------


def fold_kernel(inarray, fold_ar, count_ar, delays, maxdelay, tsamp, period, accel, total_nsamps,
                nsamps, nchans, nbins, nints, nsubs, index):
    """kernels.fold (kernels.py:404-445) in plain Python floats (IEEE double, like numba
    without fastmath). tsamp, period and accel arrive as float32 (the numba signature)."""
    tsamp, period, accel = float(np.float32(tsamp)), float(np.float32(period)), float(np.float32(accel))
    factor1 = total_nsamps / nints
    factor2 = nchans / nsubs
    tobs = total_nsamps * tsamp
    x = np.asarray(inarray)
    chans = np.arange(nchans)
    sub_band = np.array([int(c // factor2) for c in range(nchans)])
    for isamp in range(nsamps - maxdelay):
        tj = (isamp + index) * tsamp
        phase = nbins * tj * (1 + accel * (tj - tobs) / (2 * 299792458.0)) / period + 0.5
        phasebin = abs(int(phase)) % nbins
        subint = (isamp + index) // factor1
        pos1 = (subint * nbins * nsubs) + phasebin
        pos2 = (pos1 + sub_band * nbins).astype(np.int64)
        rows = isamp + np.asarray(delays, dtype=np.int64)
        idx = nchans * rows + chans
        idx = np.where(idx < 0, idx + x.size, idx)
        np.add.at(fold_ar, pos2, x[idx].astype(np.float32))
        np.add.at(count_ar, pos2, 1)


def filterbank_fold(stream, nsamples, nchans, tsamp, delays, period, accel=0.0, nbins=50, nints=32,
                    nbands=32, gulp=16384, start=0, nsamps=None):
    """Filterbank.fold (base.py:1109-1207) over the unpacked stream; returns the
    [nints, nbands, nbins] cube (float32)."""
    nbands = min(nbands, nchans)
    delays = np.asarray(delays, dtype=np.int32)
    max_delay = int(delays.max())
    gulp = max(2 * max_delay, gulp)
    fold_ar = np.zeros(nbins * nints * nbands, dtype=np.float32)
    count_ar = np.zeros(nbins * nints * nbands, dtype=np.int32)
    n = (nsamples - start) if nsamps is None else nsamps
    x = np.asarray(stream).reshape(-1, nchans)
    pos = start
    for ii, nsamps_r, skip in read_plan_blocks(n, gulp, max_delay):
        data = x[pos : pos + nsamps_r].reshape(-1)
        fold_kernel(data, fold_ar, count_ar, delays, max_delay, tsamp, period, accel, nsamples,
                    nsamps_r, nchans, nbins, nints, nbands, ii * (gulp - max_delay))
        pos += nsamps_r + skip
    with np.errstate(divide="ignore", invalid="ignore"):
        fold_ar /= count_ar
    return fold_ar.reshape(nints, nbands, nbins)


# ---------------------------------------------------------------------------
# in-memory blocks (kernels.py:1059-1193, block.py:267-343)
# ---------------------------------------------------------------------------


def roll_block(arr, shifts):
    """kernels.roll_block (kernels.py:1059-1099): row r rolled right by shifts[r], circular."""
    arr = np.asarray(arr)
    return np.stack([np.roll(arr[r], int(shifts[r])) for r in range(arr.shape[0])])


def roll_block_valid(arr, shifts):
    """kernels.roll_block_valid (kernels.py:1102-1152): the columns no row wraps into."""
    arr = np.asarray(arr)
    s = np.asarray(shifts, dtype=np.int64)
    start, end = max(0, int(s.max())), arr.shape[1] + min(0, int(s.min()))
    if end - start <= 0:
        raise ValueError("Not enough samples")
    return np.stack([arr[r, start - s[r] : end - s[r]] for r in range(arr.shape[0])])


def dmt_block(arr, dm_delays, valid=False):
    """kernels.dmt_block[_valid] (kernels.py:1155-1193): per DM, the channel sum (float32, in
    channel order from 0) of the block rolled by +delays. The valid form needs every DM's
    own valid width to equal the one of the whole table (else ValueError)."""
    arr = np.asarray(arr, dtype=np.float32)
    d = np.atleast_2d(np.asarray(dm_delays, dtype=np.int64))
    if valid:
        width = arr.shape[1] + min(0, int(d.min())) - max(0, int(d.max()))
        if width <= 0:
            raise ValueError("Not enough samples")
    rows = []
    for drow in d:
        r = roll_block_valid(arr, drow) if valid else roll_block(arr, drow)
        if valid and r.shape[1] != width:
            raise ValueError("could not broadcast")
        acc = np.zeros(r.shape[1], dtype=np.float32)
        for c in range(r.shape[0]):
            acc = acc + r[c]
        rows.append(acc)
    return np.stack(rows)


# ---------------------------------------------------------------------------
# matched filter (core/filters.py, core/kernels.py:861-1027, core/stats.py)
# ---------------------------------------------------------------------------


def good_size_real(n: int) -> int:
    """pocketfft ``good_size(n, real=True)`` (kernels.py:909-926 via rocket-fft).

    Smallest 2^a 3^b 5^c >= n.
    """
    if n <= 6:
        return n
    best = 2 * n
    f5 = 1
    while f5 < best:
        f35 = f5
        while f35 < best:
            x = f35
            while x < n:
                x *= 2
            best = min(best, x)
            f35 *= 3
        f5 *= 5
    return best


def box_width_spacing(size_max: int, spacing_factor: float = 1.5) -> np.ndarray:
    """MatchedFilter.get_box_width_spacing (filters.py:340-365)."""
    widths = [1]
    while widths[-1] < size_max:
        nxt = int(max(widths[-1] + 1, spacing_factor * widths[-1]))
        if nxt > size_max:
            break
        widths.append(nxt)
    return np.array(widths, dtype=np.float32)


def normalize_template(arr):
    """kernels.normalize_template (kernels.py:1008-1027)."""
    mean = np.mean(arr)
    arr_norm = arr - mean
    norm = np.sqrt(np.sum(arr_norm**2))
    if norm == 0:
        return arr_norm
    return arr_norm / norm


def convolve_templates_fft(data, temps, ref_bins, dtype=np.float64):
    """kernels.convolve_templates (kernels.py:861-906) with numpy FFTs in ``dtype``.

    ``np.fft.irfft`` is called with its default length, as at kernels.py:904: for an odd
    good size Ng that is Ng - 1 samples, and when n == Ng (odd) the row cannot be filled,
    so this raises ValueError as the reference's slice assignment (kernels.py:905) does.
    """
    data = np.asarray(data, dtype=dtype)
    n = data.size
    ng = good_size_real(n)
    data_pad = data[np.arange(ng) % n]
    data_fft = np.fft.rfft(data_pad)
    convs = np.empty((len(temps), n), dtype=dtype)
    for i, (t, rb) in enumerate(zip(temps, ref_bins)):
        tp = np.zeros(ng, dtype=dtype)
        tp[: len(t)] = t
        tp = np.roll(tp, -rb)
        tp = np.roll(tp[::-1], 1)
        tn = normalize_template(tp)
        conv = np.fft.irfft(data_fft * np.fft.rfft(tn))
        if conv.size < n:
            msg = f"cannot assign slice from input of different size ({conv.size} < {n})"
            raise ValueError(msg)
        convs[i] = conv[:n]
    return convs


def circular_to_reference(c_full, n: int) -> np.ndarray:
    """The reference's ``conv[:n]`` from the Ng-circular convolution ``c_full`` (last axis
    of length Ng, all Ng lags).

    kernels.py:904 computes ``irfft(rfft(pad(z)) * rfft(t))`` = ``irfft(rfft(c))`` at
    irfft's default length 2 * (Ng // 2): that is ``c`` itself for an even Ng, and for an odd
    Ng the length-(Ng - 1) inverse of the first (Ng + 1) / 2 bins of ``rfft(c)`` (the
    imaginary parts of the first and last of them dropped). Raises ValueError when those
    Ng - 1 samples cannot fill n outputs (n == Ng odd), as kernels.py:905 does.
    """
    c_full = np.asarray(c_full, dtype=np.float64)
    ng = c_full.shape[-1]
    if ng % 2 == 0:
        return c_full[..., :n]
    y = np.fft.irfft(np.fft.rfft(c_full, axis=-1), axis=-1)
    if y.shape[-1] < n:
        msg = f"cannot assign slice from input of different size ({y.shape[-1]} < {n})"
        raise ValueError(msg)
    return y[..., :n]


def template_conv_circular(z, h, ref_bin: int) -> np.ndarray:
    """Ng-circular convolution of one template, float64, at all Ng lags.

    ``c[t] = (sum_m h_m zt[(t - ref + m) mod Ng] - mu * sum(zt)) / sigma`` with ``zt`` the
    circular pad of ``z`` to Ng (kernels.py:998-1005), ``mu = sum(h)/Ng`` and
    ``sigma = sqrt(sum(h^2) - sum(h)^2/Ng)`` (normalize_template, kernels.py:1008-1027, of
    the padded, rolled and reversed template, kernels.py:896-902).
    """
    z = np.asarray(z, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    n = z.size
    ng = good_size_real(n)
    zt = z[np.arange(ng) % n]
    s = zt.sum()
    mu = h.sum() / ng
    var = (h * h).sum() - h.sum() ** 2 / ng
    t = np.arange(ng)
    acc = np.zeros(ng)
    for m in range(h.size):
        acc += h[m] * zt[(t - ref_bin + m) % ng]
    num = acc - mu * s
    return num / np.sqrt(var) if var > 0 else num


def template_conv_closed(z, h, ref_bin: int) -> np.ndarray:
    """One convolve_templates row (kernels.py:861-906) in float64: the circular closed
    form mapped to the reference's output length (``circular_to_reference``)."""
    z = np.asarray(z)
    return circular_to_reference(template_conv_circular(z, h, ref_bin), z.size)


def boxcar_conv_circular(z, widths) -> np.ndarray:
    """F10 closed form for boxcar templates (ones(w), ref 0), float64, at all Ng lags:
    ``(sum_{j<w} zt[(t + j) mod Ng] - w S / Ng) / sqrt(w (1 - w / Ng))``."""
    z = np.asarray(z, dtype=np.float64)
    n = z.size
    ng = good_size_real(n)
    zt = z[np.arange(ng) % n]
    q = np.concatenate([[0.0], np.cumsum(zt)])
    s = q[-1]
    t = np.arange(ng)
    out = np.empty((len(widths), ng))
    for i, w in enumerate(widths):
        w = int(w)
        end = t + w
        win = np.where(end <= ng, q[np.minimum(end, ng)] - q[t], (s - q[t]) + q[np.maximum(end - ng, 0)])
        out[i] = (win - w * s / ng) / math.sqrt(w * (1.0 - w / ng))
    return out


def boxcar_conv_closed(z, widths) -> np.ndarray:
    """Boxcar convolve_templates rows [len(widths), n] in float64 (the reference's output
    for any n: ``circular_to_reference`` of the closed form)."""
    z = np.asarray(z)
    return circular_to_reference(boxcar_conv_circular(z, widths), z.size)


def irfft_default_brute(spec_half) -> np.ndarray:
    """numpy's irfft at its default length m = 2 (len(spec_half) - 1), written out as a
    real inverse DFT: ``y_t = (Re X_0 + 2 sum_{0<k<m/2} Re(X_k e^{2 pi i k t/m})
    + Re X_{m/2} (-1)^t) / m`` (the imaginary parts of X_0 and X_{m/2} do not enter)."""
    x = np.asarray(spec_half, dtype=np.complex128)
    m = 2 * (x.size - 1)
    t = np.arange(m)
    y = np.full(m, x[0].real)
    for k in range(1, x.size - 1):
        y += 2.0 * (x[k] * np.exp(2j * np.pi * k * t / m)).real
    y += x[-1].real * np.where(t % 2 == 0, 1.0, -1.0)
    return y / m


def mf_argmax(convs):
    """MatchedFilter._compute argmax (filters.py:333-338): first max, template-major."""
    itemp, peak = np.unravel_index(np.argmax(convs), convs.shape)
    return int(itemp), int(peak), convs[itemp, peak]


def estimate_zscore(data, loc_method="median", scale_method="iqr"):
    """stats.estimate_zscore (stats.py:418-477) for loc {median, mean, norm} and
    scale {iqr, std, mad, norm}; returns (zscores f32, loc, scale)."""
    data = np.asarray(data, dtype=np.float32)
    if loc_method == "norm":
        loc = np.zeros(1, dtype=np.float32)
    elif loc_method == "median":
        loc = np.atleast_1d(np.median(data, axis=0, keepdims=True))
    elif loc_method == "mean":
        loc = np.atleast_1d(np.mean(data, axis=0, keepdims=True, dtype=np.float64))
    else:
        raise ValueError(loc_method)
    if scale_method == "norm":
        scale = np.ones(1, dtype=np.float32)
    elif scale_method == "iqr":
        d64 = data.astype(np.float64)
        p = np.percentile(d64, [25, 75], axis=0, keepdims=True)
        scale = np.atleast_1d(np.squeeze(np.diff(p, axis=0) / 1.3489795003921634))
    elif scale_method == "std":
        scale = np.atleast_1d(np.std(data, axis=0, dtype=np.float64))
    elif scale_method == "mad":  # stats.py:505-539
        d64 = data.astype(np.float64)
        med = np.median(d64)
        mad = np.median(np.abs(d64 - med)) / 0.6744897501960817
        if np.isclose(mad, 0):
            mad = np.mean(np.abs(d64 - med)) / np.sqrt(2 / np.pi)
        scale = np.atleast_1d(mad)
    else:
        raise ValueError(scale_method)
    scale = np.where(np.isclose(scale, 0), 1, scale)
    z = np.subtract(data, loc, dtype=np.float32)
    np.divide(z, scale, out=z)
    return z, loc, scale


def symmetric_pad_window(x, window: int):
    """Padding of stats.running_filter (stats.py:231-236): numpy 'symmetric' pad of
    (w//2, w//2) for odd w and (w//2, w//2-1) for even w."""
    pad = (window // 2, window // 2) if window % 2 else (window // 2, window // 2 - 1)
    return np.pad(np.asarray(x, dtype=np.float64), pad, "symmetric")


def running_filter_mean(x, window: int) -> np.ndarray:
    """stats.running_filter(method='mean') (stats.py:200-243) in exact float64:
    ``out[t] = mean(padded[t : t + window])``."""
    p = symmetric_pad_window(x, window)
    c = np.concatenate([[0.0], np.cumsum(p)])
    return (c[window:] - c[:-window])[: len(x)] / window


def running_normalise(x, window: int) -> np.ndarray:
    """Running mean/std normalisation of one DM series (no reference function).

    ``z = (x - m) / s`` with ``m`` the window mean and ``s`` the window standard
    deviation over the symmetric-padded series (the running_filter padding), both
    in float64; ``s == 0`` is replaced by 1 (as estimate_zscore does). float32 out.
    """
    x64 = np.asarray(x, dtype=np.float64)
    p = symmetric_pad_window(x64, window)
    ref = p[0]
    d = p - ref
    c1 = np.concatenate([[0.0], np.cumsum(d)])
    c2 = np.concatenate([[0.0], np.cumsum(d * d)])
    n = x64.size
    s1 = (c1[window:] - c1[:-window])[:n]
    s2 = (c2[window:] - c2[:-window])[:n]
    mean_d = s1 / window
    var = np.maximum(s2 / window - mean_d * mean_d, 0.0)
    std = np.sqrt(var)
    std = np.where(std == 0, 1.0, std)
    return ((x64 - (mean_d + ref)) / std).astype(np.float32)


def sweep_mf(rows, window: int, widths):
    """Per-DM running normalisation + boxcar MF + argmax (float64 closed form).

    Returns arrays (snr f32, peak int32, width_index int32) over rows.
    """
    snr, peak, wid = [], [], []
    for x in rows:
        z = running_normalise(x, window)
        conv = boxcar_conv_closed(z, widths).astype(np.float32)
        it, pk, v = mf_argmax(conv)
        snr.append(v)
        peak.append(pk)
        wid.append(it)
    return (np.array(snr, dtype=np.float32), np.array(peak, dtype=np.int32),
            np.array(wid, dtype=np.int32))


# ---------------------------------------------------------------------------
# sigproc header (io/sigproc.py:46-89, 480-523, 655-680), minimal reader
# ---------------------------------------------------------------------------

_UINT_KEYS = {"telescope_id", "machine_id", "data_type", "barycentric", "pulsarcentric",
              "nbits", "nsamples", "nchans", "nifs", "ibeam", "nbeams"}
_STR_KEYS = {"rawdatafile", "source_name"}
_DOUBLE_KEYS = {"az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp", "fch1",
                "foff", "refdm", "period"}


def read_sigproc(path):
    """Return (header dict, unpacked-ready raw data bytes) of a sigproc file."""
    with open(path, "rb") as fp:
        def rs():
            n = struct.unpack("<I", fp.read(4))[0]
            return fp.read(n).decode()
        if rs() != "HEADER_START":
            raise OSError("not sigproc")
        hdr = {}
        while True:
            k = rs()
            if k == "HEADER_END":
                break
            if k in _STR_KEYS:
                hdr[k] = rs()
            elif k in _UINT_KEYS:
                hdr[k] = struct.unpack("<I", fp.read(4))[0]
            elif k in _DOUBLE_KEYS:
                hdr[k] = struct.unpack("<d", fp.read(8))[0]
            elif k == "signed":
                hdr[k] = struct.unpack("b", fp.read(1))[0]
            else:
                raise OSError(k)
        hdr["hdrlen"] = fp.tell()
        raw = np.frombuffer(fp.read(), dtype=np.uint8)
    nfile = (8 * raw.size) // hdr["nbits"] // hdr["nchans"]
    hdr["nsamples"] = min(hdr.get("nsamples", nfile), nfile)
    return hdr, raw


def unpacked_stream(hdr, raw) -> np.ndarray:
    """The reader's data stream in file order, unpacked to one byte per sample with
    the reference default bit order (io/bits.py:175-182)."""
    nbits = hdr["nbits"]
    nbytes = hdr["nsamples"] * hdr["nchans"] * nbits // 8
    raw = raw[:nbytes]
    if nbits in (1, 2, 4):
        return unpack(raw, nbits, DEFAULT_BITORDER[nbits])
    if nbits == 8:
        return raw.copy()
    raise ValueError(nbits)
