"""SIGPROC filterbank reader (mirrors ``sigpyproc.readers.FilReader``)."""

from __future__ import annotations

import logging
from pathlib import Path

import numpy as np

from sigpyproc3_amd.base import Filterbank
from sigpyproc3_amd.header import Header
from sigpyproc3_amd.io.fileio import FileReader, allocate_buffer

logger = logging.getLogger(__name__)


def plan_blocks(nsamps: int, gulp: int, skipback: int) -> tuple[int, list[tuple[int, int, int]]]:
    """The read schedule of ``read_plan`` (readers.py:184-216).

    Returns ``(gulp, [(ii, nsamps_read, skip_samples), ...])``: ``gulp`` is clipped to
    ``nsamps``, each full read is followed by a seek back of ``skipback`` samples, and a
    last partial read shorter than ``skipback`` is merged into the previous one.
    """
    gulp = min(nsamps, gulp)
    skipback = abs(skipback)
    if skipback >= gulp:
        msg = f"readsamps ({gulp}) must be > skipback ({skipback})"
        raise ValueError(msg)
    nreads, lastread = divmod(nsamps, gulp - skipback)
    if lastread < skipback:
        nreads -= 1
        lastread = nsamps - nreads * (gulp - skipback)
    blocks = [(ii, gulp, -skipback) for ii in range(nreads)]
    if lastread != 0:
        blocks.append((nreads, lastread, 0))
    return gulp, blocks


class FilReader(Filterbank):
    """Read SIGPROC filterbank file(s) (readers.py:33-233)."""

    def __init__(self, filenames, *, check_contiguity: bool = True) -> None:
        if isinstance(filenames, str | Path):
            filenames = [filenames]
        self._filenames = filenames
        self._header = Header.from_sigproc(filenames, check_contiguity=check_contiguity)
        self._file = FileReader(self.header.stream_info, mode="r", nbits=self.header.nbits)
        super().__init__()

    @property
    def header(self) -> Header:
        return self._header

    @property
    def filename(self):
        return self._filenames[0]

    @property
    def bitsinfo(self):
        return self._file.bitsinfo

    @property
    def chan_stride(self) -> float:
        """Bytes per channel sample (readers.py:88-91)."""
        return self.bitsinfo.itemsize / self.bitsinfo.bitfact

    @property
    def samp_stride(self) -> int:
        """Bytes per time sample (readers.py:93-96)."""
        return int(self.header.nchans * self.chan_stride)

    def read_block(self, start: int, nsamps: int, fch1: float | None = None,
                   nchans: int | None = None):
        """Read ``nsamps`` samples from ``start`` as a channel-major FilterbankBlock of the
        channels from ``fch1`` (readers.py:98-130)."""
        from sigpyproc3_amd.block import FilterbankBlock

        hdr = self.header
        fch1 = fch1 if fch1 is not None else hdr.fch1
        nchans = nchans if nchans is not None else hdr.nchans
        if fch1 > hdr.fch1 or nchans > hdr.nchans:
            msg = f"requested block is out of range: fch1={fch1}, nchans={nchans}"
            raise ValueError(msg)
        if start < 0 or start + nsamps > hdr.nsamples:
            msg = f"requested block is out of range: start={start}, nsamps={nsamps}"
            raise ValueError(msg)
        self._file.seek(start * self.samp_stride)
        data = self._file.cread(hdr.nchans * nsamps)
        nread = data.size // hdr.nchans
        data = data.reshape(nread, hdr.nchans).transpose()
        c0 = int((fch1 - hdr.fch1) / hdr.foff)
        new_hdr = hdr.new_header({"tstart": hdr.mjd_after_nsamps(start), "nsamples": nread,
                                  "fch1": fch1, "nchans": nchans})
        return FilterbankBlock(data[c0 : c0 + nchans], new_hdr)

    def read_plan(self, gulp: int = 16384, start: int = 0, nsamps: int | None = None,
                  skipback: int = 0, description: str | None = None, *, quiet: bool = False,
                  allocator=None):
        """Yield ``(nsamps_read, ii, data)`` gulps of unpacked samples (readers.py:171-233).

        Same schedule, buffers, allocator hook and errors as the reference; the unpack
        of 1/2/4-bit data runs on the GPU.
        """
        for nsamps_r, ii, raw, unpacked in self._gulps(gulp, start, nsamps, skipback,
                                                       allocator=allocator, unpack=True):
            yield nsamps_r, ii, unpacked if unpacked is not None else raw

    def read_plan_packed(self, gulp: int = 16384, start: int = 0, nsamps: int | None = None,
                         skipback: int = 0, *, allocator=None):
        """Like ``read_plan`` but yields the packed bytes of each gulp (no unpack)."""
        for nsamps_r, ii, raw, _ in self._gulps(gulp, start, nsamps, skipback,
                                                allocator=allocator, unpack=False):
            yield nsamps_r, ii, raw

    def read_plan_packed_into(self, buffers, gulp: int = 16384, start: int = 0,
                              nsamps: int | None = None, skipback: int = 0, pool=None):
        """The ``read_plan_packed`` schedule, reading gulp k into ``buffers[k % len(buffers)]``
        (each at least ``gulp * samp_stride`` bytes). Yields ``(nsamps_read, ii, k % n, nbytes)``;
        the caller must be done with a buffer before asking for the gulp that reuses it. With a
        thread ``pool``, each gulp is read by its threads in parallel pieces."""
        if nsamps is None:
            nsamps = self.header.nsamples - start
        gulp, blocks = plan_blocks(nsamps, gulp, skipback)
        views = [np.frombuffer(b, dtype=np.uint8) for b in buffers]
        need = gulp * self.samp_stride
        if any(v.size < need for v in views):
            msg = f"read buffers must hold {need} bytes"
            raise ValueError(msg)
        self._file.seek(start * self.samp_stride)
        nchans = self.header.nchans
        for k, (ii, nrows, skip) in enumerate(blocks):
            i = k % len(views)
            expected = int(nrows * nchans * self.chan_stride)
            # A full gulp buffer per read, as the reference's read_plan does (so a short last
            # gulp that is not at the end of the stream fails the same byte-count check).
            if pool is not None:
                nbytes = self._file.creadinto_packed_parallel(views[i][:need], pool)
            else:
                nbytes = self._file.creadinto_packed(views[i][:need])
            if nbytes != expected:
                msg = f"Unexpected number of bytes read from file {nbytes} (actual) != {expected} (expected)"
                raise ValueError(msg)
            if skip != 0:
                self._file.seek(int(skip * nchans * self.chan_stride), whence=1)
            yield nrows, ii, i, expected

    def read_packed_to_device(self, host, dev, start: int, nsamps: int, pool, piece: int = 1 << 26) -> int:
        """Packed bytes of samples ``[start, start + nsamps)`` into the pinned ``host`` tensor and
        on to the device tensor ``dev``, ``piece`` bytes at a time: the asynchronous copy of a
        piece (on the current stream) runs while the pool reads the next. Returns the bytes
        read; a short read stops early."""
        nbytes = int(nsamps) * self.samp_stride
        buf = host.numpy()
        self._file.seek(start * self.samp_stride)
        done = 0
        while done < nbytes:
            n = min(piece, nbytes - done)
            got = self._file.creadinto_packed_parallel(buf[done : done + n], pool)
            if got > 0:
                dev[done : done + got].copy_(host[done : done + got], non_blocking=True)
            done += got
            if got < n:
                break
        return done

    def _gulps(self, gulp, start, nsamps, skipback, *, allocator, unpack: bool):
        if nsamps is None:
            nsamps = self.header.nsamples - start
        gulp, blocks = plan_blocks(nsamps, gulp, skipback)
        allocator = bytearray if allocator is None else allocator
        read_buffer = allocate_buffer(allocator, gulp * self.samp_stride)
        raw_all = np.frombuffer(read_buffer, dtype=np.uint8)
        do_unpack = unpack and self.bitsinfo.unpack
        unpack_buffer = None
        if do_unpack:
            unpack_buffer = allocate_buffer(allocator, gulp * self.header.nchans)
            data_all = np.frombuffer(unpack_buffer, dtype=self.bitsinfo.dtype)
        elif self.bitsinfo.unpack:
            data_all = None
        else:
            data_all = np.frombuffer(read_buffer, dtype=self.bitsinfo.dtype)
        self._file.seek(start * self.samp_stride)
        nchans = self.header.nchans
        for ii, nrows, skip in blocks:
            block = nrows * nchans
            logger.debug("read_plan: block %d/%d, %d elements, skipback=%d", ii, len(blocks), block, skip)
            if do_unpack:
                nbytes = self._file.creadinto(read_buffer, unpack_buffer)
            else:
                nbytes = self._file.creadinto_packed(read_buffer)
            expected = int(block * self.chan_stride)
            if nbytes != expected:
                msg = f"Unexpected number of bytes read from file {nbytes} (actual) != {expected} (expected)"
                raise ValueError(msg)
            if skip != 0:
                self._file.seek(int(skip * nchans * self.chan_stride), whence=1)
            raw = raw_all[:expected]
            if do_unpack:
                yield nrows, ii, raw, data_all[:block]
            elif data_all is not None:
                yield nrows, ii, raw, data_all[:block]
            else:
                yield nrows, ii, raw, None
