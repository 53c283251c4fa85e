"""SIGPROC data stream I/O over one or more files.

The data sections of a list of SIGPROC files (each file's bytes after its header) form one
byte stream, as in the reference's multi-file reader (io/fileio.py:149-343). Here the stream
is addressed by absolute offset: ``FileReader`` keeps a stream position and a descriptor per
file, and maps a byte range of the stream onto (file, offset) spans that it reads with
``os.preadv``, which writes straight into the caller's buffer (pinned host memory on the
dedispersion path) and releases the GIL. A large request can therefore be split into
pieces read by a thread pool, across file boundaries, with no shared file cursor.
``creadinto`` unpacks 1/2/4-bit data on the GPU (``bits.unpack``); ``creadinto_packed``
leaves it packed for the kernels that unpack in registers.
"""

from __future__ import annotations

import io
import os
import threading

import numpy as np

from sigpyproc3_amd.io.bits import BitsInfo, pack, unpack


def allocate_buffer(allocator, nbytes: int):
    """Allocate ``nbytes`` with the caller's ``allocator`` and check what it returned: the
    reference's allocator hook (io/fileio.py:25-71) with the same errors."""
    if nbytes <= 0:
        msg = f"Requested buffer size ({nbytes}) is invalid, should be > 0"
        raise ValueError(msg)
    try:
        buffer = allocator(nbytes)
    except Exception as exc:
        msg = f"Failed to allocate buffer of size {nbytes}"
        raise RuntimeError(msg) from exc
    try:
        view = memoryview(buffer)
    except TypeError as exc:
        msg = f"Allocator did not return a buffer object, got {type(buffer)}"
        raise TypeError(msg) from exc
    if view.nbytes != nbytes:
        msg = f"Allocated buffer is not the expected size: {view.nbytes} (actual) != {nbytes} (expected)"
        raise ValueError(msg)
    return buffer


_READ_POOL = None
_POOL_LOCK = threading.Lock()


def read_pool():
    """Threads shared by the parallel packed reads (at most 8)."""
    global _READ_POOL
    with _POOL_LOCK:
        if _READ_POOL is None:
            from concurrent.futures import ThreadPoolExecutor

            _READ_POOL = ThreadPoolExecutor(max_workers=max(1, min(8, len(os.sched_getaffinity(0)))),
                                            thread_name_prefix="spp-read")
    return _READ_POOL


class FileReader:
    """Positioned reads of the data stream of a ``StreamInfo`` (io/sigproc.py)."""

    def __init__(self, sinfo, mode: str = "r", nbits: int = 8) -> None:
        if "r" not in mode:
            msg = f"FileReader reads only (mode {mode!r})"
            raise ValueError(msg)
        self.sinfo = sinfo
        self.bitsinfo = BitsInfo(nbits)
        self.files = list(sinfo.get_info_list("filename"))
        self._hdrlen = [int(h) for h in sinfo.get_info_list("hdrlen")]
        lens = [int(n) for n in sinfo.get_info_list("datalen")]
        self._start = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)  # stream offset of each file
        self.total = int(self._start[-1])
        self._fds: list[int | None] = [None] * len(self.files)
        self._fd_lock = threading.Lock()
        self.pos = 0
        self._fd(0)  # fail early on a missing file

    # -------------------------------------------------------------- files
    def _fd(self, i: int) -> int:
        fd = self._fds[i]
        if fd is None:
            with self._fd_lock:
                fd = self._fds[i]
                if fd is None:
                    fd = os.open(self.files[i], os.O_RDONLY)
                    self._fds[i] = fd
        return fd

    def close(self) -> None:
        for i, fd in enumerate(self._fds):
            if fd is not None:
                os.close(fd)
                self._fds[i] = None

    def __enter__(self):
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    @property
    def cur_data_pos_stream(self) -> int:
        """Stream offset (bytes) of the next read."""
        return self.pos

    def eos(self) -> bool:
        return self.pos >= self.total

    # -------------------------------------------------------------- reads
    def _read_at(self, view: memoryview, off: int) -> int:
        """Fill ``view`` from stream offset ``off`` (up to the stream end); returns the bytes
        read. Each (file, offset) span is one or more ``preadv`` calls."""
        want = min(len(view), max(0, self.total - off))
        done = 0
        i = int(np.searchsorted(self._start, off, side="right")) - 1
        while done < want:
            within = off + done - int(self._start[i])
            span = min(want - done, int(self._start[i + 1]) - (off + done))
            if span <= 0:
                i += 1
                continue
            fd = self._fd(i)
            got = 0
            while got < span:
                n = os.preadv(fd, [view[done + got : done + span]], self._hdrlen[i] + within + got)
                if n <= 0:
                    msg = f"short read in {self.files[i]} at byte {self._hdrlen[i] + within + got}"
                    raise OSError(msg)
                got += n
            done += span
            i += 1
        return done

    def creadinto_packed(self, read_buffer) -> int:
        """Fill ``read_buffer`` with the next packed stream bytes (fewer at the stream end);
        returns the bytes read."""
        view = memoryview(read_buffer).cast("B")
        n = self._read_at(view, self.pos)
        self.pos += n
        return n

    def creadinto_packed_parallel(self, read_buffer, pool, chunk: int = 1 << 23) -> int:
        """``creadinto_packed`` with the request cut into ``chunk``-byte pieces that the
        threads of ``pool`` read at their own stream offsets."""
        view = memoryview(read_buffer).cast("B")
        total = len(view)
        if pool is None or total < 2 * chunk:
            return self.creadinto_packed(read_buffer)
        base = self.pos
        got = list(pool.map(lambda a: self._read_at(view[a : min(a + chunk, total)], base + a),
                            range(0, total, chunk)))
        n = sum(got)
        self.pos += n
        return n

    def creadinto(self, read_buffer, unpack_buffer=None) -> int:
        """Fill ``read_buffer`` from the stream and, for 1/2/4-bit data, unpack it into
        ``unpack_buffer`` (the reference's creadinto, io/fileio.py:234-296). Returns the bytes
        read."""
        if self.bitsinfo.unpack and unpack_buffer is None:
            msg = "unpack_buffer should be provided when unpacking"
            raise ValueError(msg)
        nbytes = self.creadinto_packed(read_buffer)
        if self.bitsinfo.unpack:
            read_ar = np.frombuffer(memoryview(read_buffer), dtype=np.uint8)
            unpack_ar = np.frombuffer(memoryview(unpack_buffer), dtype=np.uint8)
            unpack(read_ar, self.bitsinfo.nbits, unpack_ar, bitorder=self.bitsinfo.bitorder)
        return nbytes

    def cread(self, nunits: int) -> np.ndarray:
        """Read ``nunits`` samples (unpacked for 1/2/4-bit data)."""
        nbytes = int(nunits // self.bitsinfo.bitfact) * self.bitsinfo.itemsize
        buf = np.empty(nbytes, dtype=np.uint8)
        got = self.creadinto_packed(buf) if nbytes else 0
        data = buf[:got].view(self.bitsinfo.dtype)
        if self.bitsinfo.unpack:
            return unpack(data.copy(), self.bitsinfo.nbits, bitorder=self.bitsinfo.bitorder)
        return data.copy()

    def seek(self, offset: int, whence: int = 0) -> None:
        """Move the stream position: whence 0 from the start, 1 from the current position."""
        if whence == 0:
            target = int(offset)
        elif whence == 1:
            target = self.pos + int(offset)
        else:
            msg = "whence should be either 0 (SEEK_SET) or 1 (SEEK_CUR)"
            raise ValueError(msg)
        if target < 0 or target >= self.total:
            msg = f"offset out of bounds: {target}"
            raise ValueError(msg)
        self.pos = target


class FileWriter:
    """Write a SIGPROC file: header bytes, then samples packed to ``nbits``."""

    def __init__(self, file: str, *, mode: str = "w", nbits: int = 8, rescale: bool = False) -> None:
        self.file = file
        self._fh = io.FileIO(file, mode=mode)
        self.bitsinfo = BitsInfo(nbits)
        self.rescale = rescale

    def cwrite(self, arr: np.ndarray) -> None:
        if self.bitsinfo.nbits < 32 and self.rescale:
            arr = self.bitsinfo.quantize(arr)
        if self.bitsinfo.unpack:
            pack(np.asarray(arr, dtype=np.uint8), self.bitsinfo.nbits,
                 bitorder=self.bitsinfo.bitorder).tofile(self._fh)
        else:
            np.asarray(arr).astype(self.bitsinfo.dtype, copy=False).tofile(self._fh)

    def write(self, bo: bytes) -> None:
        self._fh.write(bo)

    def close(self) -> None:
        self._fh.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc) -> None:
        self.close()
