"""Multi-file stream reader/writer (mirrors ``sigpyproc.io.fileio``).

``FileReader`` reads a logical data stream spread over one or more SIGPROC files. The
unpack step of ``creadinto`` runs on the GPU (``bits.unpack``); ``creadinto_packed``
skips it so the dedispersion path can move packed gulps to the device and unpack them
in registers.
"""

from __future__ import annotations

import io
import os

import numpy as np

from sigpyproc3_amd.io.bits import BitsInfo, pack, unpack


def allocate_buffer(allocator, nbytes: int):
    """Allocate ``nbytes`` with ``allocator`` and validate it (io/fileio.py:25-71)."""
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


class FileBase:
    """Open one file of a list at a time (io/fileio.py:74-146)."""

    def __init__(self, files: list[str], mode: str) -> None:
        self.files = files
        self.mode = mode
        self.opener = io.FileIO
        self.ifile_cur: int | None = None
        self._open(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc) -> None:
        self._close_current()

    @property
    def file_cur(self) -> str | None:
        return None if self.ifile_cur is None else self.files[self.ifile_cur]

    def close(self) -> None:
        self._close_current()

    def eos(self) -> bool:
        eof = self.file_obj.tell() == os.fstat(self.file_obj.fileno()).st_size
        return eof and self.ifile_cur == len(self.files) - 1

    def _open(self, ifile: int) -> None:
        if ifile < 0 or ifile >= len(self.files):
            msg = f"ifile should be between 0 and {len(self.files) - 1}, got {ifile}"
            raise ValueError(msg)
        if ifile != self.ifile_cur:
            fobj = self.opener(self.files[ifile], mode=self.mode)
            self._close_current()
            self.file_obj = fobj
            self.ifile_cur = ifile

    def _close_current(self) -> None:
        if self.ifile_cur is not None:
            self.file_obj.close()
            self.ifile_cur = None


_READ_POOL = None


def read_pool():
    """Threads shared by the parallel packed reads (at most 8)."""
    global _READ_POOL
    if _READ_POOL is None:
        from concurrent.futures import ThreadPoolExecutor

        _READ_POOL = ThreadPoolExecutor(max_workers=max(1, min(8, len(os.sched_getaffinity(0)))),
                                        thread_name_prefix="spp-read")
    return _READ_POOL


class FileReader(FileBase):
    """Read a multi-file SIGPROC data stream (io/fileio.py:149-343)."""

    def __init__(self, sinfo, mode: str = "r", nbits: int = 8) -> None:
        self.sinfo = sinfo
        self.bitsinfo = BitsInfo(nbits)
        super().__init__(sinfo.get_info_list("filename"), mode)
        self._seek2hdr(0)

    @property
    def cur_data_pos_file(self) -> int | None:
        if self.ifile_cur is None:
            return None
        return self.file_obj.tell() - self.sinfo.entries[self.ifile_cur].hdrlen

    @property
    def cur_data_pos_stream(self) -> int | None:
        if self.ifile_cur is None:
            return None
        if self.ifile_cur == 0:
            return self.cur_data_pos_file
        return self.cur_data_pos_file + int(self.sinfo.cumsum_datalens[self.ifile_cur - 1])

    def _readinto(self, read_buffer) -> int:
        if self.ifile_cur is None:
            msg = "No file is open for reading"
            raise OSError(msg)
        nbytes = 0
        view = memoryview(read_buffer).cast("B")
        while True:
            got = self.file_obj.readinto(view[nbytes:])
            if got is None:
                msg = "file might in non-blocking mode"
                raise BlockingIOError(msg)
            nbytes += got
            if nbytes == len(view) or self.eos():
                break
            self._seek2hdr(self.ifile_cur + 1)
        return nbytes

    def creadinto(self, read_buffer, unpack_buffer=None) -> int:
        """Fill ``read_buffer`` from the stream, unpacking into ``unpack_buffer`` for
        1/2/4-bit data (io/fileio.py:234-296). Returns the bytes read."""
        nbytes = self._readinto(read_buffer)
        if self.bitsinfo.unpack and unpack_buffer is not None:
            read_ar = np.frombuffer(memoryview(read_buffer), dtype=np.uint8)
            unpack_ar = np.frombuffer(memoryview(unpack_buffer), dtype=np.uint8)
            unpack(read_ar, self.bitsinfo.nbits, unpack_ar, bitorder=self.bitsinfo.bitorder)
        elif self.bitsinfo.unpack:
            msg = "unpack_buffer should be provided when unpacking"
            raise ValueError(msg)
        return nbytes

    def creadinto_packed(self, read_buffer) -> int:
        """Fill ``read_buffer`` with packed stream bytes (no unpack)."""
        return self._readinto(read_buffer)

    def creadinto_packed_parallel(self, read_buffer, pool, chunk: int = 1 << 23) -> int:
        """``creadinto_packed`` with the request split into ``chunk``-byte pieces that the
        threads of ``pool`` read at explicit offsets (``os.preadv`` releases the GIL, so page
        cache copies run in parallel); the stream position then moves past the request. A
        request that does not lie inside the current file (a file boundary, the end of the
        stream) takes the sequential path."""
        if self.ifile_cur is None:
            msg = "No file is open for reading"
            raise OSError(msg)
        view = memoryview(read_buffer).cast("B")
        total = len(view)
        fd = self.file_obj.fileno()
        pos = self.file_obj.tell()
        if pool is None or total < 2 * chunk or pos + total > os.fstat(fd).st_size:
            return self._readinto(read_buffer)

        def piece(off: int) -> int:
            end = min(off + chunk, total)
            got = off
            while got < end:
                n = os.preadv(fd, [view[got:end]], pos + got)
                if n <= 0:
                    msg = f"short read at byte {pos + got}"
                    raise OSError(msg)
                got += n
            return end - off

        done = sum(pool.map(piece, range(0, total, chunk)))
        self.file_obj.seek(pos + done)
        return done

    def cread(self, nunits: int) -> np.ndarray:
        """Read ``nunits`` samples (unpacked for 1/2/4-bit data)."""
        nbytes = int(nunits // self.bitsinfo.bitfact) * self.bitsinfo.itemsize
        buf = bytearray(nbytes)
        got = self._readinto(buf) if nbytes else 0
        data = np.frombuffer(bytes(buf[:got]), dtype=self.bitsinfo.dtype)
        if self.bitsinfo.unpack:
            return unpack(data.copy(), self.bitsinfo.nbits, bitorder=self.bitsinfo.bitorder)
        return data.copy()

    def seek(self, offset: int, whence: int = 0) -> None:
        if self.ifile_cur is None:
            msg = "No file is open for reading"
            raise OSError(msg)
        if whence == 0:
            self._seek_set(offset)
        elif whence == 1:
            self._seek_set(offset + self.cur_data_pos_stream)
        else:
            msg = "whence should be either 0 (SEEK_SET) or 1 (SEEK_CUR)"
            raise ValueError(msg)

    def _seek2hdr(self, ifile: int) -> None:
        self._open(ifile)
        self.file_obj.seek(self.sinfo.entries[ifile].hdrlen)

    def _seek_set(self, offset: int) -> None:
        total = self.sinfo.get_combined("datalen")
        if offset < 0 or offset >= total:
            msg = f"offset out of bounds: {offset}"
            raise ValueError(msg)
        cums = self.sinfo.cumsum_datalens
        fileid = int(np.where(offset < cums)[0][0])
        self._seek2hdr(fileid)
        within = offset if fileid == 0 else offset - int(cums[fileid - 1])
        self.file_obj.seek(within, os.SEEK_CUR)


class FileWriter(FileBase):
    """Write SIGPROC data (io/fileio.py:346-412)."""

    def __init__(self, file: str, *, mode: str = "w", nbits: int = 8, rescale: bool = False) -> None:
        super().__init__([file], mode)
        self.bitsinfo = BitsInfo(nbits)
        self.rescale = rescale

    def cwrite(self, arr: np.ndarray) -> None:
        if self.bitsinfo.nbits < 32 and self.rescale:
            arr = self.bitsinfo.quantize(arr)
        if self.bitsinfo.unpack:
            pack(np.asarray(arr, dtype=np.uint8), self.bitsinfo.nbits,
                 bitorder=self.bitsinfo.bitorder).tofile(self.file_obj)
        else:
            np.asarray(arr).astype(self.bitsinfo.dtype, copy=False).tofile(self.file_obj)

    def write(self, bo: bytes) -> None:
        self.file_obj.write(bo)
