"""SIGPROC header parsing and encoding (the header side of ``sigpyproc.io.sigproc``).

A SIGPROC header is a sequence of length-prefixed ASCII keys, each followed by a
little-endian value whose type the key fixes, between ``HEADER_START`` and
``HEADER_END`` (io/sigproc.py:46-89, 480-523). The sample count is derived from the
file size and capped by a declared ``nsamples`` (io/sigproc.py:655-680). Sky-coordinate
parsing (astropy) is outside the dedispersion path: ``src_raj``/``src_dej`` stay floats.
"""

from __future__ import annotations

import logging
import struct
from dataclasses import dataclass, field
from pathlib import Path
from typing import BinaryIO

import numpy as np

logger = logging.getLogger(__name__)

_U32 = struct.Struct("<I")
_F64 = struct.Struct("<d")
_I8 = struct.Struct("b")

# key -> value codec (None = string), in the reference's encoding order.
SIGPROC_SCHEMA: dict[str, struct.Struct | None] = {
    "telescope_id": _U32,
    "machine_id": _U32,
    "data_type": _U32,
    "rawdatafile": None,
    "source_name": None,
    "barycentric": _U32,
    "pulsarcentric": _U32,
    "az_start": _F64,
    "za_start": _F64,
    "src_raj": _F64,
    "src_dej": _F64,
    "tstart": _F64,
    "tsamp": _F64,
    "nbits": _U32,
    "nsamples": _U32,
    "fch1": _F64,
    "foff": _F64,
    "nchans": _U32,
    "nifs": _U32,
    "refdm": _F64,
    "period": _F64,
    "signed": _I8,
    "ibeam": _U32,
    "nbeams": _U32,
}

telescope_ids = {
    "Fake": 0, "Arecibo": 1, "Ooty": 2, "Nancay": 3, "Parkes": 4, "Jodrell": 5, "GBT": 6,
    "GMRT": 7, "Effelsberg": 8, "Effelsberg LOFAR": 9, "SRT": 10, "LOFAR": 11, "VLA": 12,
    "CHIME": 20, "MWA": 30, "MeerKAT": 64, "NC": 40, "NGNC": 41,
}
machine_ids = {
    "FAKE": 0, "PSPM": 1, "WAPP": 2, "AOFTM": 3, "BPP": 4, "OOTY": 5, "SCAMP": 6, "GMRTFB": 7,
    "PULSAR2000": 8, "PARSPEC": 9, "BPSR": 10, "COBALT": 11, "GMRTNEW": 14, "CHIME": 20,
    "MWA-VCS": 30, "MWAX-VCS": 31, "MWAX-RTB": 32, "ADU": 40, "iTPM": 41,
}
telescope_names = {v: k for k, v in telescope_ids.items()}
machine_names = {v: k for k, v in machine_ids.items()}


@dataclass(frozen=True)
class FileInfo:
    """One file of a (possibly multi-file) stream (io/sigproc.py:131-168)."""

    filename: str
    hdrlen: int
    datalen: int
    nsamples: int
    tstart: float
    tsamp: float

    @property
    def tend(self) -> float:
        return self.tstart + ((self.nsamples - 1) * self.tsamp) / 86400


@dataclass
class StreamInfo:
    """Ordered files of a data stream (io/sigproc.py:171-208)."""

    entries: list[FileInfo] = field(default_factory=list)

    @property
    def cumsum_datalens(self) -> np.ndarray:
        return np.cumsum(self.get_info_list("datalen"))

    @property
    def time_gaps(self) -> np.ndarray:
        tstart = np.array(self.get_info_list("tstart"))
        tend = np.array([e.tend for e in self.entries])
        if len(tstart) == 1:
            return np.array([0])
        return (tstart[1:] - tend[:-1]) * 86400

    def add_entry(self, finfo: FileInfo) -> None:
        self.entries.append(finfo)

    def get_combined(self, key: str):
        return sum(self.get_info_list(key))

    def get_info_list(self, key: str) -> list:
        return [getattr(e, key) for e in self.entries]

    def check_contiguity(self) -> bool:
        tsamps = self.get_info_list("tsamp")
        same = len(set(tsamps)) == 1
        return same and bool(np.allclose(self.time_gaps, tsamps[0], rtol=0.1))


def _read_string(fp: BinaryIO) -> str:
    raw = fp.read(4)
    if len(raw) != 4:
        msg = "EOF while reading string length"
        raise struct.error(msg)
    n = _U32.unpack(raw)[0]
    if not 1 <= n <= 4096:
        msg = f"Implausible string length {n}"
        raise ValueError(msg)
    body = fp.read(n)
    if len(body) != n:
        msg = "EOF while reading string body"
        raise struct.error(msg)
    return body.decode("utf-8", errors="strict")


def _read_value(fp: BinaryIO, codec: struct.Struct):
    raw = fp.read(codec.size)
    if len(raw) != codec.size:
        msg = "EOF while reading a numeric field"
        raise struct.error(msg)
    return codec.unpack(raw)[0]


def _skip_unknown(fp: BinaryIO, path: Path) -> None:
    """Skip an unknown key's value by trying 4/8/1-byte widths until the next token is a
    known key (io/sigproc.py:604-652)."""
    start = fp.tell()
    for width in (4, 8, 1):
        fp.seek(start)
        if len(fp.read(width)) < width:
            continue
        peek = fp.tell()
        try:
            nxt = _read_string(fp)
        except (struct.error, UnicodeDecodeError, ValueError):
            continue
        if nxt in SIGPROC_SCHEMA or nxt == "HEADER_END":
            fp.seek(peek)
            return
    msg = f"{path}: cannot determine the size of an unknown header value at offset {start}"
    raise OSError(msg)


def _read_header(fp: BinaryIO, path: Path) -> dict:
    try:
        first = _read_string(fp)
    except (struct.error, UnicodeDecodeError, ValueError) as exc:
        msg = f"{path}: Header is not in sigproc format... Is file empty?."
        raise OSError(msg) from exc
    if first != "HEADER_START":
        msg = f"{path}: Header is not in sigproc format."
        raise OSError(msg)
    hdr: dict = {}
    while True:
        try:
            key = _read_string(fp)
        except (struct.error, UnicodeDecodeError, ValueError) as exc:
            msg = f"{path}: Header truncated while reading key at byte offset {fp.tell()}."
            raise OSError(msg) from exc
        if key == "HEADER_END":
            break
        if key not in SIGPROC_SCHEMA:
            logger.warning("%s: unknown header key %s, skipping", path, key)
            _skip_unknown(fp, path)
            continue
        codec = SIGPROC_SCHEMA[key]
        try:
            hdr[key] = _read_string(fp) if codec is None else _read_value(fp, codec)
        except (struct.error, UnicodeDecodeError, ValueError) as exc:
            msg = f"{path}: Header truncated while reading value for key {key!r}."
            raise OSError(msg) from exc
    hdr["hdrlen"] = fp.tell()
    return hdr


def _nsamples(hdr: dict) -> int:
    missing = {"nbits", "nchans", "datalen"} - hdr.keys()
    if missing:
        msg = f"Cannot compute nsamples: missing required header keys {missing}."
        raise ValueError(msg)
    nbits, nchans, datalen = int(hdr["nbits"]), int(hdr["nchans"]), int(hdr["datalen"])
    if nbits <= 0 or nchans <= 0:
        msg = f"Invalid header values: nbits={nbits}, nchans={nchans} (must be > 0)."
        raise ValueError(msg)
    from_file = (8 * datalen) // nbits // nchans
    if "nsamples" in hdr:
        return min(int(hdr["nsamples"]), from_file)
    return from_file


def parse_header(filename: str | Path) -> dict:
    """Parse one SIGPROC file's header, plus hdrlen/filelen/datalen/nsamples/filename."""
    path = Path(filename).expanduser().resolve()
    if not path.is_file():
        msg = f"{path} does not exist or is not a file"
        raise FileNotFoundError(msg)
    with path.open("rb") as fp:
        hdr = _read_header(fp, path)
        fp.seek(0, 2)
        hdr["filelen"] = fp.tell()
    hdr["datalen"] = hdr["filelen"] - hdr["hdrlen"]
    if hdr["datalen"] < 0:
        msg = f"{path}: file shorter than its header"
        raise OSError(msg)
    hdr["nsamples"] = _nsamples(hdr)
    hdr["filename"] = path.as_posix()
    return hdr


def match_header(h1: dict, h2: dict) -> None:
    for key, value in h1.items():
        if key in {"tstart", "rawdatafile"} or key not in SIGPROC_SCHEMA:
            continue
        if value != h2[key]:
            msg = f"Header key ({key} = {value}) and ({key} = {h2[key]}) do not match for file {h2['filename']}"
            raise ValueError(msg)


def parse_header_multi(filenames, *, check_contiguity: bool = True) -> tuple[dict, StreamInfo]:
    """Parse a file or a contiguous sequence of files (io/sigproc.py:279-312)."""
    if isinstance(filenames, str | Path):
        filenames = [filenames]
    hdr = parse_header(filenames[0])

    def info(h: dict) -> FileInfo:
        return FileInfo(filename=str(h["filename"]), hdrlen=int(h["hdrlen"]),
                        datalen=int(h["datalen"]), nsamples=int(h["nsamples"]),
                        tstart=float(h.get("tstart", 0.0)), tsamp=float(h["tsamp"]))

    sinfo = StreamInfo([info(hdr)])
    if len(filenames) > 1:
        for fn in filenames[1:]:
            h = parse_header(fn)
            match_header(hdr, h)
            sinfo.add_entry(info(h))
        if check_contiguity and not sinfo.check_contiguity():
            msg = f"Files {filenames} are not contiguous"
            raise ValueError(msg)
    hdr["nsamples"] = sinfo.get_combined("nsamples")
    return hdr, sinfo


def _encode_string(value: str) -> bytes:
    b = value.encode("utf-8")
    return _U32.pack(len(b)) + b


def encode_key(key: str, value=None) -> bytes:
    out = _encode_string(key)
    if value is None:
        return out
    if key not in SIGPROC_SCHEMA:
        msg = f"Unknown SIGPROC header key {key!r} cannot be encoded."
        raise KeyError(msg)
    codec = SIGPROC_SCHEMA[key]
    if codec is None:
        if not isinstance(value, str):
            msg = f"Key {key!r} expects a string value, got {value}."
            raise TypeError(msg)
        return out + _encode_string(value)
    return out + codec.pack(value)


def encode_header(header: dict, *, allow_nsamples_overwrite: bool = False) -> bytes:
    """Binary SIGPROC header in schema order (io/sigproc.py:386-410)."""
    out = _encode_string("HEADER_START")
    for key in SIGPROC_SCHEMA:
        if key == "nsamples" and not allow_nsamples_overwrite:
            continue
        if key in header:
            out += encode_key(key, header[key])
    return out + _encode_string("HEADER_END")
