"""Bit packing/unpacking and bit-depth metadata.

Mirrors ``sigpyproc.io.bits`` (io/bits.py). ``unpack`` runs on the GPU through
``spp_unpack``; ``pack`` (the write path, outside the dedispersion path) is host numpy.
"""

from __future__ import annotations

from typing import Any, ClassVar

import attrs
import numpy as np

from sigpyproc3_amd import _lib


def _check_args(array, nbits: int, bitorder: str) -> str:
    dtype = array.dtype if isinstance(array, np.ndarray) else None
    if dtype is None:
        dtype_ok = str(array.dtype) == "torch.uint8"
    else:
        dtype_ok = dtype == np.uint8
    if not dtype_ok:
        msg = f"Input array must be uint8, got {array.dtype}"
        raise ValueError(msg)
    if nbits not in {1, 2, 4}:
        msg = f"nbits must be 1, 2, or 4, got {nbits}"
        raise ValueError(msg)
    if (not bitorder) or (bitorder[0] not in {"b", "l"}):
        msg = f"bitorder must be 'big' or 'little', got {bitorder}"
        raise ValueError(msg)
    return "big" if bitorder[0] == "b" else "little"


def unpack(array, nbits: int, unpacked=None, *, bitorder: str = "big"):
    """Unpack 1, 2 and 4-bit data packed as bytes (io/bits.py:19-70).

    ``array`` may be a host numpy array or a device tensor; the result has the same
    kind (a numpy ``unpacked`` is filled in place). Raises ``ValueError`` for a non-uint8
    input, an nbits outside {1, 2, 4}, a bad bit order or a wrong-sized output.
    """
    order = _check_args(array, nbits, bitorder)
    bitfact = 8 // nbits
    size = int(array.numel()) if not isinstance(array, np.ndarray) else array.size
    if unpacked is None:
        if isinstance(array, np.ndarray):
            unpacked = np.zeros(size * bitfact, dtype=np.uint8)
        else:
            unpacked = _lib.empty((size * bitfact,), "uint8")
    else:
        usize = unpacked.size if isinstance(unpacked, np.ndarray) else int(unpacked.numel())
        if usize != size * bitfact:
            msg = f"Unpacking array must be {bitfact} x input size, got {usize}"
            raise ValueError(msg)
    if size == 0:
        return unpacked
    from sigpyproc3_amd.core import kernels

    getattr(kernels, f"unpack{nbits:d}_8_{order}")(array, unpacked)
    return unpacked


def pack(array: np.ndarray, nbits: int, packed: np.ndarray | None = None, *,
         bitorder: str = "big") -> np.ndarray:
    """Pack 1, 2 and 4-bit samples into bytes (io/bits.py:73-124), host numpy."""
    array = np.asarray(array)
    if array.dtype != np.uint8:
        msg = f"Input array must be uint8, got {array.dtype}"
        raise ValueError(msg)
    if nbits not in {1, 2, 4}:
        msg = f"nbits must be 1, 2, or 4, got {nbits}"
        raise ValueError(msg)
    if (not bitorder) or (bitorder[0] not in {"b", "l"}):
        msg = f"bitorder must be 'big' or 'little', got {bitorder}"
        raise ValueError(msg)
    bitfact = 8 // nbits
    if packed is None:
        packed = np.zeros(array.size // bitfact, dtype=np.uint8)
    elif packed.size != array.size // bitfact:
        msg = f"packing array must be input size // {bitfact}, got {packed.size}"
        raise ValueError(msg)
    # Values are shifted and OR-ed without masking, as kernels.pack* do (core/kernels.py:72-173):
    # an out-of-range sample spills into the neighbouring fields; bits above 7 are dropped.
    a = array[: packed.size * bitfact].reshape(-1, bitfact).astype(np.uint32)
    acc = np.zeros(packed.size, dtype=np.uint32)
    for k in range(bitfact):
        shift = (8 - (k + 1) * nbits) if bitorder[0] == "b" else k * nbits
        acc |= a[:, k] << shift
    packed[:] = acc.astype(np.uint8)
    return packed


@attrs.frozen(auto_attribs=True)
class BitsInfo:
    """Bit-depth metadata (io/bits.py:127-328)."""

    nbits_to_dtype: ClassVar[dict[int, str]] = {
        1: "<u1", 2: "<u1", 4: "<u1", 8: "<u1", 16: "<u2", 32: "<f4",
    }
    default_sigma: ClassVar[dict[int, float]] = {1: 0.5, 2: 1.5, 4: 6, 8: 6, 16: 6, 32: 6}
    default_bitorder: ClassVar[dict[int, str]] = {
        1: "little", 2: "big", 4: "big", 8: "invalid", 16: "invalid", 32: "invalid",
    }

    nbits: int = attrs.field(validator=attrs.validators.in_(nbits_to_dtype))
    digi_sigma: float = attrs.field(
        default=attrs.Factory(lambda self: self.default_sigma[self.nbits], takes_self=True),
    )

    @property
    def dtype(self) -> np.dtype[Any]:
        return np.dtype(self.nbits_to_dtype[self.nbits])

    @property
    def itemsize(self) -> int:
        return self.dtype.itemsize

    @property
    def unpack(self) -> bool:
        return bool(self.nbits in {1, 2, 4})

    @property
    def bitfact(self) -> int:
        return 8 // self.nbits if self.unpack else 1

    @property
    def bitorder(self) -> str:
        return self.default_bitorder[self.nbits]

    @property
    def digi_min(self) -> int:
        return 0

    @property
    def digi_max(self) -> int:
        return (1 << self.nbits) - 1

    @property
    def digi_mean(self) -> float:
        return (1 << (self.nbits - 1)) - 0.5

    @property
    def digi_scale(self) -> float:
        return self.digi_mean / self.digi_sigma

    def to_dict(self) -> dict[str, Any]:
        attributes = attrs.asdict(self)
        prop = {k: getattr(self, k) for k, v in vars(type(self)).items() if isinstance(v, property)}
        attributes.update(prop)
        return attributes

    def quantize(self, arr_norm: np.ndarray) -> np.ndarray:
        """Quantize normalised data to this bit depth (io/bits.py:306-328)."""
        arr = (arr_norm * self.digi_scale) + self.digi_mean + 0.5
        arr = arr.astype(np.int32)
        np.clip(arr, self.digi_min, self.digi_max, out=arr)
        return arr.astype(self.dtype, copy=False)
