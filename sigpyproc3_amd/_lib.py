"""ctypes binding of ``libspp_hip.so`` and device-buffer plumbing.

The shared library is the product: every hot-path operation goes through it. There is
no CPU fallback -- if the library is missing the import of a kernel entry point raises
``ImportError`` with the build command. PyTorch (ROCm) provides device memory and
streams only; no torch types cross the C ABI (plain pointers and sizes).
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libspp_hip.so"

SPP_OK = 0
SPP_EINVAL = 1
SPP_EHIP = 2

# Every function the C header declares: name -> (restype, argtypes).
_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_int = ctypes.c_int
_P64 = ctypes.POINTER(ctypes.c_int64)


class SweepInfo(ctypes.Structure):
    _fields_ = [
        ("workspace_bytes", ctypes.c_int64),
        ("max_len", ctypes.c_int64),
        ("min_len", ctypes.c_int64),
        ("total_samples", ctypes.c_int64),
        ("ndm_tiles", ctypes.c_int64),
        ("max_spread", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("tables_bytes", ctypes.c_int64),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("workspace_bytes", ctypes.c_int64),
        ("pool_bytes", ctypes.c_int64),
        ("build_items", ctypes.c_int64),
        ("sweep_items", ctypes.c_int64),
        ("rounds", ctypes.c_int32),
        ("segments", ctypes.c_int32),
        ("shape", ctypes.c_int32),
        ("time_tiles_per_item", ctypes.c_int32),
        ("dm_rows_per_item", ctypes.c_int32),
        ("chans_per_stage", ctypes.c_int32),
    ]


SIGNATURES: dict[str, tuple[object, list]] = {
    "spp_last_error": (ctypes.c_char_p, []),
    "spp_version": (_int, []),
    "spp_unpack": (_int, [_p, _i64, _p, _int, _int, _p]),
    "spp_dedisperse": (_int, [_p, _int, _int, _i64, _i64, _p, _i64, _p, _i64, _p]),
    "spp_subband": (_int, [_p, _int, _int, _i64, _i64, _p, _p, _i64, _i64, _p, _i64, _p]),
    "spp_sweep_plan_create": (_int, [_p, _i64, _i64, _i64, _int, _int, ctypes.POINTER(_p)]),
    "spp_sweep_plan_destroy": (_int, [_p]),
    "spp_sweep_variant_count": (_int, []),
    "spp_sweep_plan_set_variant": (_int, [_p, _int]),
    "spp_sweep_plan_info": (_int, [_p, ctypes.POINTER(SweepInfo)]),
    "spp_sweep_plan_row_lengths": (_int, [_p, _p]),
    "spp_sweep_plan_upload": (_int, [_p, _p, _i64, _p]),
    "spp_sweep_execute": (_int, [_p, _p, _p, _i64, _p, _i64, _p]),
    "spp_sweep_build": (_int, [_p, _p, _p, _i64, _p]),
    "spp_sweep_run": (_int, [_p, _p, _i64, _p, _i64, _p]),
    "spp_sweep_execute_n": (_int, [_p, _p, _i64, _p, _i64, _p, _i64, _p]),
    "spp_sweep_g_bytes": (_int, [_p, _i64, _P64]),
    "spp_sweep_execute_g": (_int, [_p, _p, _i64, _p, _i64, _p, _i64, _p, _i64, _p]),
    "spp_batch_create": (_int, [_p, _i32, _i64, ctypes.POINTER(_p)]),
    "spp_batch_destroy": (_int, [_p]),
    "spp_batch_shape_count": (_int, []),
    "spp_batch_set_shape": (_int, [_p, _int, _i32]),
    "spp_batch_get_info": (_int, [_p, ctypes.POINTER(BatchInfo)]),
    "spp_batch_rounds": (_int, [_p, _p]),
    "spp_batch_upload": (_int, [_p, _p, _i64, _p]),
    "spp_batch_bind": (_int, [_p, _p, _p, _p, _p, _i64, _p]),
    "spp_batch_run": (_int, [_p, _p, _i64, _p]),
    "spp_batch_capture": (_int, [_p, _p, _i64]),
    "spp_batch_replay": (_int, [_p, _p]),
    "spp_batch_errors": (_int, [_p, _p, _i64, _p, ctypes.POINTER(_i32)]),
    "spp_probe_lds": (_int, [_int, _i64, _i32, _p, _p]),
    "spp_bandpass_workspace_size": (_int, [_int, _i64, _i64, _P64]),
    "spp_bandpass": (_int, [_p, _int, _int, _i64, _i64, _p, _int, _p, _i64, _p]),
    "spp_collapse": (_int, [_p, _int, _int, _i64, _i64, _p, _i64, _p]),
    "spp_mask_bytes": (_int, [_p, _i64, _i64, _p, _p, _p]),
    "spp_remove_zerodm": (_int, [_p, _int, _int, _i64, _i64, _p, _p, _p, _p]),
    "spp_fold_workspace_size": (_int, [_int, _i64, _i64, _i32, _P64]),
    "spp_fold": (_int, [_p, _int, _int, _i64, _i64, _p, _p, _i64, _p, _p, _i32, _p, _i32, _i32, _p, _p, _p, _i64,
                        _p]),
    "spp_roll_rows": (_int, [_p, _i64, _i64, _p, _p, _i64, _p, _p]),
    "spp_dmt": (_int, [_p, _i64, _i64, _p, _p, _i64, _i64, _p, _p]),
    "spp_running_normalise_workspace_size": (_int, [_i64, _p, _i32, _P64]),
    "spp_running_normalise": (_int, [_p, _i64, _i64, _p, _i32, _p, _i64, _p, _i64, _p]),
    "spp_boxcar_mf_workspace_size": (_int, [_i64, _p, _p, _i32, _P64]),
    "spp_boxcar_mf": (_int, [_p, _i64, _i64, _p, _p, _i32, _p, _p, _p, _p, _i64, _p]),
    "spp_boxcar_mf_noalloc": (_int, [_p, _i64, _i64, _p, _p, _i32, _p, _p, _p, _p]),
    "spp_normalise_boxcar_mf_workspace_size": (_int, [_i64, _p, _p, _i32, _P64]),
    "spp_normalise_boxcar_mf": (_int, [_p, _i64, _i64, _p, _i32, _p, _i32, _p, _p, _p, _p, _i64, _p]),
    "spp_convolve_templates_workspace_size": (_int, [_i64, _i32, _i64, _P64]),
    "spp_convolve_templates": (_int, [_p, _i64, _p, _p, _p, _i32, _p, _p, _i64, _p]),
}

_lib = None
_lock = threading.Lock()


def load(path: str | Path | None = None) -> ctypes.CDLL:
    """Load the library (once) and bind every declared entry point."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            msg = (
                f"{p} is missing: the HIP extension is required (there is no CPU path). "
                "Build it with `python -m sigpyproc3_amd.build`."
            )
            raise ImportError(msg)
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    """Map a C status to the reference's exception types."""
    if rc == SPP_OK:
        return
    msg = (load().spp_last_error() or b"").decode(errors="replace")
    if rc == SPP_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg or f"libspp_hip error {rc}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def workspace_size(name: str, *args) -> int:
    """Bytes reported by a ``*_workspace_size`` entry point (its last argument is the output)."""
    out = np.zeros(1, dtype=np.int64)
    call(name, *args, out.ctypes.data_as(_P64))
    return int(out[0])


def workspace(nbytes: int):
    """A device byte buffer of at least ``nbytes`` (one byte when none is needed)."""
    return empty((max(int(nbytes), 1),), "uint8")


# ----------------------------------------------------------------------------- torch plumbing

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as _t

        _torch = _t
    return _torch


def require_gpu() -> None:
    t = torch()
    if not t.cuda.is_available():
        msg = "no GPU visible: libspp_hip runs on an AMD Instinct GPU (gfx950); there is no CPU path"
        raise RuntimeError(msg)


def device():
    require_gpu()
    return torch().device("cuda", torch().cuda.current_device())


def stream_handle(stream=None) -> ctypes.c_void_p:
    """hipStream_t of ``stream`` (default: torch's current stream on the current device)."""
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return ctypes.c_void_p(int(s.cuda_stream))


def ptr(x) -> ctypes.c_void_p:
    """Device address of a torch tensor, or host address of a numpy array."""
    if isinstance(x, np.ndarray):
        return ctypes.c_void_p(x.ctypes.data)
    return ctypes.c_void_p(int(x.data_ptr()))


_NP_TO_TORCH = {
    np.dtype(np.uint8): "uint8",
    np.dtype(np.int32): "int32",
    np.dtype(np.int64): "int64",
    np.dtype(np.float32): "float32",
    np.dtype(np.float64): "float64",
}


def to_device(a: np.ndarray):
    """Copy a host numpy array to a new contiguous device tensor."""
    t = torch()
    a = np.ascontiguousarray(a)
    return t.from_numpy(a).to(device(), non_blocking=False)


def empty(shape, dtype: str):
    t = torch()
    return t.empty(shape, dtype=getattr(t, dtype), device=device())


def zeros(shape, dtype: str):
    t = torch()
    return t.zeros(shape, dtype=getattr(t, dtype), device=device())


def is_device_tensor(x) -> bool:
    t = torch() if _torch is not None or _torch_importable() else None
    return t is not None and isinstance(x, t.Tensor) and x.is_cuda


def _torch_importable() -> bool:
    try:
        torch()
    except ImportError:
        return False
    return True
