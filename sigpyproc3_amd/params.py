"""Dispersion delays and header code tables.

Mirrors ``sigpyproc.params`` (params.py) for the parts the dedispersion path uses.
The delay table is computed once on the host: it is the only floating-point step of
the path and both the GPU kernels and any checker consume the same int32 table.
"""

from __future__ import annotations

import hashlib

import numpy as np

DM_CONSTANT_LK = 4.148808e3  # L&K Handbook of Pulsar Astronomy (params.py:17)

# SIGPROC data_type codes (params.py:182-193).
data_types = {
    0: "raw data",
    1: "filterbank",
    2: "time series",
    3: "pulse profiles",
    4: "amplitude spectrum",
    5: "complex spectrum",
    6: "dedispersed subbands",
    10: "PSRFITS",
}
data_types_inv = {v: k for k, v in data_types.items()}


def compute_dmdelays(freqs, dm, tsamp: float, ref_freq: float, *, in_samples: bool = True):
    """Dispersion delays of each channel for DM value(s) (params.py:26-62).

    The arithmetic must round exactly as the reference does, so the order and the
    types are the reference's: float32 channel frequencies and DMs; Python-float
    constants enter as weak scalars (numpy >= 2), so ``dm * K``, ``freqs**-2`` and the
    subtraction of ``ref_freq**-2`` are float32 operations; the quotient by ``tsamp``
    is float32; ``round`` is half-to-even; the result is int32. Returns ``[nchans]``
    for a scalar DM and ``[ndm, nchans]`` for an array.
    """
    f32 = np.atleast_1d(freqs).astype(np.float32)
    dms = np.atleast_1d(dm)[:, np.newaxis].astype(np.float32)
    scaled = dms * DM_CONSTANT_LK
    delays = scaled * ((f32**-2) - (ref_freq**-2))
    if in_samples:
        delays = (delays / tsamp).round().astype(np.int32)
    return delays.squeeze()


def table_digest(table: np.ndarray) -> str:
    """sha1 of an int32 delay table (the table depends on the host's numpy SIMD path)."""
    return hashlib.sha1(np.ascontiguousarray(table, dtype=np.int32).tobytes()).hexdigest()


def numpy_simd_features() -> str:
    """The CPU features numpy dispatches to on this host (float32 pow depends on them)."""
    try:
        from numpy._core._multiarray_umath import __cpu_features__ as feats  # numpy >= 2
    except ImportError:  # pragma: no cover
        try:
            from numpy.core._multiarray_umath import __cpu_features__ as feats
        except ImportError:
            return "unknown"
    return ",".join(k for k, v in sorted(feats.items()) if v)
