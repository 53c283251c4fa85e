"""Host side of the batched sweep (spp_batch_*): rounds under a G budget and queue sizes.

These calls plan on the host only (no device memory, no kernels), so they run without a GPU.
The item counts are restated here from the queue's definition: stage 1 takes 128-word time
blocks x 32-byte column blocks of each stream, stage 2 takes (DM sub-tile, group of time
tiles) pairs, with 16-row DM tiles cut into sub-tiles of the shape's DM rows.
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from sigpyproc3_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def _plan(ndm, nchans, nsamps, nbits, seed):
    rng = np.random.default_rng(seed)
    lib = _lib.load()
    # smooth, increasing delays (a dispersion-like table) keep DM tiles at 16 rows
    slope = rng.uniform(0.5, 3.0)
    chan = np.arange(nchans, dtype=np.float64)
    table = np.rint(np.outer(np.arange(ndm) * slope, chan / nchans * 8.0)).astype(np.int32)
    h = ctypes.c_void_p()
    _lib.check(lib.spp_sweep_plan_create(_lib.ptr(table), ndm, nchans, nsamps, nbits, int(nbits in (2, 4)),
                                         ctypes.byref(h)))
    info = _lib.SweepInfo()
    _lib.check(lib.spp_sweep_plan_info(h, ctypes.byref(info)))
    return h, info, table


def _g_bytes(nchans, nsamps):
    gpitch = (nsamps + 63) // 64 * 64
    return (4 * nchans * gpitch + 255) // 256 * 256


def _units(n, run, x, q=8):
    cyc = run * q
    return (n // cyc) * run + min(max(n % cyc - x * run, 0), run)


@pytest.fixture(scope="module")
def plans():
    cfgs = [(40, 64, 9000, 2), (17, 96, 20000, 1), (70, 48, 5000, 8), (33, 128, 12000, 4), (5, 24, 3000, 8)]
    out = [(_plan(*c, seed=i), c) for i, c in enumerate(cfgs)]
    yield out
    for (h, _, _), _ in out:
        _lib.load().spp_sweep_plan_destroy(h)


def _batch(plans, budget):
    lib = _lib.load()
    arr = (ctypes.c_void_p * len(plans))(*[h.value for (h, _, _), _ in plans])
    b = ctypes.c_void_p()
    _lib.check(lib.spp_batch_create(arr, len(plans), budget, ctypes.byref(b)))
    return b


def _info(b):
    info = _lib.BatchInfo()
    _lib.check(_lib.load().spp_batch_get_info(b, ctypes.byref(info)))
    return info


def test_batch_info_struct_layout():
    text = (ROOT / "include" / "spp_hip.h").read_text()
    body = text[text.index("typedef struct spp_batch spp_batch;") : text.index("} spp_batch_info;")]
    fields = re.findall(r"(int64_t|int32_t)\s+(\w+);", body)
    assert [f[1] for f in fields] == [f[0] for f in _lib.BatchInfo._fields_]


@pytest.mark.parametrize("shape", range(5))
@pytest.mark.parametrize("ntt", [1, 3])
def test_queue_item_counts(plans, shape, ntt):
    lib = _lib.load()
    b = _batch(plans, 0)
    try:
        _lib.check(lib.spp_batch_set_shape(b, shape, ntt))
        info = _info(b)
        assert (info.rounds, info.segments, info.shape, info.time_tiles_per_item) == (1, 2, shape, ntt)
        rows = info.dm_rows_per_item
        assert rows in (8, 16) and 16 % rows == 0
        build = sweep = 0
        for (_, pinfo, _), (ndm, nchans, nsamps, nbits) in plans:
            gpitch = (nsamps + 63) // 64 * 64
            nby = -(-(nchans * nbits // 8) // 32)
            build += -(-gpitch // 128) * nby
            T = (8 if nbits <= 2 else 4) * 256
            ntiles_t = -(-int(pinfo.max_len) // T)
            groups = -(-ntiles_t // ntt)
            sweep += groups * int(pinfo.ndm_tiles) * (16 // rows)
        assert info.build_items == build
        assert info.sweep_items == sweep
    finally:
        lib.spp_batch_destroy(b)


def test_units_partition():
    # every unit belongs to exactly one of the 8 sub-queues
    for n in [0, 1, 5, 31, 32, 33, 100, 257, 1000]:
        for run in [1, 4, 32]:
            assert sum(_units(n, run, x) for x in range(8)) == n


def test_rounds_respect_the_budget(plans):
    lib = _lib.load()
    sizes = [_g_bytes(c[1], c[2]) for _, c in plans]
    total = sum(sizes)
    # all resident: one round
    b = _batch(plans, total)
    try:
        info = _info(b)
        assert info.rounds == 1 and info.pool_bytes == total
    finally:
        lib.spp_batch_destroy(b)
    # a budget of twice the largest copy: several rounds of two alternating pools
    budget = 2 * max(sizes)
    b = _batch(plans, budget)
    try:
        info = _info(b)
        assert info.rounds >= 2 and info.segments == 2 * info.rounds
        cap = budget // 2 // 256 * 256
        assert info.pool_bytes == cap
        rounds = np.zeros(len(plans), dtype=np.int32)
        _lib.check(lib.spp_batch_rounds(b, _lib.ptr(rounds)))
        for r in range(info.rounds):
            assert sum(s for s, rr in zip(sizes, rounds) if rr == r) <= cap
        assert set(rounds.tolist()) == set(range(info.rounds))
        assert info.workspace_bytes >= 2 * cap
    finally:
        lib.spp_batch_destroy(b)


def test_budget_too_small_is_refused(plans):
    lib = _lib.load()
    sizes = [_g_bytes(c[1], c[2]) for _, c in plans]
    arr = (ctypes.c_void_p * len(plans))(*[h.value for (h, _, _), _ in plans])
    b = ctypes.c_void_p()
    rc = lib.spp_batch_create(arr, len(plans), max(sizes), ctypes.byref(b))
    assert rc == _lib.SPP_EINVAL
    assert b"budget" in lib.spp_last_error()
    assert b.value is None


def test_bad_shape_and_empty_batch_are_refused(plans):
    lib = _lib.load()
    b = _batch(plans, 0)
    try:
        assert lib.spp_batch_set_shape(b, lib.spp_batch_shape_count(), 1) == _lib.SPP_EINVAL
        assert lib.spp_batch_set_shape(b, -1, 1) == _lib.SPP_EINVAL
        assert lib.spp_batch_set_shape(b, 0, 0) == _lib.SPP_EINVAL
        with pytest.raises(ValueError, match="not the buffer"):
            _lib.call("spp_batch_run", b, ctypes.c_void_p(1 << 20), _info(b).workspace_bytes, None)
    finally:
        lib.spp_batch_destroy(b)
    out = ctypes.c_void_p()
    assert lib.spp_batch_create(None, 0, 0, ctypes.byref(out)) == _lib.SPP_EINVAL
