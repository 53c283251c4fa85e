"""GPU parity of the batched sweep (one persistent launch for many plans).

Every plane of a batch must equal its own plan's one-plan sweep (SweepPlan.execute, the
plain GPU form) bit for bit, for every block shape and time-tile grouping, when the G budget
cuts the batch into rounds, and on every replay of a captured HIP graph with new data copied
into the bound streams. A few rows are also checked against the C oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_oracle
from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu

# (nbits, nchans, nsamps, ndm, max_dm, tsamp): 1-bit rows of 42 bytes (unaligned), 32-bit
# lanes for 8-bit data and for 4-bit data at 8192 channels and 2-bit data at 22000 channels
INSTANCES = [
    (1, 336, 30000, 40, 600.0, 256e-6),
    (2, 1024, 50000, 70, 400.0, 64e-6),
    (4, 336, 24000, 33, 900.0, 256e-6),
    (8, 1024, 40000, 52, 300.0, 64e-6),
    (2, 4096, 20000, 16, 100.0, 64e-6),
    (8, 336, 9000, 5, 200.0, 64e-6),
    (4, 8192, 12000, 20, 60.0, 64e-6),
    (2, 22000, 6000, 9, 20.0, 64e-6),
    (1, 1024, 70000, 100, 1500.0, 256e-6),
]


def _table(nchans, ndm, max_dm, tsamp):
    f = O.chan_freqs(nchans, -300.0 / nchans, 1550.0)
    dms = np.linspace(0.0, max_dm, ndm)
    return np.atleast_2d(O.compute_dmdelays(f, dms, tsamp, 1550.0)).astype(np.int32)


def _unpack(raw, nbits, nchans):
    return O.unpack(raw, nbits, "big" if nbits in (2, 4) else "little").reshape(-1, nchans) if nbits < 8 \
        else raw.reshape(-1, nchans)


@pytest.fixture(scope="module")
def case():
    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.sweep import SweepPlan

    rng = np.random.default_rng(5)
    plans, packed, host, ref = [], [], [], []
    for k, (nbits, nchans, nsamps, ndm, max_dm, tsamp) in enumerate(INSTANCES):
        table = _table(nchans, ndm, max_dm, tsamp)
        raw = rng.integers(0, 256, nsamps * nchans * nbits // 8, dtype=np.uint8)
        plan = SweepPlan(table, nsamps, nbits, int(nbits in (2, 4)))
        dev = _lib.to_device(raw)
        plans.append(plan)
        packed.append(dev)
        host.append(raw)
        ref.append(plan.execute(dev).clone())
        print(f"instance {k}: {nbits}-bit {nchans} ch x {nsamps}, {ndm} DMs, plain plane ready", flush=True)
    torch.cuda.synchronize()
    return plans, packed, host, ref


def _compare(planes, ref):
    import torch

    for k, (a, b) in enumerate(zip(planes, ref)):
        assert torch.equal(a, b), f"instance {k}: batch plane differs from the one-plan sweep"


@pytest.mark.parametrize("shape,ntt", [(0, 1), (1, 2), (2, 1), (3, 4), (0, 3), (4, 1), (4, 2)])
def test_batch_equals_one_plan_sweeps(case, shape, ntt):
    from sigpyproc3_amd.batch import SweepBatch

    plans, packed, _, ref = case
    print(f"batch shape {shape} x {ntt} time tiles per item", flush=True)
    b = SweepBatch(plans)
    assert b.info.rounds == 1
    b.set_shape(shape, ntt)
    planes = b.execute(packed)
    _compare(planes, ref)


def test_batch_rows_against_oracle(case):
    from sigpyproc3_amd.batch import SweepBatch

    plans, packed, host, _ = case
    b = SweepBatch(plans)
    planes = b.execute(packed)
    for k in (0, 3, 6):
        nbits, nchans = INSTANCES[k][0], INSTANCES[k][1]
        x = _unpack(host[k], nbits, nchans)
        rows = [0, plans[k].ndm // 2, plans[k].ndm - 1]
        got = planes[k].cpu().numpy()
        exp = c_oracle.sweep(x, nchans, plans[k].table[rows], ld=got.shape[1])
        for j, r in enumerate(rows):
            n = int(plans[k].row_len[r])
            np.testing.assert_array_equal(got[r, :n], exp[j, :n], err_msg=f"instance {k} row {r}")
            assert not got[r, n:].any()


def test_batch_in_rounds(case):
    from sigpyproc3_amd.batch import SweepBatch

    plans, packed, _, ref = case
    sizes = [4 * p.nchans * ((p.nsamps + 63) // 64 * 64) for p in plans]
    b = SweepBatch(plans, g_budget=2 * max(sizes) + 4096)
    print(f"rounds {b.info.rounds}, pool {b.info.pool_bytes} bytes", flush=True)
    assert b.info.rounds >= 3
    for shape, ntt in [(0, 1), (2, 2)]:
        b.set_shape(shape, ntt)
        planes = b.execute(packed)
        _compare(planes, ref)


def test_batch_graph_replays_new_gulps(case):
    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.batch import SweepBatch

    plans, packed, _, _ = case
    rng = np.random.default_rng(11)
    bufs = [torch.empty_like(p) for p in packed]
    b = SweepBatch(plans)
    planes = b.bind(bufs)
    b.capture()
    for it in range(3):
        # the next set of gulps, copied into the bound streams
        fresh = [_lib.to_device(rng.integers(0, 256, int(p.numel()), dtype=np.uint8)) for p in packed]
        for dst, src in zip(bufs, fresh):
            dst.copy_(src)
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        assert b.errors() == 0
        expect = [pl.execute(src).clone() for pl, src in zip(plans, fresh)]
        torch.cuda.synchronize()
        print(f"replay {it}: comparing", flush=True)
        _compare(planes, expect)


def test_batch_tuner_keeps_results(case, tmp_path, monkeypatch):
    from sigpyproc3_amd.batch import SweepBatch

    monkeypatch.setenv("SPP_TUNE_CACHE", str(tmp_path / "tune.json"))
    plans, packed, _, ref = case
    b = SweepBatch(plans)
    b.bind(packed)
    best = b.tune(repeats=1)
    assert b.shape == best
    planes = b.execute(packed)
    _compare(planes, ref)
    # the cached choice is applied without timing
    b2 = SweepBatch(plans)
    b2.bind(packed)
    assert b2.tune(repeats=1) == best
