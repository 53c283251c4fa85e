"""The 64 uneven sweep instances (1/2/4/8-bit, 336/1024/4096 channels, 2^16..2^19 samples,
256..4096 DM trials up to DM 500..3000, tsamp 64/256 us), all at full size in ONE persistent
batch launch.

The planes of all 64 instances stay resident (about 185 GB); the stage-1 copies get what HBM
is left as their budget, so the launch runs in rounds of two alternating pools. Per instance:
every plane row equals the plain one-plan sweep (SweepPlan.execute) bit for bit; the exact
prefix-sum checksum of every row (independent torch integer ops on the unpacked stream); the
zero tail of every row; the injected pulse found at its DM row and time; and for every fourth
instance one row bit-exact against the C oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_oracle

pytestmark = pytest.mark.gpu

NINST = 64


def _unpack(packed, nbits: int, torch):
    if nbits == 8:
        return packed
    p = packed.to(torch.int32)
    per = 8 // nbits
    mask = (1 << nbits) - 1
    big = nbits in (2, 4)
    shifts = [(8 - (k + 1) * nbits) if big else k * nbits for k in range(per)]
    parts = [(p >> s) & mask for s in shifts]
    return torch.stack(parts, dim=1).reshape(-1).to(torch.uint8)


def _row_sums(x, table, row_len, torch):
    n, nch = x.shape
    d = torch.from_numpy(table.astype(np.int64)).cuda()
    L = torch.from_numpy(row_len).cuda()
    expect = torch.zeros(table.shape[0], dtype=torch.int64, device="cuda")
    for cb in range(0, nch, 256):
        blk = x[:, cb : cb + 256].to(torch.int64)
        pre = torch.zeros((n + 1, blk.shape[1]), dtype=torch.int64, device="cuda")
        pre[1:] = torch.cumsum(blk, dim=0)
        dd = d[:, cb : cb + 256]
        expect += (torch.gather(pre, 0, (dd + L[:, None]).clamp(max=n)) - torch.gather(pre, 0, dd)).sum(dim=1)
        del blk, pre
    return expect


@pytest.fixture(scope="module")
def batch_run():
    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.batch import SweepBatch
    from sigpyproc3_amd.sweep import SweepPlan
    from sigpyproc3_amd.synth import cfg_batch_instance, generate_packed

    cfgs, plans, packed = [], [], []
    for i in range(NINST):
        cfg, dms = cfg_batch_instance(i)
        table = cfg.delays(dms).astype(np.int32)
        plans.append(SweepPlan(table, cfg.nsamps, cfg.nbits, 1 if cfg.nbits in (2, 4) else 0, allocate=False))
        packed.append(generate_packed(cfg, device=_lib.device()))
        cfgs.append((cfg, dms))
        if i % 8 == 7:
            print(f"planned and generated {i + 1} instances", flush=True)
    planes = [_lib.empty((p.ndm, p.max_len), "float32") for p in plans]
    free, _ = torch.cuda.mem_get_info()
    g_total = sum(4 * p.nchans * ((p.nsamps + 63) // 64 * 64) for p in plans)
    budget = min(g_total, free - (12 << 30))
    batch = SweepBatch(plans, g_budget=budget)
    print(f"planes {sum(int(x.numel()) * 4 for x in planes) / 2**30:.1f} GiB; G {g_total / 2**30:.1f} GiB in "
          f"{batch.info.rounds} rounds of {batch.info.pool_bytes / 2**30:.1f} GiB pools; "
          f"{batch.info.sweep_items} stage-2 items", flush=True)
    batch.bind(packed, planes)
    batch.run()
    torch.cuda.synchronize()
    assert batch.errors() == 0
    rounds = batch.rounds
    del batch
    torch.cuda.empty_cache()
    print("batch launch done", flush=True)
    return cfgs, plans, packed, planes, rounds


@pytest.mark.parametrize("i", range(NINST))
def test_instance(batch_run, i):
    import torch

    cfgs, plans, packed, planes, rounds = batch_run
    cfg, dms = cfgs[i]
    plan, plane = plans[i], planes[i]
    print(f"instance {i} (round {rounds[i]}): {cfg.nbits}-bit {cfg.nchans} ch x {cfg.nsamps}, {len(dms)} DMs "
          f"to {dms[-1]:.0f}", flush=True)
    # the plain one-plan form on the same stream
    plan.allocate()
    plain = plan.execute(packed[i])
    assert torch.equal(plain, plane)
    del plain
    plan.release()
    x = _unpack(packed[i], cfg.nbits, torch).reshape(cfg.nsamps, cfg.nchans)
    expect = _row_sums(x, plan.table, plan.row_len, torch)
    L = torch.from_numpy(plan.row_len).cuda()
    cols = torch.arange(plane.shape[1], device="cuda")[None, :]
    for r0 in range(0, plan.ndm, 256):  # row blocks keep the temporaries small
        blk = plane[r0 : r0 + 256]
        valid = cols < L[r0 : r0 + 256, None]
        got = torch.where(valid, blk, torch.zeros((), device="cuda")).double().sum(dim=1)
        assert torch.equal(got.to(torch.int64), expect[r0 : r0 + 256])
        assert float(torch.where(valid, torch.zeros((), device="cuda"), blk).abs().max()) == 0.0
    (p,) = cfg.pulses
    k = int(np.argmin(np.abs(dms - p.dm)))
    row = plane[k, : int(plan.row_len[k])]
    assert p.t <= int(row.argmax()) < p.t + p.width
    if i % 4 == 0:
        r = int(len(dms) - 1 - i % 7)
        ref = c_oracle.sweep(x.cpu().numpy(), cfg.nchans, plan.table[[r]], ld=plane.shape[1])
        n = int(plan.row_len[r])
        np.testing.assert_array_equal(plane[r, :n].cpu().numpy(), ref[0, :n])
    del x
    torch.cuda.empty_cache()
