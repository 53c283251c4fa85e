"""The 2-bit 4096-channel 2^20-sample 2048-DM sweep at full size.

At this size the oracle cannot recompute every row, so the checks are:
  * an exact size-independent invariant for EVERY row: the sum of row i equals
    sum_c (P_c[d_ic + L_i] - P_c[d_ic]) with P_c the per-channel prefix sums of the
    unpacked stream (computed independently with torch integer ops);
  * bit-exact rows against the C oracle and against the one-DM-per-launch kernel;
  * the three injected pulses are found at their DM rows and times.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_oracle

pytestmark = pytest.mark.gpu


def _unpack_2bit_big(packed, torch):
    p = packed.to(torch.int32)
    parts = [(p >> s) & 3 for s in (6, 4, 2, 0)]
    return torch.stack(parts, dim=1).reshape(-1).to(torch.uint8)


@pytest.fixture(scope="module")
def sweep_2bit_2048dm():
    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.sweep import SweepPlan
    from sigpyproc3_amd.synth import cfg_sweep_2bit, generate_packed

    cfg = cfg_sweep_2bit()
    dms = np.arange(2048, dtype=np.float64)
    table = cfg.delays(dms).astype(np.int32)
    packed = generate_packed(cfg, device=_lib.device())
    plan = SweepPlan(table, cfg.nsamps, cfg.nbits, 1)
    plane = plan.execute(packed)
    torch.cuda.synchronize()
    yield cfg, table, packed, plan, plane
    del plane, plan, packed
    torch.cuda.empty_cache()


def test_every_row_checksum(sweep_2bit_2048dm):
    import torch

    cfg, table, packed, plan, plane = sweep_2bit_2048dm
    n, nch = cfg.nsamps, cfg.nchans
    x = _unpack_2bit_big(packed, torch).reshape(n, nch)
    d = torch.from_numpy(table.astype(np.int64)).cuda()            # [ndm, nch]
    L = torch.from_numpy(plan.row_len).cuda()                        # [ndm]
    expect = torch.zeros(table.shape[0], dtype=torch.int64, device="cuda")
    for cb in range(0, nch, 256):
        blk = x[:, cb : cb + 256].to(torch.int32)
        pre = torch.zeros((n + 1, blk.shape[1]), dtype=torch.int64, device="cuda")
        pre[1:] = torch.cumsum(blk, dim=0)
        dd = d[:, cb : cb + 256]
        hi = torch.gather(pre, 0, (dd + L[:, None]).clamp(max=n))
        lo = torch.gather(pre, 0, dd)
        expect += (hi - lo).sum(dim=1)
        del blk, pre
    rows = torch.arange(plane.shape[1], device="cuda")[None, :] < L[:, None]
    got = torch.where(rows, plane, torch.zeros((), device="cuda")).double().sum(dim=1)
    assert torch.equal(got.to(torch.int64), expect)
    tail = torch.where(rows, torch.zeros((), device="cuda"), plane)
    assert float(tail.abs().max()) == 0.0
    assert int(plan.total_samples) == 2117087959


def test_rows_vs_c_oracle_and_single_dm_kernel(sweep_2bit_2048dm):
    import torch

    from sigpyproc3_amd import _lib

    cfg, table, packed, plan, plane = sweep_2bit_2048dm
    x = _unpack_2bit_big(packed, torch).cpu().numpy()
    rows = [0, 1, 150, 1700, 2047]
    ref = c_oracle.sweep(x, cfg.nchans, table[rows], ld=plane.shape[1])
    for k, i in enumerate(rows):
        n = int(plan.row_len[i])
        np.testing.assert_array_equal(plane[i, :n].cpu().numpy(), ref[k, :n], err_msg=f"row {i}")
    # one DM per launch over the whole stream (one gulp)
    for i in (150, 1700):
        out = torch.zeros(cfg.nsamps, dtype=torch.float32, device="cuda")
        d = _lib.to_device(table[i])
        _lib.call("spp_dedisperse", _lib.ptr(packed), 2, 1, cfg.nchans, cfg.nsamps, _lib.ptr(d),
                  int(table[i].max()), _lib.ptr(out), 0, _lib.stream_handle())
        n = int(plan.row_len[i])
        assert torch.equal(out[:n], plane[i, :n])


def test_injected_pulses_found(sweep_2bit_2048dm):
    cfg, table, packed, plan, plane = sweep_2bit_2048dm
    for p in cfg.pulses:
        i = int(p.dm)
        row = plane[i, : int(plan.row_len[i])]
        t = int(row.argmax())
        assert p.t <= t < p.t + p.width, (p, t)
