"""The 8-bit, 4096-channel, 2^19-sample, 1024-DM sweep followed by the matched filter.

Full size: DM 0..1023 step 1 (max delay 14 834 samples, 509 454+ valid samples per row),
running mean/std normalisation over 4096 samples, boxcar widths {1, 2, ..., 256}, argmax
per DM row. Checks:
  * every row of the sweep: the exact prefix-sum checksum (independent torch integer ops);
  * sampled rows bit-exact against the C oracle;
  * the normalised series against the float64 oracle restatement (rtol 1e-6, atol 1e-5);
  * per-row (snr, peak, width) against the float64 closed-form boxcar MF of the same z:
    snr within rel 1e-5 + abs sqrt(w)*1e-5, (peak, width) exact;
  * the injected DM-500 pulse is the strongest detection, at its time and DM.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_oracle
from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu

WIDTHS = (1, 2, 4, 8, 16, 32, 64, 128, 256)
ROWS = (0, 1, 499, 500, 501, 777, 1023)


@pytest.fixture(scope="module")
def sweep_8bit_1024dm():
    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.sweep import SweepPlan, boxcar_mf, running_normalise
    from sigpyproc3_amd.synth import cfg_sweep_mf, generate_packed

    cfg = cfg_sweep_mf()
    dms = np.arange(1024, dtype=np.float64)
    table = cfg.delays(dms).astype(np.int32)
    packed = generate_packed(cfg, device=_lib.device())
    plan = SweepPlan(table, cfg.nsamps, cfg.nbits, 0)
    plane = plan.execute(packed)
    z = running_normalise(plane, plan.row_len, 4096)
    snr, peak, widx = boxcar_mf(z, plan.row_len, WIDTHS)
    torch.cuda.synchronize()
    yield cfg, table, packed, plan, plane, z, (snr, peak, widx)
    del z, plane, plan, packed
    torch.cuda.empty_cache()


def test_sweep_every_row_checksum(sweep_8bit_1024dm):
    import torch

    cfg, table, packed, plan, plane, _, _ = sweep_8bit_1024dm
    n, nch = cfg.nsamps, cfg.nchans
    x = packed.reshape(n, nch)
    d = torch.from_numpy(table.astype(np.int64)).cuda()
    L = torch.from_numpy(plan.row_len).cuda()
    expect = torch.zeros(table.shape[0], dtype=torch.int64, device="cuda")
    for cb in range(0, nch, 256):
        pre = torch.zeros((n + 1, 256), dtype=torch.int64, device="cuda")
        pre[1:] = torch.cumsum(x[:, cb : cb + 256].to(torch.int64), dim=0)
        dd = d[:, cb : cb + 256]
        expect += (torch.gather(pre, 0, (dd + L[:, None]).clamp(max=n)) - torch.gather(pre, 0, dd)).sum(dim=1)
        del pre
    rows = torch.arange(plane.shape[1], device="cuda")[None, :] < L[:, None]
    got = torch.where(rows, plane, torch.zeros((), device="cuda")).double().sum(dim=1)
    assert torch.equal(got.to(torch.int64), expect)
    assert int(plan.row_len.min()) == n - int(table.max())


def test_sweep_rows_vs_c_oracle(sweep_8bit_1024dm):
    cfg, table, packed, plan, plane, _, _ = sweep_8bit_1024dm
    x = packed.cpu().numpy()
    rows = list(ROWS)
    ref = c_oracle.sweep(x, cfg.nchans, table[rows], ld=plane.shape[1])
    for k, i in enumerate(rows):
        n = int(plan.row_len[i])
        np.testing.assert_array_equal(plane[i, :n].cpu().numpy(), ref[k, :n], err_msg=f"row {i}")


def test_normalise_and_mf_vs_fp64_oracle(sweep_8bit_1024dm):
    _, _, _, plan, plane, z, (snr, peak, widx) = sweep_8bit_1024dm
    worst = 0.0
    for i in ROWS:
        n = int(plan.row_len[i])
        x = plane[i, :n].cpu().numpy()
        zg = z[i, :n].cpu().numpy()
        np.testing.assert_allclose(zg, O.running_normalise(x, 4096), rtol=1e-6, atol=1e-5,
                                   err_msg=f"normalise row {i}")
        conv = O.boxcar_conv_closed(zg, WIDTHS).astype(np.float32)
        it, pk, v = O.mf_argmax(conv)
        assert (int(widx[i]), int(peak[i])) == (it, pk), f"row {i}"
        tol = 1e-5 * abs(v) + 1e-5 * np.sqrt(WIDTHS[it])
        assert abs(float(snr[i]) - float(v)) <= tol, (i, snr[i], v)
        worst = max(worst, abs(float(snr[i]) - float(v)) / max(abs(float(v)), 1e-30))
    print(f"max relative snr difference vs float64 oracle: {worst:.3e}")


def test_injected_pulse_is_best_detection(sweep_8bit_1024dm):
    cfg, _, _, plan, _, _, (snr, peak, widx) = sweep_8bit_1024dm
    (p,) = cfg.pulses
    best = int(np.argmax(snr))
    assert abs(best - int(p.dm)) <= 1
    assert p.t - WIDTHS[widx[best]] <= peak[best] <= p.t + p.width
    assert snr[int(p.dm)] > 50
