"""GPU tests of the streaming service (continuous batching under an HBM budget).

Several streams with their own DM lists, bit depths and gulp lengths run through one
server, their gulps interleaved and queued behind a budget that holds only a few gulps at a
time. Every gulp's new output samples must equal, byte for byte, (a) the oracle's
dedispersion of the whole stream at the same absolute times and (b) the same stream run
alone through a fresh server. The reference's streaming semantics (read_plan's skip-back of
the largest delay, readers.py:207-233, base.py:402-411) fix what each gulp must produce.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_oracle
from oracle import spp_oracle as O

pytestmark = pytest.mark.gpu

# (nbits, nchans, ndm, max_dm, gulp sizes)
STREAMS = [
    (8, 256, 24, 80.0, [9000, 3001, 6500, 4000]),
    (2, 512, 40, 150.0, [12000, 5000, 5000]),
    (4, 128, 12, 40.0, [4000, 2500, 7000, 3000, 3000]),
    (8, 384, 60, 120.0, [15000, 2000, 8000]),
    (1, 1024, 17, 300.0, [20000, 9000]),
]


def _stream(k, nbits, nchans, ndm, max_dm, sizes):
    rng = np.random.default_rng(100 + k)
    order = O.DEFAULT_BITORDER.get(nbits, "big")
    dms = np.linspace(0, max_dm, ndm)
    f = O.chan_freqs(nchans, -300.0 / nchans, 1550.0)
    table = np.atleast_2d(O.compute_dmdelays(f, dms, 64e-6, 1550.0)).astype(np.int32)
    total = sum(sizes)
    x = rng.integers(0, 1 << nbits, (total, nchans), dtype=np.uint8)
    packed = x.reshape(-1) if nbits == 8 else O.pack(x.reshape(-1), nbits, order)
    return table, x, packed, 1 if order == "big" else 0


@pytest.fixture(scope="module")
def streams():
    return [_stream(k, *s) for k, s in enumerate(STREAMS)]


def _gulps(packed, nchans, nbits, sizes, torch):
    rb = nchans * nbits // 8
    out, start = [], 0
    for n in sizes:
        out.append(torch.from_numpy(np.ascontiguousarray(packed[start * rb : (start + n) * rb])).pin_memory())
        start += n
    return out


def _run(server, picks, streams, order, torch):
    """Open the picked streams, submit their gulps in the given interleaved order and return
    {stream: [rows of each gulp]} (host copies of the new samples)."""
    handles = {}
    for k in picks:
        nbits, nchans, _, _, sizes = STREAMS[k]
        table, _, _, big = streams[k]
        handles[k] = server.open_stream(table, nbits, big, max(sizes))
    srcs = {k: _gulps(streams[k][2], STREAMS[k][1], STREAMS[k][0], STREAMS[k][4], torch) for k in picks}
    nxt = {k: 0 for k in picks}
    reqs = []
    for k in order:
        if k not in picks:
            continue
        g = nxt[k]
        reqs.append((k, g, server.submit(handles[k], srcs[k][g], STREAMS[k][4][g])))
        nxt[k] += 1
    out = {k: [None] * len(STREAMS[k][4]) for k in picks}
    for k, g, r in reqs:
        out[k][g] = r.rows()
        r.release()
    for h in handles.values():
        server.close_stream(h)
    return out


def _need(table, nbits, nchans, n_new, first):
    """Working set of one gulp as the server counts it (input + stage-1 copy + plane)."""
    al = lambda b: (int(b) + 255) // 256 * 256  # noqa: E731
    md = table.max(axis=1)
    n = n_new + (0 if first else int(md.max()))
    return al(n * nchans * nbits // 8) + al(4 * nchans * ((n + 63) // 64 * 64)) + al(4 * table.shape[0] * (n - int(md.min())))


def _order(seed):
    rng = np.random.default_rng(seed)
    seq = [k for k, s in enumerate(STREAMS) for _ in s[4]]
    rng.shuffle(seq)
    return seq


def test_streams_alone_and_in_the_mix_match_the_oracle(streams):
    import torch

    from sigpyproc3_amd.serve import Server

    order = _order(7)
    # resident tables and overlaps plus 1.5 of the largest gulp's working set: at most one
    # large gulp is on the device at a time and the others queue behind it
    big_need = max(_need(streams[k][0], STREAMS[k][0], STREAMS[k][1], n, g == 0)
                   for k in range(len(STREAMS)) for g, n in enumerate(STREAMS[k][4]))
    budget = int(1.5 * big_need) + (16 << 20)
    with Server(budget) as srv:
        mixed = _run(srv, list(range(len(STREAMS))), streams, order, torch)
        print(f"mix: admitted {srv.admitted}, peak {srv.peak_used} of {srv.arena.total} bytes", flush=True)
        assert srv.peak_used <= srv.arena.total
        assert srv.used == 0
    for k, (nbits, nchans, _, _, sizes) in enumerate(STREAMS):
        table, x, _, _ = streams[k]
        md = table.max(axis=1).astype(np.int64)
        total = sum(sizes)
        ref = c_oracle.sweep(x.reshape(-1), nchans, table, ld=total)
        with Server(budget) as srv:
            alone = _run(srv, [k], streams, _order(7), torch)[k]
        start = 0
        for g, n in enumerate(sizes):
            for i in range(table.shape[0]):
                got = mixed[k][g][i]
                t_abs = 0 if g == 0 else start - md[i]
                np.testing.assert_array_equal(got, ref[i, t_abs : t_abs + got.size],
                                              err_msg=f"stream {k} gulp {g} row {i} vs oracle")
                assert got.size == (n - md[i] if g == 0 else n)
                assert got.tobytes() == alone[g][i].tobytes(), f"stream {k} gulp {g} row {i}: alone != mix"
            start += n
        print(f"stream {k}: {len(sizes)} gulps equal alone, in the mix and to the oracle", flush=True)


def test_budget_queues_instead_of_failing(streams):
    import torch

    from sigpyproc3_amd.serve import QueueFull, Server

    nbits, nchans, _, _, sizes = STREAMS[0]
    table, _, packed, big = streams[0]
    srcs = _gulps(packed, nchans, nbits, sizes, torch)
    one = max(_need(table, nbits, nchans, n, g == 0) for g, n in enumerate(sizes))
    with Server(budget_bytes=int(1.2 * one) + (4 << 20), max_queued=2) as srv:
        h = srv.open_stream(table, nbits, big, max(sizes))
        r0 = srv.submit(h, srcs[0], sizes[0])
        r1 = srv.submit(h, srcs[1], sizes[1])
        # at most max_queued gulps wait: a third submission while both are queued is refused
        # or accepted once the first was admitted; either way nothing fails on the device
        try:
            r2 = srv.submit(h, srcs[2], sizes[2])
        except QueueFull:
            r0.wait()
            r2 = srv.submit(h, srcs[2], sizes[2])
        for r in (r0, r1, r2):
            r.rows()
            r.release()
        srv.close_stream(h)
        assert srv.used == 0
    with Server(budget_bytes=1 << 20) as srv, pytest.raises((ValueError, RuntimeError), match="budget"):
        h = srv.open_stream(table, nbits, big, max(sizes))
        srv.submit(h, srcs[0], sizes[0])
