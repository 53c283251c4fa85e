#!/usr/bin/env python3
"""Benchmark: the 2-bit, 4096-channel, 2^20-sample, 2048-DM (0..2047) dedispersion sweep.

One step = one full sweep of the packed stream resident in HBM (stage 1 + stage 2 of
``spp_sweep_execute``). Prints one JSON line (rank 0): metric, value, unit, timing and config
plus ``roofline`` (the dominant kernel, k_sweep, timed with HIP events on its stream)
and ``cpu_baseline`` (the C/OpenMP restatement of kernels.dedisperse_par, timed on a
bounded DM sample on this host and extrapolated per DM-trial sample).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "DM-trial·samples/s, 4096ch×2048-DM sweep (bit-exact); p50/p99 gulp latency"
UNIT = "DM-trial·samples/s"
HBM_PEAK_GBS = 8000.0
# Published MI355X vector rate: 157.3 TFLOP/s FP32, where an FMA counts 2 and packed FP32
# another 2; one plain 32-bit lane op per lane and cycle is a quarter of it.
VALU_LANE_OPS_PEAK_T = 157.3 / 4


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_counters(path: Path) -> dict:
    """Per-dispatch counters of k_sweep from the committed rocprofv3 PMC summary (key: value)."""
    vals = {}
    if path.exists():
        for line in path.read_text().splitlines():
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                try:
                    vals[k.strip()] = float(v.strip())
                except ValueError:
                    pass
    return vals


def lds_probe() -> dict:
    """LDS read rates of this GPU (k_probe_lds, HIP events on its stream): TB/s of
    ds_read_b128 (the best LDS read rate) and of ds_read2st64_b32 (the sweep's pattern)."""
    import torch

    from sigpyproc3_amd import _lib

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    blocks, iters = cus * 4, 20000
    out = _lib.empty((blocks * 512,), "int32")
    res = {}
    for pattern, width, name in ((0, 16, "ds_read_b128"), (1, 8, "ds_read2st64_b32")):
        best = None
        for _ in range(4):  # the first launch warms up
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            _lib.call("spp_probe_lds", pattern, iters, blocks, _lib.ptr(out), _lib.stream_handle())
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        nbytes = blocks * 512 * iters * 8 * width
        res[name] = round(nbytes / (best * 1e-3) / 1e12, 2)
    return res


def plain_one_dm(packed, cfg, table, row_len, plane, ndm_sample: int = 32) -> dict:
    """The plain one-DM-per-launch GPU form: spp_dedisperse (the Filterbank.dedisperse kernel)
    once per DM over the whole resident stream, on an evenly spaced DM sample, extrapolated per
    DM-trial sample. Its last row is compared with the sweep's plane."""
    import torch

    from sigpyproc3_amd import _lib

    picks = np.linspace(0, len(table) - 1, ndm_sample).astype(int)
    dev = _lib.device()
    rows = [torch.from_numpy(np.ascontiguousarray(table[i])).to(dev) for i in picks]
    mds = [int(table[i].max()) for i in picks]
    out = _lib.empty((cfg.nsamps,), "float32")
    big = 1 if cfg.nbits in (2, 4) else 0

    def one(d, md):
        _lib.call("spp_dedisperse", _lib.ptr(packed), cfg.nbits, big, cfg.nchans, cfg.nsamps, _lib.ptr(d), md,
                  _lib.ptr(out), 0, _lib.stream_handle())

    one(rows[0], mds[0])  # warm
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for d, md in zip(rows, mds):
        one(d, md)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    last = int(picks[-1])
    n = int(row_len[last])
    exact = bool(torch.equal(out[:n], plane[last, :n]))
    samples = float(sum(int(row_len[i]) for i in picks))
    return {"value": samples / (ms * 1e-3), "unit": UNIT, "ms_per_dm": round(ms / ndm_sample, 3),
            "dm_sample": f"{ndm_sample} DM rows evenly spaced over the {len(table)}, one spp_dedisperse "
                         f"launch each over the whole {cfg.nsamps}-sample stream in HBM; extrapolated per "
                         "DM-trial sample",
            "last_row_bit_exact_vs_sweep": exact}


def cpu_model() -> dict:
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                info[k] = v.strip()
    except OSError:
        pass
    info["logical_cpus_in_affinity"] = len(os.sched_getaffinity(0))
    info["os_cpu_count"] = os.cpu_count()
    return info


def cpu_baseline(packed_dev, cfg, table, row_len, plane, budget_s: float):
    """Time the C/OpenMP restatement of dedisperse_par on a bounded DM sample.

    Per DM the reference's Filterbank.dedisperse re-reads and serially unpacks the whole
    file, then runs dedisperse_par; both costs are timed (the unpack on one thread on a
    64 MiB slice, scaled). The rows computed are also compared with the GPU plane.
    """
    from oracle import c_oracle

    packed = packed_dev.cpu().numpy()
    threads = c_oracle.max_threads()
    log(f"cpu baseline: unpacking {packed.size / 2**30:.2f} GiB on {threads} threads")
    stream = c_oracle.unpack(packed, cfg.nbits, "big")
    # serial unpack rate (numba's unpack kernels are single-threaded loops)
    c_oracle.set_threads(1)
    sl = packed[: 64 << 20]
    t0 = time.perf_counter()
    c_oracle.unpack(sl, cfg.nbits, "big")
    t_unpack = (time.perf_counter() - t0) * packed.size / sl.size
    c_oracle.set_threads(threads)
    nch = cfg.nchans
    out = np.zeros(cfg.nsamps, dtype=np.float32)
    picks = [0]
    t_rows = []
    t0 = time.perf_counter()
    c_oracle.dedisperse(stream, out, table[0], nch, cfg.nsamps, 0)
    t_rows.append(time.perf_counter() - t0)
    exact = bool(np.array_equal(out[: row_len[0]], plane[0, : row_len[0]].cpu().numpy()))
    per_dm = t_rows[0] + t_unpack
    extra = max(0, min(7, int(budget_s / per_dm) - 1))
    for i in np.linspace(0, len(table) - 1, extra + 2)[1:].astype(int)[:extra]:
        t0 = time.perf_counter()
        c_oracle.dedisperse(stream, out, table[i], nch, cfg.nsamps, 0)
        t_rows.append(time.perf_counter() - t0)
        picks.append(int(i))
        n = int(row_len[i])
        exact &= bool(np.array_equal(out[:n], plane[i, :n].cpu().numpy()))
        log(f"cpu baseline: DM row {i} in {t_rows[-1]:.2f} s")
    samples = float(sum(row_len[i] for i in picks))
    t_compute = float(sum(t_rows))
    secs = t_compute + t_unpack * len(picks)
    return {
        "value": samples / secs,
        "unit": UNIT,
        "cores": threads,
        "kind": "port",
        "sample": (f"{len(picks)} DM rows {picks} of the 2048, each over the full 2^20-sample stream "
                   f"(C/OpenMP restatement of kernels.dedisperse_par, {threads} threads, "
                   f"{np.mean(t_rows):.2f} s/row) plus the serial unpack of the 1 GiB packed file per "
                   f"DM ({t_unpack:.2f} s, 1 thread, timed on a 64 MiB slice and scaled) as "
                   f"Filterbank.dedisperse incurs it; extrapolated per DM-trial sample"),
        "compute_only": {"value": samples / t_compute, "s_per_dm": round(t_compute / len(picks), 3),
                         "threads": threads},
        "unpack_only": {"s_per_dm": round(t_unpack, 3), "threads": 1,
                        "share_of_value": round(t_unpack * len(picks) / secs, 3)},
        "cpu": cpu_model(),
        "rows_bit_exact_vs_gpu": exact,
    }


def bench_serve(args) -> int:
    """16 streams (8-bit, 4096 ch), each with its own DM list of U[512, 4096] trials to DM
    U[500, 3000]; Poisson gulp arrivals of U[2^14, 2^18] new samples at 0.5/0.9/1.2 x the
    measured capacity, overlapped by each stream's largest delay, through the continuous-
    batching server under an HBM budget. Latency = arrival -> sweep complete, also as a
    fraction of the gulp's own duration; sampled rows are checked against the oracle."""
    import torch

    from oracle import c_oracle
    from sigpyproc3_amd import _lib, serve
    from sigpyproc3_amd.params import compute_dmdelays

    _lib.load()
    _lib.require_gpu()
    nstreams, lo, hi = 16, 1 << 14, 1 << 18
    nchans, fch1, foff, tsamp = 4096, 1550.0, -0.0732421875, 64e-6
    freqs = np.arange(nchans, dtype=np.float32) * np.float32(foff) + np.float32(fch1)
    rng = np.random.default_rng(5)
    budget = int(args.serve_budget_gib * 2**30)
    srv = serve.Server(budget, max_queued=args.serve_max_queued)
    handles, tables = [], []
    t_plan = time.perf_counter()
    for _s in range(nstreams):
        ndm = int(rng.integers(512, 4097))
        dms = np.linspace(0.0, float(rng.uniform(500.0, 3000.0)), ndm)
        table = np.atleast_2d(compute_dmdelays(freqs, dms, tsamp, fch1)).astype(np.int32)
        tables.append(table)
        handles.append(srv.open_stream(table, 8, 0, hi))
    t_plan = time.perf_counter() - t_plan
    source = serve.make_source(hi, nchans)
    src_host = source.numpy()
    history: list[list[int]] = [[] for _ in range(nstreams)]   # accepted gulp sizes per stream
    log(f"serve: {nstreams} streams, DM lists {[t.shape[0] for t in tables]}, overlaps "
        f"{[h.overlap for h in handles]}, plans {t_plan:.1f} s, budget {args.serve_budget_gib} GiB")
    # first gulp of every stream (no overlap) and one warm gulp each
    warm = [serve.Gulp(s, hi) for s in range(nstreams)] + [serve.Gulp(s, lo) for s in range(nstreams)]
    for g in warm:
        history[g.stream].append(g.nsamps)
    serve.run_arrivals(srv, handles, source, warm)

    def record(g):
        history[g.stream].append(g.nsamps)

    def samples_of(sched):
        return sum(tables[g.stream].shape[0] * g.nsamps for g in sched if not g.rejected)

    burst = serve.poisson_schedule(nstreams, args.serve_gulps, 0.0, lo, hi, seed=11)
    srv.max_queued = len(burst) + 1   # the capacity run queues every gulp
    wall = serve.run_arrivals(srv, handles, source, burst, on_submit=record)
    srv.max_queued = args.serve_max_queued
    assert not any(g.rejected for g in burst)
    capacity = len(burst) / wall
    cap_samples = samples_of(burst) / wall
    log(f"serve: capacity {capacity:.1f} gulps/s ({cap_samples:.3e} DM-trial samples/s) over {len(burst)} "
        f"gulps; peak budget use {srv.peak_used / 2**30:.1f} GiB")
    runs = {}
    checks = []
    for frac in (0.5, 0.9, 1.2):
        sched = serve.poisson_schedule(nstreams, args.serve_gulps, frac * capacity, lo, hi, seed=int(frac * 100))
        picks = set(range(5, len(sched), max(1, len(sched) // 3)))

        def sample(g, sched=sched, picks=picks):
            k = sched.index(g)
            if k in picks:
                i = int(k % tables[g.stream].shape[0])
                r = g.request
                row = r.plane[i, int(r.row_offset[i]) : int(r.row_offset[i]) + g.nsamps].cpu().numpy()
                checks.append((g.stream, r._hist_index, i, g.nsamps, row))

        def record_idx(g):
            g.request._hist_index = len(history[g.stream])
            history[g.stream].append(g.nsamps)

        wall = serve.run_arrivals(srv, handles, source, sched, on_done=sample, on_submit=record_idx)
        summ = serve.latency_summary(sched, tsamp)
        summ["offered_gulps_per_s"] = frac * capacity
        summ["dm_trial_samples_per_s"] = samples_of(sched) / wall
        runs[f"{frac}x"] = summ
        log(f"serve {frac}x: p50 {summ['p50_ms']:.1f} ms, p99 {summ['p99_ms']:.1f} ms, latency/duration p99 "
            f"{summ['latency_over_duration_p99']:.3f}, rejected {summ['rejected']}; queue p50/p99 "
            f"{summ['queue_ms_p50']:.1f}/{summ['queue_ms_p99']:.1f} ms, service p50/p99 "
            f"{summ['service_ms_p50']:.1f}/{summ['service_ms_p99']:.1f} ms")
    # sampled rows against the oracle: each gulp's input is the stream's last `overlap` samples
    # followed by the payload (the first n samples of the source)
    exact = True
    for s_, j, i, n, row in checks:
        ov = handles[s_].overlap
        need, parts = ov, []
        for m in reversed(history[s_][:j]):   # the last ov samples of the previous payloads
            take = min(m, need)
            parts.append(src_host[(m - take) * nchans : m * nchans])
            need -= take
            if need == 0:
                break
        inp = np.concatenate(parts[::-1] + [src_host[: n * nchans]])
        d = tables[s_][i]
        out = np.zeros(ov + n, dtype=np.float32)
        c_oracle.dedisperse(inp, out, d, nchans, ov + n, 0)
        md = int(d.max())
        exact &= bool(np.array_equal(row, out[ov - md : ov - md + n]))
    log(f"serve: {len(checks)} sampled rows bit-exact vs the oracle: {exact}")
    for h in handles:
        srv.close_stream(h)
    srv.shutdown()
    result = {
        "metric": "p50/p99 gulp latency at 0.5/0.9/1.2 x capacity (16-stream continuous-batching sweep service)",
        "value": runs["0.9x"]["dm_trial_samples_per_s"],
        "unit": UNIT,
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "16 streams x 8-bit x 4096 ch (1550 MHz, foff -0.0732, 64 us), gulps "
                               "U[2^14, 2^18] + overlap, per-stream DM lists U[512, 4096] to DM U[500, 3000]",
                   "dm_lists": [int(t.shape[0]) for t in tables], "overlaps": [h.overlap for h in handles],
                   "hbm_budget_bytes": budget, "max_queued": args.serve_max_queued},
        "capacity_gulps_per_s": capacity,
        "capacity_dm_trial_samples_per_s": cap_samples,
        "latency": runs,
        "peak_budget_bytes": srv.peak_used,
        "sampled_rows_bit_exact_vs_oracle": exact,
        "sampled_rows": len(checks),
    }
    print(json.dumps(result), flush=True)
    return 0 if exact else 1


def bench_batch(args) -> int:
    """The 64 uneven instances (1/2/4/8-bit, 336/1024/4096 ch, 2^16..2^19 samples, 256..4096
    DMs) in ONE persistent launch per step, against the sum of per-instance one-plan sweeps."""
    import torch

    from sigpyproc3_amd import _lib, tune
    from sigpyproc3_amd.batch import SweepBatch
    from sigpyproc3_amd.sweep import SweepPlan
    from sigpyproc3_amd.synth import cfg_batch_instance, generate_packed

    _lib.load()
    _lib.require_gpu()
    dev = _lib.device()
    plans, packed, desc = [], [], []
    t0 = time.perf_counter()
    for i in range(64):
        cfg, dms = cfg_batch_instance(i)
        table = cfg.delays(dms).astype(np.int32)
        plans.append(SweepPlan(table, cfg.nsamps, cfg.nbits, 1 if cfg.nbits in (2, 4) else 0, allocate=False))
        packed.append(generate_packed(cfg, device=dev))
        desc.append([cfg.nbits, cfg.nchans, cfg.nsamps, len(dms)])
        if i % 16 == 15:
            log(f"batch: {i + 1} instances planned and generated ({time.perf_counter() - t0:.1f} s)")
    planes = [_lib.empty((p.ndm, p.max_len), "float32") for p in plans]
    torch.cuda.synchronize()
    free, _ = torch.cuda.mem_get_info()
    g_total = sum(4 * p.nchans * ((p.nsamps + 63) // 64 * 64) for p in plans)
    budget = min(g_total, free - (12 << 30))
    batch = SweepBatch(plans, g_budget=budget)
    batch.bind(packed, planes)
    log(f"batch: {batch.info.rounds} rounds of {batch.info.pool_bytes / 2**30:.1f} GiB G pools, "
        f"{batch.info.build_items} stage-1 and {batch.info.sweep_items} stage-2 items")
    tune_ms = None
    if args.variant >= 0:
        batch.set_shape(args.variant, 1)
    else:
        batch.tune(repeats=1, cache=False)
        tune_ms = tune.last_times()
        log(f"batch autotune: shape {batch.shape}; ms {tune_ms}")
    for _ in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record()
        batch.run()
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    errors = batch.errors()
    launch_ms = np.array([a.elapsed_time(b) for a, b in ev])
    total = batch.total_samples
    sec_per_step = wall / args.steps
    value = total / sec_per_step
    info = batch.info
    shape = batch.shape
    del batch
    torch.cuda.empty_cache()
    log(f"batch: {sec_per_step * 1e3:.1f} ms per launch, {value:.3e} DM-trial samples/s; errors {errors}")
    # The plain form: each instance's own one-plan sweep (stage 1 + stage 2 launches).
    plain_ms, exact = [], True
    for k, (p, x, pl) in enumerate(zip(plans, packed, planes)):
        p.allocate()
        out = p.execute(x)  # warm (and the comparison)
        exact &= bool(torch.equal(out, pl))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        p.execute(x, out)
        e1.record()
        e1.synchronize()
        plain_ms.append(e0.elapsed_time(e1))
        del out
        p.release()
        if k % 16 == 15:
            torch.cuda.empty_cache()
    plain_s = float(sum(plain_ms)) * 1e-3
    log(f"plain per-instance sweeps: {plain_s * 1e3:.1f} ms in total; planes bit-exact: {exact}")
    result = {
        "metric": "DM-trial·samples/s, 64 uneven filterbanks in one persistent launch (bit-exact)",
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "64 filterbanks: nchans {336, 1024, 4096}, nsamps U[2^16, 2^19], nbits {1, 2, 4, 8}, "
                               "tsamp {64, 256} us, 1250-1550 MHz, DM lists of U[256, 4096] trials to DM "
                               "U[500, 3000]; float32 DM x time planes (all resident)",
                   "instances": desc, "dm_trial_samples": total,
                   "g_budget_bytes": int(budget), "rounds": int(info.rounds),
                   "pool_bytes": int(info.pool_bytes), "stage1_items": int(info.build_items),
                   "stage2_items": int(info.sweep_items)},
        "batch_shape": {"shape": shape[0], "time_tiles_per_item": shape[1],
                        "dm_rows_per_item": int(info.dm_rows_per_item),
                        "chans_per_stage": int(info.chans_per_stage)},
        "tune_ms": tune_ms,
        "launch_ms": {"mean": round(float(launch_ms.mean()), 3), "min": round(float(launch_ms.min()), 3)},
        "queue_errors": errors,
        "plain_per_instance": {"value": total / plain_s, "total_ms": round(plain_s * 1e3, 2),
                               "note": "each instance's own SweepPlan.execute (stage 1 + stage 2, default "
                                       "block shape), timed one after another with HIP events; planes "
                                       "compared with the batch's"},
        "speedup_vs_plain_per_instance": value / (total / plain_s),
        "planes_bit_exact_vs_plain": exact,
    }
    print(json.dumps(result), flush=True)
    return 0 if exact and errors == 0 else 1


def bench_mf(args) -> int:
    """8-bit, 4096 ch, 2^19 samples, DM 0..1023: sweep, then running normalisation over 4096
    samples and the boxcar matched filter (widths 1..256) with per-DM argmax in one pass over
    the plane. One step = both. The two-pass form (normalised plane written, then read by the
    MF) is timed beside it."""
    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.sweep import DEFAULT_WIDTHS, SweepPlan, boxcar_mf, normalise_boxcar_mf, running_normalise
    from sigpyproc3_amd.synth import cfg_sweep_mf, generate_packed

    _lib.load()
    _lib.require_gpu()
    cfg = cfg_sweep_mf()
    table = cfg.delays(np.arange(1024, dtype=np.float64)).astype(np.int32)
    packed = generate_packed(cfg, device=_lib.device())
    plan = SweepPlan(table, cfg.nsamps, cfg.nbits, 0)
    plane = _lib.empty((plan.ndm, plan.max_len), "float32")
    if args.variant >= 0:
        plan.variant = args.variant
    else:
        from sigpyproc3_amd import tune

        plan.tune(packed, plane, cache=False)
        log(f"mf: autotune: variant {plan.variant}; k_sweep ms per variant {tune.last_times()}")
    for _ in range(args.warmup):
        plan.execute(packed, plane)
        normalise_boxcar_mf(plane, plan.row_len, 4096, DEFAULT_WIDTHS)
    torch.cuda.synchronize()
    t_sweep, t_mf = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        plan.execute(packed, plane)
        torch.cuda.synchronize()
        b = time.perf_counter()
        snr, peak, widx = normalise_boxcar_mf(plane, plan.row_len, 4096, DEFAULT_WIDTHS)  # synchronises
        c = time.perf_counter()
        t_sweep.append(b - a)
        t_mf.append(c - b)
    wall = (time.perf_counter() - t0) / args.steps
    # the two-pass form on the same plane: same detections, z through HBM
    t_two = []
    for _ in range(max(2, args.steps // 4)):
        c = time.perf_counter()
        z = running_normalise(plane, plan.row_len, 4096)
        snr2, peak2, widx2 = boxcar_mf(z, plan.row_len, DEFAULT_WIDTHS)
        t_two.append(time.perf_counter() - c)
        del z
    same = bool(np.array_equal(peak, peak2) and np.array_equal(widx, widx2)
                and np.allclose(snr, snr2, rtol=2e-6, atol=1e-6))
    total = plan.total_samples
    best = int(np.argmax(snr))
    result = {
        "metric": "DM-trial samples/s through sweep + running normalisation + boxcar MF",
        "value": total / wall,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "u32 sweep, f64 normalisation/MF sums, f32 out",
        "data": "synthetic",
        "config": {"workload": "8-bit, 4096 ch x 2^19 samples (1550 MHz, foff -0.0732, 64 us), DM 0..1023 "
                               "step 1, window 4096, boxcar widths 1..256",
                   "dm_trial_samples": total, "sweep_variant": plan.variant},
        "stages_ms": {"sweep": 1e3 * float(np.mean(t_sweep)),
                      "normalise_boxcar_mf": 1e3 * float(np.mean(t_mf)),
                      "two_pass_normalise_then_boxcar_mf": 1e3 * float(np.mean(t_two))},
        "one_pass_matches_two_pass": same,
        "best_detection": {"dm": float(best), "snr": float(snr[best]), "peak": int(peak[best]),
                           "width": int(DEFAULT_WIDTHS[int(widx[best])])},
    }
    log(f"mf: sweep {result['stages_ms']['sweep']:.2f} ms, one-pass normalise+mf "
        f"{result['stages_ms']['normalise_boxcar_mf']:.2f} ms (two-pass "
        f"{result['stages_ms']['two_pass_normalise_then_boxcar_mf']:.2f} ms, same result: {same})")
    print(json.dumps(result), flush=True)
    return 0


def bench_single(args) -> int:
    """The one-DM path at 8-bit, 1024 ch, 65 536 samples, DM 300 (the gulp schedule of
    Filterbank.dedisperse: gulp 16384, skipback = max delay), with the gulps already in HBM;
    one 2-bit 4096-channel gulp of 2^18 samples (the headline band) at DM 0, 100 and 1000;
    and the 2-bit unpack of a 256 MiB packed buffer. Each is priced against HBM: (bytes read
    + bytes written) per launch / kernel time against 8 TB/s."""
    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.readers import plan_blocks
    from sigpyproc3_amd.synth import cfg_single_dm, generate_packed

    _lib.load()
    _lib.require_gpu()
    stream = _lib.stream_handle()
    cfg = cfg_single_dm()
    delays = cfg.delays([300.0])[0].astype(np.int32)
    md = int(delays.max())
    packed = generate_packed(cfg, device=_lib.device())
    d_dev = _lib.to_device(delays)
    gulp = max(2 * md, 16384)
    _, blocks = plan_blocks(cfg.nsamps, gulp, md)
    tim = _lib.zeros((cfg.nsamps - md,), "float32")
    starts, pos = [], 0
    for ii, n, skip in blocks:
        starts.append((ii, n, pos))
        pos += n + skip

    def one_pass(st=stream):
        for ii, n, p0 in starts:
            _lib.call("spp_dedisperse", _lib.ptr(packed[p0 * cfg.nchans:]), 8, 0, cfg.nchans, n,
                      _lib.ptr(d_dev), md, _lib.ptr(tim), ii * (gulp - md), st)

    for _ in range(max(1, args.warmup)):
        one_pass()
    torch.cuda.synchronize()
    reps = 50
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        one_pass()
    e1.record()
    torch.cuda.synchronize()
    t_pass = e0.elapsed_time(e1) / reps * 1e-3
    bytes_pass = sum(n * cfg.nchans for _, n, _ in starts) + 4 * (cfg.nsamps - md)
    row = tim.cpu().numpy()
    peak_t = int(np.argmax(row))
    # The stream resident in HBM as a whole: one launch over every output sample (what
    # sweep.dedisperse_resident does for non-negative delays), compared with the gulp pass.
    tim1 = _lib.zeros((cfg.nsamps - md,), "float32")

    def one_launch(st=stream):
        _lib.call("spp_dedisperse", _lib.ptr(packed), 8, 0, cfg.nchans, cfg.nsamps, _lib.ptr(d_dev), md,
                  _lib.ptr(tim1), 0, st)

    one_launch()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        one_launch()
    e1.record()
    torch.cuda.synchronize()
    t_one = e0.elapsed_time(e1) / reps * 1e-3
    one_equal = bool(np.array_equal(tim1.cpu().numpy(), row))
    bytes_one = cfg.nsamps * cfg.nchans + 4 * (cfg.nsamps - md)
    # The same pass captured once as a HIP graph and replayed: the gulp launches without the
    # host's per-call cost (the kernels allocate nothing and clear nothing).
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        one_pass(_lib.stream_handle())
    tim.zero_()
    graph.replay()
    torch.cuda.synchronize()
    graph_equal = bool(np.array_equal(tim.cpu().numpy(), row))
    e0.record()
    for _ in range(reps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    t_graph = e0.elapsed_time(e1) / reps * 1e-3
    # One 2-bit gulp of the 4096-channel band: 2^18 samples (256 MiB packed), one launch per DM.
    from sigpyproc3_amd.synth import cfg_sweep_2bit

    c2 = cfg_sweep_2bit()
    c2.nsamps = 1 << 18
    c2.pulses = []
    packed2 = generate_packed(c2, device=_lib.device())
    wide = {}
    for dm in (0.0, 100.0, 1000.0):
        d2 = c2.delays([dm])[0].astype(np.int32)
        md2 = int(d2.max())
        n_out = c2.nsamps - md2
        d2_dev = _lib.to_device(d2)
        out2 = _lib.empty((n_out,), "float32")

        def gulp2(d2_dev=d2_dev, md2=md2, out2=out2):
            _lib.call("spp_dedisperse", _lib.ptr(packed2), 2, 1, c2.nchans, c2.nsamps, _lib.ptr(d2_dev),
                      md2, _lib.ptr(out2), 0, stream)

        gulp2()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(20):
            gulp2()
        e1.record()
        torch.cuda.synchronize()
        t2 = e0.elapsed_time(e1) / 20 * 1e-3
        b2 = c2.nsamps * c2.nchans // 4 + 4 * n_out
        wide[f"dm{int(dm)}"] = {"us": t2 * 1e6, "max_delay": md2, "hbm_bytes": b2, "hbm_gbs": b2 / t2 / 1e9,
                                "hbm_frac": b2 / t2 / 8e12, "channel_adds_per_s": n_out * c2.nchans / t2}
    del packed2
    # unpack: 2-bit, 256 MiB packed -> 1 GiB
    nb = 256 << 20
    src = torch.randint(0, 256, (nb,), dtype=torch.uint8, device=_lib.device())
    dst = _lib.empty((nb * 4,), "uint8")
    _lib.call("spp_unpack", _lib.ptr(src), nb, _lib.ptr(dst), 2, 1, stream)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        _lib.call("spp_unpack", _lib.ptr(src), nb, _lib.ptr(dst), 2, 1, stream)
    e1.record()
    torch.cuda.synchronize()
    t_unpack = e0.elapsed_time(e1) / 10 * 1e-3
    result = {
        "metric": "one-DM dedispersion (gulp path) and 2-bit unpack, priced against HBM",
        "value": (cfg.nsamps - md) / t_pass,
        "unit": "output samples/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "8-bit, 1024 ch x 65536 samples, DM 300 (max delay %d), gulp %d, %d gulps"
                               % (md, gulp, len(starts))},
        "dedisperse": {"us_per_pass": t_pass * 1e6, "hbm_bytes_per_pass": bytes_pass,
                       "hbm_gbs": bytes_pass / t_pass / 1e9, "hbm_frac": bytes_pass / t_pass / 8e12,
                       "peak_sample": peak_t, "pulse_at": [20000, 20015],
                       "graph_us_per_pass": t_graph * 1e6, "graph_hbm_gbs": bytes_pass / t_graph / 1e9,
                       "graph_output_equal": graph_equal,
                       "resident_one_launch_us": t_one * 1e6, "resident_hbm_gbs": bytes_one / t_one / 1e9,
                       "resident_hbm_frac": bytes_one / t_one / 8e12, "resident_output_equal": one_equal},
        "dedisperse_2bit_4096ch_2p18": wide,
        "unpack_2bit": {"ms": t_unpack * 1e3, "bytes": nb * 5, "hbm_gbs": nb * 5 / t_unpack / 1e9,
                        "hbm_frac": nb * 5 / t_unpack / 8e12},
    }
    log(f"single: resident stream in one launch {t_one * 1e6:.1f} us (equal: {one_equal}); "
        f"{t_pass * 1e6:.1f} us per dedisperse pass ({result['dedisperse']['hbm_gbs']:.0f} GB/s), "
        f"{t_graph * 1e6:.1f} us replayed as a graph (output equal: {graph_equal}); 2-bit 4096 ch gulp "
        + ", ".join(f"{k} {v['us']:.0f} us" for k, v in wide.items()) + "; "
        f"unpack {result['unpack_2bit']['hbm_gbs']:.0f} GB/s")
    print(json.dumps(result), flush=True)
    return 0


def bench_file(args) -> int:
    """The 2-bit 4096-ch 2^20-sample stream as a SIGPROC file (1 GiB of data): one-DM
    Filterbank.dedisperse through the double-buffered gulp pipeline and through a single
    caller buffer, and Filterbank.dedisperse_sweep over DM 0..2047 straight from the file."""
    import tempfile

    import torch

    from sigpyproc3_amd import FilReader, _lib
    from sigpyproc3_amd.synth import cfg_sweep_2bit, generate_packed, write_sigproc

    _lib.load()
    _lib.require_gpu()
    cfg = cfg_sweep_2bit()
    with tempfile.TemporaryDirectory() as tmp:
        path = f"{tmp}/sweep_2bit.fil"
        write_sigproc(path, cfg, generate_packed(cfg, device=_lib.device()))
        torch.cuda.synchronize()
        fil = FilReader(path)
        nbytes = fil.header.nsamples * fil.samp_stride
        with open(path, "rb") as fp:        # page-cache warm, as a re-read of a recent file
            while fp.read(1 << 26):
                pass
        dm = 1000.0
        runs = {}
        for name, kw in (("pipeline", {}), ("single_buffer", {"allocator": bytearray})):
            fil.dedisperse(dm, gulp=65536, **kw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tim = fil.dedisperse(dm, gulp=65536, **kw)
            runs[name] = time.perf_counter() - t0
            runs[name + "_first_value"] = float(tim.data[0])
        dms = np.arange(2048, dtype=np.float64)
        fil.dedisperse_sweep(dms[:16])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = fil.dedisperse_sweep(dms)
        torch.cuda.synchronize()
        t_sweep = time.perf_counter() - t0
        total = int(res.row_len.sum())
        del res
    result = {
        "metric": "file-resident workloads: one-DM dedisperse and the 2048-DM sweep from a SIGPROC file",
        "value": total / t_sweep,
        "unit": UNIT,
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic SIGPROC file in the page cache",
        "config": {"workload": "2-bit, 4096 ch x 2^20 samples file (1 GiB of data)", "dm": dm, "gulp": 65536},
        "dedisperse_s": {"pipeline": runs["pipeline"], "single_buffer": runs["single_buffer"]},
        "dedisperse_file_gbs": {"pipeline": nbytes / runs["pipeline"] / 1e9,
                                "single_buffer": nbytes / runs["single_buffer"] / 1e9},
        "sweep_from_file_s": t_sweep,
        "outputs_equal": runs["pipeline_first_value"] == runs["single_buffer_first_value"],
    }
    log(f"file: dedisperse {runs['pipeline']:.3f} s (pipeline) / {runs['single_buffer']:.3f} s (single "
        f"buffer); sweep from file {t_sweep:.3f} s")
    print(json.dumps(result), flush=True)
    return 0


def bench_extras(args) -> int:
    """The gulp pre-steps, fold and block operations on resident data: a 2-bit, 4096-channel,
    2^18-sample stream (256 MiB packed) and a float32 block of 4096 channels x 16384 samples.
    Each kernel is timed with HIP events (median of 10 launches) and reported as HBM bytes
    per launch / time against 8 TB/s; the numpy oracle of the same operation is timed on a
    bounded slice on the host beside it (per byte of packed input)."""
    import time

    import torch

    from oracle import spp_oracle as O
    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.foldedcube import band_starts, fold_cells
    from sigpyproc3_amd.synth import SynthConfig, generate_packed

    _lib.load()
    _lib.require_gpu()
    st = _lib.stream_handle()
    cfg = SynthConfig(nchans=4096, nsamps=1 << 18, nbits=2, fch1=1550.0, foff=-0.0732421875,
                      tsamp=64e-6, seed=7)
    n, nch = cfg.nsamps, cfg.nchans
    row_bytes = nch * cfg.nbits // 8
    packed = generate_packed(cfg, device=_lib.device())
    nbytes = int(packed.numel())
    rng = np.random.default_rng(11)

    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return float(np.median(ts))

    # host slice for the oracle timings: the first 4096 samples, unpacked
    ns = 4096
    host_slice = O.unpack(packed[: ns * row_bytes].cpu().numpy(), 2, "big")
    slice_bytes = ns * row_bytes

    def cpu_rate(fn):
        t0 = time.perf_counter()
        fn()
        return slice_bytes / (time.perf_counter() - t0)

    lines = {}

    def record(name, sec, hbm_bytes, cpu_bps, note):
        lines[name] = {"ms": round(sec * 1e3, 4), "hbm_bytes": int(hbm_bytes),
                       "hbm_gbs": round(hbm_bytes / sec / 1e9, 1), "hbm_frac": round(hbm_bytes / sec / 8e12, 4),
                       "input_gbs": round(nbytes / sec / 1e9, 1),
                       "cpu_oracle_gbs": round(cpu_bps / 1e9, 4), "note": note}
        log(f"extras: {name}: {sec * 1e3:.3f} ms, {hbm_bytes / sec / 1e9:.0f} GB/s "
            f"({hbm_bytes / sec / 8e12:.3f} of 8 TB/s); oracle {cpu_bps / 1e9:.3f} GB/s of input")

    # collapse (extract_tim): per-sample channel sums
    tim = _lib.zeros((n,), "float32")
    t = timed(lambda: _lib.call("spp_collapse", _lib.ptr(packed), 2, 1, nch, n, _lib.ptr(tim), 0, st))
    record("collapse", t, nbytes + 4 * n,
           cpu_rate(lambda: O.collapse(host_slice, ns, nch)), "one gulp of 2^18 samples")
    # bandpass, float32 time order (extract_bpass) and per-gulp float64 (extract_bpass_par)
    bp = _lib.zeros((nch,), "float32")
    bp_ws = _lib.workspace(_lib.workspace_size("spp_bandpass_workspace_size", 2, nch, n))
    for par in (0, 1):
        t = timed(lambda: _lib.call("spp_bandpass", _lib.ptr(packed), 2, 1, nch, n, _lib.ptr(bp), par,
                                    _lib.ptr(bp_ws), bp_ws.numel(), st))
        record("bandpass_parallel" if par else "bandpass", t, nbytes + 4 * nch + 2 * bp_ws.numel(),
               cpu_rate(lambda: O.bandpass(host_slice, ns, nch, parallel=bool(par))),
               "per-channel sums over the gulp (HBM bytes include the chunk partials written and read)")
    # channel mask on the packed bytes (in place on a copy)
    mask = rng.random(nch) < 0.1
    keep = np.zeros(row_bytes, np.uint32)
    for k in range(4):
        keep |= np.where(mask[k::4], 0, 3 << (8 - 2 * (k + 1))).astype(np.uint32)
    keep_d = _lib.to_device(keep.astype(np.uint8))
    set_d = _lib.zeros((row_bytes,), "uint8")
    work = packed.clone()
    t = timed(lambda: _lib.call("spp_mask_bytes", _lib.ptr(work), nbytes, row_bytes, _lib.ptr(keep_d),
                                _lib.ptr(set_d), st))
    record("mask_channels", t, 2 * nbytes,
           cpu_rate(lambda: O.mask_channels(host_slice, nch, mask, 0)), "read + write the packed bytes")
    # zero-DM removal: unpacked uint8 out
    bpass = rng.uniform(1e5, 2e5, nch).astype(np.float32)
    w_d = _lib.to_device(bpass / bpass.sum())
    b_d = _lib.to_device(bpass)
    zout = _lib.empty((n * nch,), "uint8")
    t = timed(lambda: _lib.call("spp_remove_zerodm", _lib.ptr(packed), 2, 1, nch, n, _lib.ptr(w_d),
                                _lib.ptr(b_d), _lib.ptr(zout), st))
    zs = 64  # the oracle restates the reference's per-sample loop: a short slice
    t0 = time.perf_counter()
    O.remove_zerodm(host_slice[: zs * nch], nch, bpass, np.uint8)
    zerodm_cpu = zs * row_bytes / (time.perf_counter() - t0)
    record("remove_zerodm", t, nbytes + n * nch, zerodm_cpu,
           "packed in, one byte per sample out; the oracle timed on 64 samples")
    del zout, work
    # fold: period 0.5 s, 64 bins, 16 subints, 32 bands, DM 100; cell lists built on the host
    delays = cfg.delays([100.0])[0].astype(np.int32)
    md = int(delays.max())
    nvalid = n - md
    nbins, nints, nbands = 64, 16, 32
    subint, phasebin = fold_cells(0, nvalid, n, cfg.tsamp, 0.5, 0.0, nbins, nints)
    key = subint * nbins + phasebin
    order = np.argsort(key, kind="stable").astype(np.int32)
    keys, first = np.unique(key[order], return_index=True)
    key_start = np.append(first, nvalid).astype(np.int32)
    key_pos = ((keys // nbins) * nbins * nbands + keys % nbins).astype(np.int32)
    bufs = [_lib.to_device(x) for x in (delays, order, key_start, key_pos, band_starts(nch, nbands))]
    fold_d = _lib.zeros((nbins * nints * nbands,), "float32")
    count_d = _lib.zeros((nbins * nints * nbands,), "int32")
    fold_ws = _lib.workspace(_lib.workspace_size("spp_fold_workspace_size", 2, nch, n, nbands))

    def fold_once():
        fold_d.zero_()
        count_d.zero_()
        _lib.call("spp_fold", _lib.ptr(packed), 2, 1, nch, n, _lib.ptr(bufs[0]), _lib.ptr(bufs[1]), nvalid,
                  _lib.ptr(bufs[2]), _lib.ptr(bufs[3]), len(keys), _lib.ptr(bufs[4]), nbands, nbins,
                  _lib.ptr(fold_d), _lib.ptr(count_d), _lib.ptr(fold_ws), fold_ws.numel(), st)

    t = timed(fold_once)
    fs = 512
    fa, fc = np.zeros(nbins * nints * nbands, np.float32), np.zeros(nbins * nints * nbands, np.int32)
    t0 = time.perf_counter()
    O.fold_kernel(host_slice[: fs * nch], fa, fc, delays.astype(np.int32) * 0, 0, cfg.tsamp, 0.5, 0.0, n,
                  fs, nch, nbins, nints, nbands, 0)
    fold_cpu = fs * row_bytes / (time.perf_counter() - t0)
    record("fold", t, nbytes + 4 * order.size + 8 * fold_d.numel() + 8 * n * nbands, fold_cpu,
           "band sums of every sample (fused gulp kernel), then per (subint, band, bin) cell totals; "
           "the oracle is the reference's pure-Python loop on 512 samples")
    # FilterbankBlock: roll every channel (dedisperse) and a 64-DM DM-time transform
    blk_cols = 16384
    blk = torch.randint(0, 4, (nch, blk_cols), dtype=torch.int32, device=_lib.device()).float()
    bd = np.atleast_2d(cfg.delays([40.0])).astype(np.int64)[0]
    off_h = np.ascontiguousarray(np.mod(bd, blk_cols), dtype=np.int64)
    off_d = _lib.to_device(off_h)
    rolled = _lib.empty((nch, blk_cols), "float32")
    t = timed(lambda: _lib.call("spp_roll_rows", _lib.ptr(blk), nch, blk_cols, _lib.ptr(off_h), _lib.ptr(off_d),
                                blk_cols, _lib.ptr(rolled), st))
    hb = blk[:, :1024].cpu().numpy()
    t0 = time.perf_counter()
    O.roll_block(hb, -bd)
    roll_cpu = hb.nbytes / (time.perf_counter() - t0)
    record("block_dedisperse", t, 2 * blk.numel() * 4, roll_cpu,
           "float32 block 4096 x 16384 rolled per channel (block bytes in + out)")
    ndm = 64
    dmd = np.atleast_2d(cfg.delays(40.0 + np.linspace(-40.0, 40.0, ndm))).astype(np.int64)
    offs_h = np.ascontiguousarray(np.mod(-dmd, blk_cols), dtype=np.int64)
    offs_d = _lib.to_device(offs_h)
    dmt = _lib.empty((ndm, blk_cols), "float32")
    t = timed(lambda: _lib.call("spp_dmt", _lib.ptr(blk), nch, blk_cols, _lib.ptr(offs_h), _lib.ptr(offs_d), ndm,
                                blk_cols, _lib.ptr(dmt), st), reps=5)
    t0 = time.perf_counter()
    O.dmt_block(hb, dmd[:4])
    dmt_cpu = 4 * hb.nbytes / (time.perf_counter() - t0)
    record("block_dmt_64dm", t, ndm * blk.numel() * 4 + dmt.numel() * 4, dmt_cpu,
           "each of 64 DMs reads the whole block (float32 channel sums in channel order)")
    result = {
        "metric": "gulp pre-steps, fold and block operations: HBM GB/s per kernel",
        "value": lines["remove_zerodm"]["hbm_gbs"],
        "unit": "GB/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "u8/f32",
        "data": "synthetic",
        "config": {"workload": "2-bit 4096 ch x 2^18 samples (256 MiB packed) resident in HBM; float32 block "
                               "4096 x 16384; value = remove_zerodm's HBM rate"},
        "kernels": lines,
    }
    print(json.dumps(result), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="wall-clock budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", type=int, default=-1,
                    help="stage-2 block shape; -1 = autotune (time every shape, keep the fastest)")
    ap.add_argument("--copy-calibration-gib", type=float, default=0.0,
                    help="first run a plain device copy of this many GiB three times (the "
                         "reference dispatch for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--workload", choices=("sweep", "batch", "serve", "mf", "file", "single", "extras"),
                    default="sweep",
                    help="sweep: the 2-bit 2048-DM sweep (the headline line); serve: the "
                         "16-stream gulp service with p50/p99 latency; mf: the 8-bit 1024-DM "
                         "sweep + normalisation + boxcar matched filter; file: the same stream "
                         "read from a SIGPROC file; single: the one-DM gulp path and unpack; "
                         "batch: 64 uneven filterbanks in one persistent launch; "
                         "extras: the gulp pre-steps, fold and block operations")
    ap.add_argument("--serve-gulps", type=int, default=160, help="gulps per serve run")
    ap.add_argument("--serve-budget-gib", type=float, default=96.0, help="HBM budget of the serving run")
    ap.add_argument("--serve-max-queued", type=int, default=48,
                    help="queued gulps beyond which the server refuses new ones (backpressure)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              "(the sweep would shard by DM list, one process per GPU)", file=sys.stderr)
        return 1

    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "mf":
        return bench_mf(args)
    if args.workload == "file":
        return bench_file(args)
    if args.workload == "single":
        return bench_single(args)
    if args.workload == "extras":
        return bench_extras(args)

    import torch

    from sigpyproc3_amd import _lib
    from sigpyproc3_amd.params import numpy_simd_features, table_digest
    from sigpyproc3_amd.sweep import SweepPlan
    from sigpyproc3_amd.synth import cfg_sweep_2bit, generate_packed

    _lib.load()
    _lib.require_gpu()
    dev = _lib.device()
    if args.copy_calibration_gib > 0:
        nb = int(args.copy_calibration_gib * 2**30)
        src = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {nb} bytes read + written")
        del src, dst
    cfg = cfg_sweep_2bit()
    dms = np.arange(2048, dtype=np.float64)
    table = cfg.delays(dms).astype(np.int32)
    log(f"generating {cfg.nchans} ch x {cfg.nsamps} samples, {cfg.nbits}-bit, on {torch.cuda.get_device_name()}")
    packed = generate_packed(cfg, device=dev)
    plan = SweepPlan(table, cfg.nsamps, cfg.nbits, 1)
    plane = _lib.empty((plan.ndm, plan.max_len), "float32")
    torch.cuda.synchronize()
    log(f"plan: {plan.info.ndm_tiles} DM tiles, max spread {plan.info.max_spread}, "
        f"workspace {plan.info.workspace_bytes / 2**30:.1f} GiB")
    tune_ms = None
    if args.variant >= 0:
        plan.variant = args.variant
    else:
        from sigpyproc3_amd import tune

        plan.tune(packed, plane, cache=False)
        tune_ms = tune.last_times()
        log(f"autotune: variant {plan.variant}; k_sweep ms per variant {tune_ms}")
    for _ in range(args.warmup):
        plan.execute(packed, plane)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record()
        plan.build(packed)
        e1.record()
        plan.run(plane)
        e2.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    build_ms = np.array([a.elapsed_time(b) for a, b, _ in ev])
    sweep_ms = np.array([b.elapsed_time(c) for _, b, c in ev])
    step_ms = np.array([a.elapsed_time(c) for a, _, c in ev])
    total = plan.total_samples
    common = int(plan.info.min_len) * plan.ndm
    sec_per_step = wall / args.steps
    value = total / sec_per_step
    log(f"{args.steps} steps: {sec_per_step * 1e3:.2f} ms/step; k_sweep {sweep_ms.mean():.2f} ms, "
        f"k_build_g {build_ms.mean():.2f} ms")

    # Host-resident input: one pinned-memory copy of the packed stream to HBM per sweep.
    host = torch.empty(packed.numel(), dtype=torch.uint8, pin_memory=True)
    host.copy_(packed)
    h0, h1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    h0.record()
    packed.copy_(host, non_blocking=True)
    h1.record()
    torch.cuda.synchronize()
    h2d_ms = h0.elapsed_time(h1)
    del host

    adds = float(total) * cfg.nchans
    t_sweep = float(sweep_ms.mean()) * 1e-3
    counters = read_counters(ROOT / "profiles" / "k_sweep_2bit_4096ch_2048dm_traffic.txt")
    traffic = (int(counters["fetch_bytes_per_dispatch"] + counters["write_bytes_per_dispatch"])
               if "fetch_bytes_per_dispatch" in counters and "write_bytes_per_dispatch" in counters else None)
    # LDS bytes the interleaved layout reads per sample-add: one 32-bit word holds IL samples.
    il = 8 if cfg.nbits <= 2 else 4
    lds_bytes = adds * 4.0 / il
    lds_tbs = lds_bytes / t_sweep / 1e12
    probe = lds_probe()
    log(f"LDS probe: {probe} TB/s")
    hbm_alg = 4.0 * total + cfg.nsamps * cfg.nchans * cfg.nbits / 8
    roofline = {
        "bound": "lds",
        "achieved": round(lds_tbs, 2),
        "peak": probe["ds_read_b128"],
        "unit": "TB/s",
        "frac": round(lds_tbs / probe["ds_read_b128"], 4),
        "traffic": traffic,
        "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel and shape, scaled by a copy "
                           "calibration: profiles/k_sweep_2bit_4096ch_2048dm_traffic.txt (not measured in "
                           "this run)"),
        "kernel": "k_sweep",
        "kernel_ms": round(t_sweep * 1e3, 3),
        "algorithmic_work": (f"{total} DM-trial samples x {cfg.nchans} channel adds = {adds:.4e} adds per launch; "
                             f"LDS reads {lds_bytes:.4e} B (one 4-byte word per {il} sample-adds)"),
        "peak_source": ("k_probe_lds in this run: 32 KiB of LDS per block read with ds_read_b128 (16 aligned "
                        "bytes per lane), 4 blocks per CU"),
        "lds_pattern_peak": probe["ds_read2st64_b32"],
        "lds_pattern_frac": round(lds_tbs / probe["ds_read2st64_b32"], 4),
        "lds_pattern_note": "the same probe with the sweep's own read instruction (ds_read2st64_b32)",
        "dm_tile_reuse": {
            "lds_bytes_per_hbm_byte": round(lds_bytes / hbm_alg, 1),
            "algorithmic_hbm_bytes": hbm_alg,
            "note": "LDS bytes read per algorithmic HBM byte (4 B out per DM-trial sample + the packed "
                    "stream once); each staged window word is read by the 16 DM rows of a tile",
        },
        "adds": {"achieved_tadds": round(adds / t_sweep / 1e12, 1),
                 "note": "integer sample-adds per second; one 32-bit lane op adds 8 nibble or 4 byte fields"},
    }
    if traffic is not None:
        roofline["hbm"] = {"achieved_gbs": round(traffic / t_sweep / 1e9, 1), "peak_gbs": HBM_PEAK_GBS,
                           "frac": round(traffic / t_sweep / 1e9 / HBM_PEAK_GBS, 4),
                           "traffic_over_algorithmic": round(traffic / hbm_alg, 1)}
    if "sq_insts_valu_per_dispatch" in counters:
        lane_ops = counters["sq_insts_valu_per_dispatch"] * 64
        roofline["valu"] = {"lane_ops_per_launch": lane_ops,
                            "achieved_t_lane_ops": round(lane_ops / t_sweep / 1e12, 2),
                            "peak_t_lane_ops": round(VALU_LANE_OPS_PEAK_T, 2),
                            "frac": round(lane_ops / t_sweep / 1e12 / VALU_LANE_OPS_PEAK_T, 4),
                            "source": "SQ_INSTS_VALU x 64 lanes from the committed profile, this run's kernel time"}
    plain = plain_one_dm(packed, cfg, table, plan.row_len, plane)
    log(f"plain one-DM-per-launch form: {plain['value']:.3e} DM-trial samples/s ({plain['ms_per_dm']} ms per DM)")
    stage1_bytes = plan.nsamps * cfg.nchans * cfg.nbits // 8 + 4 * cfg.nchans * (-(-cfg.nsamps // 64) * 64)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "2-bit packed filterbank, 4096 ch x 2^20 samples (64 us), DM 0..2047 step 1 "
                        "(2048 trials), float32 DM x time plane",
            "nchans": cfg.nchans, "nsamps": cfg.nsamps, "nbits": cfg.nbits, "ndm": plan.ndm,
            "fch1": cfg.fch1, "foff": cfg.foff, "tsamp": cfg.tsamp,
            "dm_trial_samples_ragged": total, "dm_trial_samples_common": common,
            "delay_table_sha1": table_digest(table), "numpy_simd": numpy_simd_features(),
        },
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99)),
                       "unit": "one 2^20-sample gulp (67.1 s of data) through the sweep"},
        "stages_ms": {"k_build_g": round(float(build_ms.mean()), 3), "k_sweep": round(float(sweep_ms.mean()), 3),
                      "k_build_g_gbs": round(stage1_bytes / (build_ms.mean() * 1e-3) / 1e9, 1)},
        "common_length_value": common / sec_per_step,
        "pcie_inclusive": {"h2d_ms": round(h2d_ms, 3), "h2d_gbs": round(packed.numel() / (h2d_ms * 1e-3) / 1e9, 1),
                           "value": total / (sec_per_step + h2d_ms * 1e-3),
                           "note": "packed stream copied from pinned host memory before each sweep"},
        "sweep_variant": plan.variant,
        "tune_ms": tune_ms,
        "roofline": roofline,
        "plain_one_dm_per_launch": plain,
        "speedup_vs_plain_one_dm_per_launch": value / plain["value"],
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(packed, cfg, table, plan.row_len, plane, args.cpu_seconds)
        cb = result["cpu_baseline"]
        result["speedup_vs_cpu_baseline"] = value / cb["value"] if cb["value"] else None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
