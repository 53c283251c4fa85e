"""Seeded synthetic filterbanks for benchmarks and tests.

Generates quantised Gaussian noise with dispersed boxcar pulses directly in packed
SIGPROC layout (time-major ``[nsamps][nchans]``, ``nbits`` per sample, bit order as
sigpyproc reads it: 1-bit little, 2/4-bit big), on any torch device -- on the GPU the
stream is born in HBM. Pulses are placed with the same host delay table the
dedispersion uses, so a pulse at DM ``d`` and time ``t`` lands in channel ``c`` at
``t + delay_c(d)``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from sigpyproc3_amd import params
from sigpyproc3_amd.io import sigproc

TWO_BIT_EDGE = 0.9816  # 4-level optimum thresholds at {-0.9816, 0, +0.9816} sigma


@dataclass
class Pulse:
    dm: float
    t: int
    width: int
    amplitude: float


@dataclass
class SynthConfig:
    nchans: int
    nsamps: int
    nbits: int
    fch1: float
    foff: float
    tsamp: float
    seed: int
    mean: float = 128.0   # 8-bit noise mean
    sigma: float = 16.0   # 8-bit noise sigma
    pulses: list[Pulse] = field(default_factory=list)

    @property
    def chan_freqs(self) -> np.ndarray:
        return np.arange(self.nchans, dtype=np.float32) * self.foff + self.fch1

    def delays(self, dms, ref_freq: float | None = None) -> np.ndarray:
        ref = self.fch1 if ref_freq is None else ref_freq
        return np.atleast_2d(params.compute_dmdelays(self.chan_freqs, dms, self.tsamp, float(ref)))

    def header_dict(self, source: str = "synthetic") -> dict:
        return {"telescope_id": 0, "machine_id": 0, "data_type": 1, "source_name": source,
                "tstart": 60000.0, "tsamp": self.tsamp, "nbits": self.nbits, "fch1": self.fch1,
                "foff": self.foff, "nchans": self.nchans, "nifs": 1}


def _bitorder_big(nbits: int) -> bool:
    return nbits in (2, 4)


def quantise(x, nbits: int, cfg: SynthConfig):
    """Map standard-normal floats to unsigned samples of ``nbits``."""
    t = _torch()
    if nbits == 8:
        return t.clamp(t.round(x * cfg.sigma + cfg.mean), 0, 255).to(t.uint8)
    if nbits == 2:
        q = (x > -TWO_BIT_EDGE).to(t.uint8) + (x > 0).to(t.uint8) + (x > TWO_BIT_EDGE).to(t.uint8)
        return q
    if nbits == 1:
        return (x > 0).to(t.uint8)
    if nbits == 4:
        return t.clamp(t.round(x * 3.0 + 7.5), 0, 15).to(t.uint8)
    msg = f"nbits {nbits} not supported"
    raise ValueError(msg)


def pack_rows(samples, nbits: int):
    """Pack uint8 samples ``[rows, nchans]`` (values < 2^nbits) along channels."""
    t = _torch()
    if nbits == 8:
        return samples.reshape(-1)
    per = 8 // nbits
    s = samples.reshape(-1, per).to(t.int32)
    out = t.zeros(s.shape[0], dtype=t.int32, device=s.device)
    big = _bitorder_big(nbits)
    for k in range(per):
        shift = (8 - (k + 1) * nbits) if big else k * nbits
        out |= s[:, k] << shift
    return out.to(t.uint8)


def generate_packed(cfg: SynthConfig, device="cpu", chunk: int = 1 << 16):
    """Packed stream as a flat uint8 torch tensor on ``device``."""
    t = _torch()
    dev = t.device(device)
    gen = t.Generator(device=dev)
    gen.manual_seed(int(cfg.seed))
    row_bytes = cfg.nchans * cfg.nbits // 8
    out = t.empty(cfg.nsamps * row_bytes, dtype=t.uint8, device=dev)
    maxv = (1 << cfg.nbits) - 1
    pul = []
    for p in cfg.pulses:
        d = t.from_numpy(cfg.delays([p.dm])[0].astype(np.int64)).to(dev)
        pul.append((p, d))
    chans = t.arange(cfg.nchans, device=dev)
    for a in range(0, cfg.nsamps, chunk):
        b = min(cfg.nsamps, a + chunk)
        x = t.randn((b - a, cfg.nchans), generator=gen, device=dev, dtype=t.float32)
        q = quantise(x, cfg.nbits, cfg).to(t.int32)
        for p, d in pul:
            for k in range(p.width):
                rows = p.t + k + d - a
                ok = (rows >= 0) & (rows < b - a)
                if bool(ok.any()):
                    r, c = rows[ok], chans[ok]
                    q[r, c] = t.clamp(q[r, c] + int(round(p.amplitude)), 0, maxv)
        out[a * row_bytes : b * row_bytes] = pack_rows(q.to(t.uint8), cfg.nbits)
    return out


def write_sigproc(path, cfg: SynthConfig, packed) -> None:
    """Write a SIGPROC file: header + the packed stream."""
    hdr = sigproc.encode_header(cfg.header_dict())
    data = packed.cpu().numpy() if hasattr(packed, "cpu") else np.asarray(packed, dtype=np.uint8)
    with open(path, "wb") as fp:
        fp.write(hdr)
        fp.write(data.tobytes())


# Benchmark and test workloads.
def cfg_single_dm() -> SynthConfig:
    """8-bit, 1024 ch, 65536 samples, DM-300 boxcar (width 16, +24) at sample 20000."""
    return SynthConfig(nchans=1024, nsamps=65536, nbits=8, fch1=1500.0, foff=-0.390625,
                       tsamp=64e-6, seed=1, pulses=[Pulse(300.0, 20000, 16, 24.0)])


def cfg_sweep_2bit() -> SynthConfig:
    """2-bit, 4096 ch, 2^20 samples, pulses at DM 150/800/1700 (+1 level, width 8)."""
    return SynthConfig(nchans=4096, nsamps=1 << 20, nbits=2, fch1=1550.0, foff=-0.0732421875,
                       tsamp=64e-6, seed=2,
                       pulses=[Pulse(150.0, 1 << 18, 8, 1.0), Pulse(800.0, 1 << 19, 8, 1.0),
                               Pulse(1700.0, 3 << 18, 8, 1.0)])


def cfg_sweep_mf() -> SynthConfig:
    """8-bit, 4096 ch, 2^19 samples (the DM 0..1023 sweep + matched filter workload)."""
    return SynthConfig(nchans=4096, nsamps=1 << 19, nbits=8, fch1=1550.0, foff=-0.0732421875,
                       tsamp=64e-6, seed=4, pulses=[Pulse(500.0, 200000, 32, 10.0)])


_t = None


def _torch():
    global _t
    if _t is None:
        import torch

        _t = torch
    return _t


def cfg_batch_instance(i: int) -> tuple[SynthConfig, np.ndarray]:
    """Instance ``i`` of the 64 uneven sweep instances (seed 3 + i).

    nchans in {336, 1024, 4096}, nsamps ~ U[2^16, 2^19], nbits in {1, 2, 4, 8}, tsamp in
    {64, 256} us, band 1250-1550 MHz (fch1 1550, foff -300/nchans), a DM list of
    U[256, 4096] trials spaced linearly from 0 to a maximum DM ~ U[500, 3000].
    Returns the configuration and the DM list.
    """
    rng = np.random.default_rng(3 + i)
    nchans = int(rng.choice([336, 1024, 4096]))
    nsamps = int(rng.integers(1 << 16, (1 << 19) + 1))
    nbits = int(rng.choice([1, 2, 4, 8]))
    tsamp = float(rng.choice([64e-6, 256e-6]))
    ndm = int(rng.integers(256, 4097))
    max_dm = float(rng.uniform(500.0, 3000.0))
    dms = np.linspace(0.0, max_dm, ndm)
    cfg = SynthConfig(nchans=nchans, nsamps=nsamps, nbits=nbits, fch1=1550.0, foff=-300.0 / nchans,
                      tsamp=tsamp, seed=3 + i)
    amp = 1.0 if nbits <= 2 else (3.0 if nbits == 4 else 12.0)
    cfg.pulses = [Pulse(float(dms[ndm // 2]), nsamps // 4, 16, amp)]
    return cfg, dms
