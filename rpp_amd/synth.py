"""Seeded synthetic audio batches for the five workloads (tests and benchmark).

A counter-based SplitMix64 stream gives the same numbers on any machine: uniform floats take
the top 24 bits, normals use Box-Muller.  Seeds are 0x5EED0000 + workload id.

  0  batch 8 mono 16 kHz, L ~ U{16000..48000}, x = U[-1,1] + 0.3 sin(2 pi 440 t)
  1  batch 512 mono 16 kHz, L log-uniform 0.5-20 s, x = N(0, 0.1^2) + 0.5 * chirp(f0 -> f1)
  2  batch 256 mono 44.1 kHz, L ~ U{2..10 s}, x = U[-1,1]
  3  served stream: Poisson arrivals, L log-uniform 0.25-30 s, x as workload 0
  4  batch 64 stereo 48 kHz int16-range / 32768, 60 s each (interleaved)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SEED0 = 0x5EED0000
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 starting at counter `offset` (vectorised)."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Sequential draws from one SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = seed & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        z = splitmix64(self.seed, n, self.pos)
        self.pos += n
        return z

    def uniform(self, n: int, lo=0.0, hi=1.0) -> np.ndarray:
        """float64 in [lo, hi) from the top 24 bits (exactly representable as float32 steps)."""
        u = (self.u64(n) >> np.uint64(40)).astype(np.float64) / float(1 << 24)
        return lo + (hi - lo) * u

    def normal(self, n: int, sigma=1.0) -> np.ndarray:
        m = (n + 1) // 2
        u1 = np.maximum(self.uniform(m), 1.0 / (1 << 25))
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
        return sigma * z

    def integers(self, n: int, lo: int, hi_incl: int) -> np.ndarray:
        span = hi_incl - lo + 1
        return lo + (self.u64(n) % np.uint64(span)).astype(np.int64)


@dataclass
class Batch:
    x: np.ndarray              # [B, stride] float32 (interleaved channels for stereo)
    lengths: np.ndarray        # [B] int32 samples (per channel)
    sample_rate: float
    channels: int = 1
    params: dict = field(default_factory=dict)

    @property
    def audio_seconds(self) -> float:
        return float(self.lengths.sum()) / self.sample_rate


FEATURES_16K = dict(nfft=512, W=400, H=160, nmel=80, sample_rate=16000.0, fmin=0.0, fmax=8000.0)
FEATURES_44K = dict(nfft=1024, W=1024, H=256, nmel=128, sample_rate=44100.0, fmin=0.0, fmax=22050.0)


def _pad_stride(n: int) -> int:
    return (n // 8) * 8 + 8


def workload0(batch: int = 8, seed: int = SEED0 + 0) -> Batch:
    s = Stream(seed)
    L = s.integers(batch, 16000, 48000).astype(np.int32)
    S = _pad_stride(int(L.max()))
    x = np.zeros((batch, S), np.float32)
    for b in range(batch):
        n = int(L[b])
        t = np.arange(n, dtype=np.float64) / 16000.0
        x[b, :n] = (s.uniform(n, -1.0, 1.0) + 0.3 * np.sin(2 * np.pi * 440.0 * t)).astype(np.float32)
    return Batch(x, L, 16000.0, params=dict(FEATURES_16K))


def log_uniform_lengths(s: Stream, n: int, lo: int, hi: int) -> np.ndarray:
    return np.round(np.exp(s.uniform(n, np.log(lo), np.log(hi)))).astype(np.int32)


def workload1(batch: int = 512, seed: int = SEED0 + 1, max_seconds: float = 20.0) -> Batch:
    s = Stream(seed)
    L = log_uniform_lengths(s, batch, 8000, int(max_seconds * 16000))
    f0 = s.uniform(batch, 50.0, 4000.0)
    f1 = s.uniform(batch, 50.0, 8000.0)
    S = _pad_stride(int(L.max()))
    x = np.zeros((batch, S), np.float32)
    for b in range(batch):
        n = int(L[b])
        t = np.arange(n, dtype=np.float64) / 16000.0
        T = n / 16000.0
        chirp = 0.5 * np.sin(2 * np.pi * (f0[b] * t + (f1[b] - f0[b]) * t * t / (2 * T)))
        x[b, :n] = (s.normal(n, 0.1) + chirp).astype(np.float32)
    return Batch(x, L, 16000.0, params=dict(FEATURES_16K, f0=f0, f1=f1))


def workload2(batch: int = 256, seed: int = SEED0 + 2) -> Batch:
    s = Stream(seed)
    L = s.integers(batch, 88200, 441000).astype(np.int32)
    S = _pad_stride(int(L.max()))
    x = np.zeros((batch, S), np.float32)
    for b in range(batch):
        n = int(L[b])
        x[b, :n] = s.uniform(n, -1.0, 1.0).astype(np.float32)
    return Batch(x, L, 44100.0, params=dict(FEATURES_44K))


def arrivals(rate: float, duration_s: float, seed: int = SEED0 + 3):
    """Poisson arrival times (s) and log-uniform 0.25-30 s lengths at 16 kHz."""
    s = Stream(seed)
    n = int(rate * duration_s * 1.5) + 16
    gaps = -np.log(np.maximum(s.uniform(n), 1e-12)) / rate
    t = np.cumsum(gaps)
    t = t[t < duration_s]
    L = log_uniform_lengths(s, len(t), 4000, 480000)
    return t, L


def clip_for_request(i: int, n: int, seed: int = SEED0 + 3) -> np.ndarray:
    """Request i's samples (workload-0 signal model), independent of batching."""
    s = Stream(seed * 1000003 + i)
    t = np.arange(n, dtype=np.float64) / 16000.0
    return (s.uniform(n, -1.0, 1.0) + 0.3 * np.sin(2 * np.pi * 440.0 * t)).astype(np.float32)


def workload4(batch: int = 64, seconds: float = 60.0, seed: int = SEED0 + 4) -> Batch:
    s = Stream(seed)
    n = int(seconds * 48000)
    x = np.empty((batch, 2 * n), np.float32)
    for b in range(batch):
        v = s.integers(2 * n, -32768, 32767).astype(np.float32) / np.float32(32768.0)
        x[b] = v
    L = np.full(batch, n, np.int32)
    return Batch(x, L, 48000.0, channels=2, params=dict(FEATURES_16K))


def kaiser_sinc_lut(coeffs: int = 4091, lobes: int = 16, beta: float = 8.0):
    """Resampling table for workload 4: a Kaiser-windowed sinc (16 zero-crossings each side) in the
    layout RpptResamplingWindow expects (the reference test suite's builder,
    utilities/test_suite/rpp_test_suite_audio.h:390-430, with the Hann envelope replaced by a
    Kaiser one): lookupSize = coeffs + 5, lut[1 + i] for i < coeffs, zero elsewhere.  With
    coeffs = 4091 the table has 4096 entries.  Returns (lut float32, scale, center)."""
    scale = np.float32(2.0 * lobes / (coeffs - 1))
    center = int((coeffs - 1) * 0.5)
    lut = np.zeros(coeffs + 5, np.float32)
    i = np.arange(coeffs)
    x = (i - center).astype(np.float64) * scale
    y = (i - center).astype(np.float64) / center
    sinc = np.sinc(x)
    env = np.i0(beta * np.sqrt(np.clip(1.0 - y * y, 0.0, 1.0))) / np.i0(beta)
    lut[1:coeffs + 1] = (sinc * env).astype(np.float32)
    return lut, float(np.float32(1.0) / scale), float(center + 1)


def make(workload: int, batch: int | None = None, **kw) -> Batch:
    fn = {0: workload0, 1: workload1, 2: workload2, 4: workload4}[workload]
    return fn(batch, **kw) if batch is not None else fn(**kw)
