"""Synthetic DDM workloads: the analyses the engine is measured and tested on.

Each workload fixes a synthetic video (seeded, counter-based generators of the
engine library, so every run and every box sees the same frames) and the
analysis parameters the reference's ./ddm would be given for it (-S scales,
-T lags, -Q lambdas, -t tolerance, -C chunk):

* ``uniform128``: 64 frames of 128^2 uint8 i.i.d. uniform values, scale 128,
  lags 1..16, 20 linearly spaced lambdas 4..64, tolerance 20 %, chunk 30.
* ``brownian512``: 1024 frames of 512^2 uint8 Brownian video (2000 Gaussian
  spots, sigma 2 px, amplitude 80, D 0.5 px^2/frame, background 100, noise
  sigma 5), scales 512/256/128/64, 32 log lags 1..512, 30 log lambdas 4..256.
* ``brownian2048_u16``: 1000 frames of 2048^2 uint16 (12-bit) Brownian video
  (8000 spots, D 1.0, background 1000, noise sigma 30), ROI 1024^2 at the
  origin, scales 1024..32, 64 log lags 1..999, 40 log lambdas 2..512.
* the serving mix (``serving_streams``): streams of 256^2 / 512^2 / 1024^2
  uint8 Brownian video with D drawn from 0.2..2.0 px^2/frame, scales from the
  frame size down to 32, 24 log lags up to 300, 30 log lambdas 4..256,
  chunks of 30 frames.

Lag and lambda lists follow the integer log-spacing rule of ``log_lags`` /
``log_lambdas``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List

import numpy as np

from .engine import PIX_U8, PIX_U16LE, lin_lambdas, log_lags, log_lambdas, synth_brownian, synth_uniform


def host_threads() -> int:
    """Host threads a worker pool should use: the CPUs this process may run
    on, at most 16 (the share of one GPU on a multi-GPU host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


@dataclass
class Workload:
    name: str
    frames: int
    size: int              # frame width = height
    bits: int              # 8: uint8 pixels; 9..16: little-endian uint16
    scales: List[int]
    lag_max: int
    lag_count: int
    lam_lo: float
    lam_hi: float
    lam_count: int
    lam_linear: bool = False
    tol: float = 1.2
    chunk: int = 30
    generator: str = "brownian"
    gen: dict = field(default_factory=dict)
    seed: int = 0
    text: str = ""

    @property
    def bpp(self) -> int:
        return 1 if self.bits <= 8 else 2

    @property
    def pixfmt(self) -> int:
        return PIX_U8 if self.bits <= 8 else PIX_U16LE

    @property
    def lags(self) -> List[int]:
        if self.lag_count == self.lag_max:
            return list(range(1, self.lag_max + 1))
        return log_lags(self.lag_max, self.lag_count)

    @property
    def lambdas(self) -> List[float]:
        if self.lam_linear:
            return lin_lambdas(self.lam_lo, self.lam_hi, self.lam_count)
        return log_lambdas(self.lam_lo, self.lam_hi, self.lam_count)

    def make_frames(self, n_frames: int = 0, nthreads: int = 0) -> np.ndarray:
        """The workload's first n_frames frames (all of them by default),
        [n][size][size] uint8 or uint16."""
        n = n_frames or self.frames
        if self.generator == "uniform":
            return synth_uniform(n, self.size, self.size, self.bits, seed=self.seed)
        g = self.gen
        return synth_brownian(n, self.size, self.size, self.bits, g["spots"], g["sigma"], g["amplitude"],
                              g["diffusion"], g["background"], g["noise"], self.seed, nthreads or host_threads())

    def engine_args(self):
        """Positional arguments of ``Engine`` (width .. pixfmt)."""
        return (self.size, self.size, self.scales, self.lags, self.lambdas, self.tol, self.chunk, self.bpp,
                self.pixfmt)


WORKLOADS = {
    "uniform128": Workload(
        "uniform128", frames=64, size=128, bits=8, scales=[128], lag_max=16, lag_count=16,
        lam_lo=4.0, lam_hi=64.0, lam_count=20, lam_linear=True, generator="uniform", seed=1,
        text="64x128^2 u8 i.i.d. uniform, scale 128, lags 1-16, 20 lambdas 4-64 px, tol 20%, chunk 30"),
    "brownian512": Workload(
        "brownian512", frames=1024, size=512, bits=8, scales=[512, 256, 128, 64], lag_max=512, lag_count=32,
        lam_lo=4.0, lam_hi=256.0, lam_count=30, seed=2026,
        gen=dict(spots=2000, sigma=2.0, amplitude=80.0, diffusion=0.5, background=100.0, noise=5.0),
        text="single-stream Brownian 1024x512^2 u8, scales 512/256/128/64, 32 log lags 1-512, "
             "30 log lambdas 4-256 px, tol 20%, chunk 30"),
    "brownian2048_u16": Workload(
        "brownian2048_u16", frames=1000, size=2048, bits=12, scales=[1024, 512, 256, 128, 64, 32], lag_max=999,
        lag_count=64, lam_lo=2.0, lam_hi=512.0, lam_count=40, seed=2027,
        gen=dict(spots=8000, sigma=2.0, amplitude=80.0, diffusion=1.0, background=1000.0, noise=30.0),
        text="single-stream Brownian 1000x2048^2 u16 (12-bit), ROI 1024^2, scales 1024..32, 64 log lags "
             "1-999, 40 log lambdas 2-512 px, tol 20%, chunk 30"),
}

# the serving mix
SERVING_SIZES = (256, 512, 1024)
SERVING_LAGS = log_lags(300, 24)
SERVING_LAMBDAS = log_lambdas(4.0, 256.0, 30)
SERVING_CHUNK = 30


@dataclass
class ServingStream:
    size: int
    diffusion: float
    seed: int

    @property
    def scales(self) -> List[int]:
        return [1 << k for k in range(int(np.log2(self.size)), 4, -1)]

    def desc(self, lag_block: int = 32) -> dict:
        return dict(width=self.size, height=self.size, scales=self.scales, lags=SERVING_LAGS,
                    lambdas=SERVING_LAMBDAS, chunk=SERVING_CHUNK, lag_block=lag_block)

    def make_frames(self, n_frames: int, nthreads: int = 0) -> np.ndarray:
        # 2000 spots per 512^2, as in the single-stream workload
        spots = max(1, 2000 * self.size * self.size // (512 * 512))
        return synth_brownian(n_frames, self.size, self.size, 8, spots, 2.0, 80.0, self.diffusion, 100.0, 5.0,
                              self.seed, nthreads or host_threads())


def serving_streams(count: int, seed: int = 7) -> List[ServingStream]:
    """`count` streams of the serving mix: frame size uniform over
    {256^2, 512^2, 1024^2}, D uniform in 0.2..2.0 px^2/frame."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice(SERVING_SIZES, size=count)
    ds = rng.uniform(0.2, 2.0, size=count)
    return [ServingStream(int(s), float(d), seed * 1000 + i) for i, (s, d) in enumerate(zip(sizes, ds))]
