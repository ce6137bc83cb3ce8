"""Seeded job mixes for the batched and served workloads.

Mixed-size jobs (64 by default): square images, side log-uniform in
[256, 2048] px, pixel spacing uniform in [0.25, 1] m, pulses uniform in
[512, 4096], range samples uniform in [2048, 8192], complex64, X-band radar
(fc 9.6 GHz, 4x upsampled 200 MHz sampling: dr = 0.1874 m), straight track at
3 km with 0.5 m seeded deviations, 1 kHz PRF at 150 m/s, image centred at 4 km
ground range.  The served stream adds a fraction of large jobs (8192 x 8192
image, 8192 pulses x 16384 samples).

Back-projection cost does not depend on the sample values, so the bulk of a
mix reads its profiles from shared seeded random buffers (one per size
class).  JobSpec.simulate() gives a job its own simulated scene instead --
point targets over complex-Gaussian clutter and AWGN, as the realistic
4096-pulse workload -- with the job's sizes, spacing and radar; the served and
batched checks against the oracle run on such jobs.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import native
from . import scene

C_LIGHT = 299792458.0
FC = 9.6e9
DR = C_LIGHT / (2 * 4 * 200e6)
V, PRF, ALT, YC = 150.0, 1000.0, 3000.0, 4000.0


@dataclass
class JobSpec:
    side: int
    spacing: float
    n_pulses: int
    n_samples: int
    large: bool
    seed: int

    @property
    def work(self) -> int:
        return self.side * self.side * self.n_pulses

    def positions(self) -> np.ndarray:
        rng = np.random.default_rng(self.seed)
        m = np.arange(self.n_pulses) - 0.5 * (self.n_pulses - 1)
        pos = np.zeros((self.n_pulses, 3))
        pos[:, 0] = V * m / PRF
        pos[:, 2] = ALT
        k = 200
        for ax in (1, 2):
            w = rng.standard_normal(self.n_pulses + k)
            s = np.convolve(w, np.ones(k) / k, mode="valid")[: self.n_pulses]
            s -= s.mean()
            pos[:, ax] += 0.5 * s / max(s.std(), 1e-12)
        return pos

    def grid(self) -> native.Grid:
        o = [-(self.side // 2) * self.spacing, YC - (self.side // 2) * self.spacing, 0.0]
        return native.make_grid(o, [self.spacing, 0, 0], [0, self.spacing, 0], self.side, self.side)

    @property
    def n_raw(self) -> int:
        """Raw samples per pulse whose compression covers the n_samples window:
        the window at 1/4 the upsampled rate plus one 2 us chirp (400 samples)."""
        return -(-self.n_samples // 4) + 400

    def radar(self) -> dict:
        """Range compression of this job's raw echoes (LFM 150 MHz / 2 us, 200 MHz
        complex sampling, 4x upsampling): the output axis is params()' r0, dr."""
        r0 = float(np.hypot(YC, ALT)) - 0.5 * self.n_samples * DR
        return dict(c=C_LIGHT, bandwidth=150e6, pulse_width=2e-6, fs=200e6, t0=2.0 * r0 / C_LIGHT,
                    n_raw=self.n_raw, upsample=4, crop_offset=0, n_out=self.n_samples)

    def scene_config(self, n_targets: int = 8, n_clutter: int = 1024) -> dict:
        """A scene config (scene.make_scene) with this job's sizes: the
        realistic workload's radar, track (0.5 m deviations low-passed over
        200 pulses) and scene statistics, n_targets point targets and
        n_clutter clutter scatterers, seeded by the job's seed."""
        cfg = scene.load_config("config1")
        cfg = scene.scaled_config(cfg, n_pulses=self.n_pulses, nx=self.side, ny=self.side,
                                  n_samples=self.n_samples, n_clutter=n_clutter,
                                  n_random_targets=n_targets)
        cfg["name"] = f"job-{self.seed}"
        cfg["seed"] = int(self.seed)
        cfg["image"]["spacing"] = float(self.spacing)
        cfg["scene"]["target_margin"] = max(24.0, 40.0 * self.spacing)
        return cfg

    def simulate(self, n_targets: int = 8, n_clutter: int = 1024) -> "scene.Scene":
        """This job's own simulated scene (its positions, grid and range axis
        replace positions(), grid() and params()' r0)."""
        return scene.make_scene(self.scene_config(n_targets, n_clutter))

    def params(self, interp="sinc8", precision="ref", dtype="c64") -> native.Params:
        r0 = float(np.hypot(YC, ALT)) - 0.5 * self.n_samples * DR
        return native.make_params(c=C_LIGHT, fc=FC, r0=r0, dr=DR, n_pulses=self.n_pulses,
                                  n_samples=self.n_samples, interp=interp, precision=precision,
                                  dtype=dtype, kaiser_beta=6.0)


def mixed_jobs(n_jobs=64, seed=20260031, large_fraction=0.0, large=(8192, 8192, 16384),
               side_range=(256, 2048), spacing_range=(0.25, 1.0), pulses_range=(512, 4096),
               samples_range=(2048, 8192)) -> list:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_jobs):
        if large_fraction > 0 and rng.uniform() < large_fraction:
            out.append(JobSpec(large[0], 0.25, large[1], large[2], True, seed + 1 + i))
            continue
        side = int(round(np.exp(rng.uniform(np.log(side_range[0]), np.log(side_range[1])))))
        out.append(JobSpec(side, float(rng.uniform(*spacing_range)),
                           int(rng.integers(pulses_range[0], pulses_range[1] + 1)),
                           int(rng.integers(samples_range[0], samples_range[1] + 1)), False,
                           seed + 1 + i))
    return out


def random_profiles(n_complex: int, seed: int) -> np.ndarray:
    """Seeded complex64 noise (host)."""
    rng = np.random.default_rng(seed)
    out = np.empty(n_complex, dtype=np.complex64)
    step = 1 << 24
    for i in range(0, n_complex, step):
        n = min(step, n_complex - i)
        v = rng.standard_normal(2 * n, dtype=np.float32)
        out[i:i + n] = v[0::2] + 1j * v[1::2]
    return out
