"""Seeded synthetic DEMs: spectral (1/f^beta) fractal terrain plus Gaussian
ridges, rescaled to int16 metres.  Deterministic for a given seed (numpy's
PCG64 stream and FFT), so tests and benchmarks regenerate identical inputs.

The workloads (sizes, spacings, height ranges, line of sight) are the ones
the engine is measured and checked on:

  small-201   201 x 201 at 100 m, beta 2.6, [0, 1500] m, los 6.7 km
  main-1201   1201 x 1201 at 90 m, beta 2.4 + 8 ridges, [0, 3000] m, los 36 km
  tiles-64    64 tiles, sides from {257, 401, 601, 801, 1201}, 30 or 90 m,
              beta 2.4, [0, 4000] m, los = (w div 3) * spacing, seed 1000+i
  large-3601  3601 x 3601 at 30 m, beta 2.2 + 20 ridges, [0, 4500] m, los 40 km
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def spectral_fractal(width: int, beta: float, rng: np.random.Generator) -> np.ndarray:
    """White noise on a power-of-two grid >= width, shaped by |k|^(-beta/2)
    in the Fourier domain (k = 0 removed), cropped to width x width."""
    n = 1
    while n < width:
        n <<= 1
    noise = rng.standard_normal((n, n))
    spec = np.fft.rfft2(noise)
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = np.inf
    spec *= k ** (-beta / 2.0)
    field = np.fft.irfft2(spec, s=(n, n))[:width, :width]
    return field


def add_ridges(field_m: np.ndarray, spacing: float, count: int,
               rng: np.random.Generator, sigma_km=(3.0, 10.0),
               amplitude_m=(500.0, 1500.0)) -> np.ndarray:
    """Gaussian ridges about random lines through the tile (metres)."""
    w = field_m.shape[0]
    yy, xx = np.mgrid[0:w, 0:w].astype(np.float64) * spacing
    out = field_m.copy()
    for _ in range(count):
        cx, cy = rng.uniform(0, w * spacing, size=2)
        theta = rng.uniform(0, np.pi)
        sigma = rng.uniform(*sigma_km) * 1000.0
        amp = rng.uniform(*amplitude_m)
        d = (xx - cx) * np.sin(theta) - (yy - cy) * np.cos(theta)
        out += amp * np.exp(-0.5 * (d / sigma) ** 2)
    return out


def rescale_int16(field: np.ndarray, hmax: float) -> np.ndarray:
    lo, hi = float(field.min()), float(field.max())
    span = hi - lo if hi > lo else 1.0
    return np.rint((field - lo) / span * hmax).astype(np.int16)


def fractal_dem(width: int, spacing: float, beta: float, hmax: float, seed: int,
                ridges: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    f = spectral_fractal(width, beta, rng)
    lo, hi = float(f.min()), float(f.max())
    f = (f - lo) / (hi - lo if hi > lo else 1.0) * hmax
    if ridges:
        f = add_ridges(f, spacing, ridges, rng)
    return rescale_int16(f, hmax)


@dataclass
class Workload:
    name: str
    width: int
    spacing: float
    max_line_of_sight: int
    elevations: np.ndarray          # int16, width*width, row-major
    observer_height: float = 1.65
    rings_per_km: float = 5.0


def small_201(seed: int = 0) -> Workload:
    return Workload("small-201", 201, 100.0, 6700,
                    fractal_dem(201, 100.0, 2.6, 1500.0, seed).reshape(-1))


def main_1201(seed: int = 0) -> Workload:
    return Workload("main-1201", 1201, 90.0, 36000,
                    fractal_dem(1201, 90.0, 2.4, 3000.0, seed, ridges=8).reshape(-1))


def large_3601(seed: int = 0) -> Workload:
    return Workload("large-3601", 3601, 30.0, 40000,
                    fractal_dem(3601, 30.0, 2.2, 4500.0, seed, ridges=20).reshape(-1))


TILE_SIDES = (257, 401, 601, 801, 1201)


def tiles_64(count: int = 64, first_seed: int = 1000) -> list[Workload]:
    out = []
    for i in range(count):
        rng = np.random.default_rng(first_seed + i)
        w = int(rng.choice(TILE_SIDES))
        spacing = 30.0 if rng.integers(0, 2) == 0 else 90.0
        los = int((w // 3) * spacing)
        out.append(Workload(f"tile-{i}", w, spacing, los,
                            fractal_dem(w, spacing, 2.4, 4000.0, first_seed + i).reshape(-1)))
    return out
