"""Seeded synthetic workloads of the split-operator path.

The reference has no seeding convention for these problems (its tests use no
RNG; scenes use std::mt19937, scenario/src/harmonic1d_director.ixx:52), and
std::uniform_real_distribution is not portable across standard libraries.
Parameters are drawn here from a splitmix64 stream mapped to doubles with 53
random bits, so the same seed gives the same inputs on every machine.

Workloads (atomic units, periodic grids [lo, hi)):

* ``ho1d``      1D harmonic oscillator, N = 65536 on [-100, 100), w = 0.1,
                coherent packet displaced by x0 ~ U(-20, 20), dt = 0.01, 2000 steps.
* ``hydrogen``  3D hydrogen-like atom, 128^3 on [-20, 20)^3, V = -1/max(r, h/2),
                packet sigma = 1.5 at r0 ~ U(2, 4) in a random direction with
                momentum |k| ~ U(0, 0.5) in a random direction, dt = 0.005, 1000 steps.
* ``barrier``   3D barrier tunnelling, 256^3 on [-40, 40)^3, V0 ~ U(0.5, 2) on
                0 <= x < 2, sin^2 absorbing mask 8 wide, packet sigma = 3 at
                (-15, 0, 0) with kx ~ U(0.8, 1.5), dt = 0.01, 2000 steps, FP32 state.

The coherent state of the oscillator has amplitude width 1/sqrt(w); in the
reference's Gaussian convention (sigma = density standard deviation,
core/src/wavepacket.ixx:16-25) that is sigma = 1/sqrt(2 w).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) with 53-bit doubles in [0, 1)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        u = (self.next_u64() >> 11) * (1.0 / (1 << 53))
        return a + (b - a) * u

    def unit_vector(self):
        z = self.uniform(-1.0, 1.0)
        phi = self.uniform(0.0, 2.0 * math.pi)
        s = math.sqrt(max(0.0, 1.0 - z * z))
        return (s * math.cos(phi), s * math.sin(phi), z)

    def choice(self, items):
        return items[self.next_u64() % len(items)]


@dataclass
class Solve:
    """One solve: grid, potential recipe, packet, step size and count."""
    name: str
    n: tuple            # (nx, ny, nz)
    lo: tuple
    hi: tuple
    potential: str      # harmonic | coulomb_floor | regularized_coulomb | soft_coulomb | barrier | free
    params: dict
    r0: tuple
    sigma: tuple
    k0: tuple
    dt: float
    steps: int
    precision: str = "fp64"
    mask_width: float = 0.0
    mass: float = 1.0
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def points(self) -> int:
        return int(np.prod(self.n))

    def grid(self):
        from . import Grid1D, Grid3D
        if self.n[1] == 1 and self.n[2] == 1:
            return Grid1D(self.lo[0], self.hi[0], self.n[0])
        return Grid3D(Grid1D(self.lo[0], self.hi[0], self.n[0]),
                      Grid1D(self.lo[1], self.hi[1], self.n[1]),
                      Grid1D(self.lo[2], self.hi[2], self.n[2]))

    def build_potential(self) -> np.ndarray:
        from . import (barrier_potential, coulomb_floor_potential, harmonic_potential,
                       regularized_coulomb_potential, soft_coulomb_potential)
        g = self.grid()
        p = self.params
        if self.potential == "harmonic":
            c = p.get("center", (0.0, 0.0, 0.0))
            return harmonic_potential(g, p["omega"], c[0] if self.n[1] == self.n[2] == 1 else c)
        if self.potential == "coulomb_floor":
            return coulomb_floor_potential(g, p["Z"], p.get("center", (0.0, 0.0, 0.0)))
        if self.potential == "regularized_coulomb":
            return regularized_coulomb_potential(g, p["Z"], p.get("center", (0.0, 0.0, 0.0)))
        if self.potential == "barrier":
            return barrier_potential(g, p["V0"], p["x_lo"], p["x_hi"])
        if self.potential == "soft_coulomb":
            c = p.get("center", (0.0, 0.0, 0.0))
            return soft_coulomb_potential(g, p["Z"], p["a"],
                                          c[0] if self.n[1] == self.n[2] == 1 else c)
        if self.potential == "free":
            return np.zeros(self.points)
        raise ValueError(f"unknown potential {self.potential}")

    def build_mask(self):
        if self.mask_width <= 0.0:
            return None
        from . import absorbing_mask
        return absorbing_mask(self.grid(), self.mask_width)

    def recipe(self):
        """The same inputs as a device-build recipe (``Propagator.build``)."""
        from .capi import TDSE_POT, Recipe
        r = Recipe()
        p = self.params
        if self.potential == "harmonic":
            r.potential, r.strength = TDSE_POT["harmonic"], p["omega"]
        elif self.potential in ("regularized_coulomb", "coulomb_floor"):
            r.potential, r.strength = TDSE_POT[self.potential], p["Z"]
        elif self.potential == "barrier":
            r.potential, r.strength = TDSE_POT["barrier"], p["V0"]
            r.x_lo, r.x_hi = p["x_lo"], p["x_hi"]
        elif self.potential == "soft_coulomb":
            r.potential, r.strength, r.softening = TDSE_POT["soft_coulomb"], p["Z"], p["a"]
        elif self.potential == "free":
            r.potential = TDSE_POT["none"]
        else:
            raise ValueError(f"unknown potential {self.potential}")
        c = p.get("center", (0.0, 0.0, 0.0))
        for a in range(3):
            r.center[a] = c[a]
            r.r0[a], r.sigma[a], r.k0[a] = self.r0[a], self.sigma[a], self.k0[a]
        r.mask_width = self.mask_width
        return r

    def build_psi0(self) -> np.ndarray:
        from . import gaussian_wavepacket
        g = self.grid()
        if self.n[1] == 1 and self.n[2] == 1:
            return gaussian_wavepacket(g, self.r0[0], self.sigma[0], self.k0[0])
        return gaussian_wavepacket(g, self.r0, self.sigma, self.k0)


def ho1d(seed: int = 1, n: int = 65536, steps: int = 2000) -> Solve:
    rng = SplitMix64(seed)
    omega = 0.1
    x0 = rng.uniform(-20.0, 20.0)
    s = 1.0 / math.sqrt(2.0 * omega)
    return Solve("ho1d", (n, 1, 1), (-100.0, 0.0, 0.0), (100.0, 1.0, 1.0), "harmonic",
                 {"omega": omega}, (x0, 0.0, 0.0), (s, 1.0, 1.0), (0.0, 0.0, 0.0), 0.01, steps,
                 "fp64", seed=seed)


def hydrogen(seed: int = 2, n: int = 128, steps: int = 1000, precision: str = "fp64") -> Solve:
    rng = SplitMix64(seed)
    r = rng.uniform(2.0, 4.0)
    u = rng.unit_vector()
    kappa = rng.uniform(0.0, 0.5)
    v = rng.unit_vector()
    half = 20.0 * n / 128.0 if n != 128 else 20.0
    return Solve("hydrogen", (n, n, n), (-half,) * 3, (half,) * 3, "coulomb_floor", {"Z": 1.0},
                 tuple(r * c for c in u), (1.5, 1.5, 1.5), tuple(kappa * c for c in v), 0.005,
                 steps, precision, seed=seed)


def barrier(seed: int = 3, n: int = 256, steps: int = 2000, precision: str = "fp32") -> Solve:
    rng = SplitMix64(seed)
    v0 = rng.uniform(0.5, 2.0)
    kx = rng.uniform(0.8, 1.5)
    half = 40.0 * n / 256.0
    return Solve("barrier", (n, n, n), (-half,) * 3, (half,) * 3, "barrier",
                 {"V0": v0, "x_lo": 0.0, "x_hi": 2.0}, (-15.0, 0.0, 0.0), (3.0, 3.0, 3.0),
                 (kx, 0.0, 0.0), 0.01, steps, precision, mask_width=8.0, seed=seed)


BATCH_SHAPES = [(64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 160, 160),
                (512, 512, 1), (1024, 1024, 1), (65536, 1, 1)]


def batch_instances(seed: int = 4, count: int = 48, steps: int = 500,
                    precision: str = "fp32", shapes=None) -> list:
    """Uneven independent solves: shape from the list above (3D boxes of 40 a.u.,
    2D of 80 a.u., 1D of 200 a.u.), harmonic (w ~ U(0.05, 0.5)) or
    regularized-Coulomb (Z ~ U(0.5, 2), 3D/2D only) potential, Gaussian packet
    with random centre, width and momentum."""
    rng = SplitMix64(seed)
    shapes = shapes or BATCH_SHAPES
    out = []
    for i in range(count):
        n = rng.choice(shapes)
        dims = sum(1 for v in n if v > 1)
        half = {3: 20.0, 2: 40.0, 1: 100.0}[dims]
        lo = tuple(-half if v > 1 else 0.0 for v in n)
        hi = tuple(half if v > 1 else 1.0 for v in n)
        coulomb = dims > 1 and rng.uniform() < 0.5
        if coulomb:
            pot, params, dt = "regularized_coulomb", {"Z": rng.uniform(0.5, 2.0)}, 0.005
        else:
            pot, params, dt = "harmonic", {"omega": rng.uniform(0.05, 0.5)}, 0.01
        r0 = tuple(rng.uniform(-0.2 * half, 0.2 * half) if v > 1 else 0.0 for v in n)
        sig = rng.uniform(1.0, 2.5)
        sigma = tuple(sig if v > 1 else 1.0 for v in n)
        k0 = tuple(rng.uniform(-0.8, 0.8) if v > 1 else 0.0 for v in n)
        out.append(Solve(f"batch{i}", n, lo, hi, pot, params, r0, sigma, k0, dt, steps,
                         precision, seed=seed))
    return out
