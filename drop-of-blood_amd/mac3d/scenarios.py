"""Seeded workloads of the MAC step (host logic shared by tests and bench).

The reference seeds its scene from a float `globalSeed` packed into the
uniform block (uniformLayout.ts:93, consumed by FLUID_INIT_VEL fluid.ts:92);
the workloads here instead place spherical blood blobs from a splitmix64
stream.  The same generator exists in js/scenarios.js and both must agree
draw for draw (tests/test_host.py checks it through node).

Blob positions are in cell units: cell (i, j, k) has its centre at
(i + 0.5, j + 0.5, k + 0.5); y is vertical.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import List

MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) with 53-bit uniform doubles."""

    def __init__(self, seed: int):
        self.state = seed & MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def range(self, a: float, b: float) -> float:
        return a + (b - a) * self.uniform()


@dataclass
class Blob:
    cx: float
    cy: float
    cz: float
    radius: float
    w0: float


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dt: float
    steps: int
    levels: int
    cycles: int
    precision: int  # 0 fp32, 1 mixed, 2 fp64
    blobs: List[Blob] = field(default_factory=list)
    seed: int = 0
    nu1: int = 2
    nu2: int = 2
    ir_tol: float = 1e-8

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


def _blobs(rng: SplitMix64, nx, ny, nz, count, r_lo, r_hi, stagger, w_lo, w_hi):
    out = []
    y0 = None
    for b in range(count):
        r = rng.range(r_lo, r_hi) if r_hi > r_lo else r_lo
        cx = rng.range(r + 2.0, nx - r - 2.0)
        cz = rng.range(r + 2.0, nz - r - 2.0)
        if y0 is None:
            y0 = rng.range(ny * 2.0 / 3.0, ny - r - 2.0)
        cy = y0 - stagger * b
        w0 = rng.range(w_lo, w_hi) if w_hi > w_lo else w_lo
        out.append(Blob(cx, cy, cz, r, w0))
    return out


def workload(name: str, seed: int = 1) -> Workload:
    """The five workloads of the MAC step (sizes in cells, SI units).

    tiny  : 32x48x32,   dx 1 mm,     dt 0.5 ms,  20 steps, 1 blob r=4, w 0.5
    ref   : 192x256x192, dx 0.5 mm,  dt 0.25 ms, 200 steps, 3 blobs r=7+-1
    mixed : same as ref, mixed precision projection
    adjoint: 64x96x64,  dx 1 mm,     dt 0.5 ms,  50 steps, 1 blob r=6
    large : 256x384x256, dx 0.375 mm, 5 levels,  5 blobs r in [6,10]
    """
    rng = SplitMix64(seed)
    if name == "tiny":
        w = Workload(name, 32, 48, 32, 1e-3, 5e-4, 20, 3, 8, 0, seed=seed)
        w.blobs = _blobs(rng, w.nx, w.ny, w.nz, 1, 4.0, 4.0, 0.0, 0.5, 0.5)
    elif name in ("ref", "mixed"):
        w = Workload(name, 192, 256, 192, 5e-4, 2.5e-4, 200, 3, 8,
                     1 if name == "mixed" else 0, seed=seed)
        w.blobs = _blobs(rng, w.nx, w.ny, w.nz, 3, 6.0, 8.0, 40.0, 0.3, 0.8)
    elif name == "adjoint":
        w = Workload(name, 64, 96, 64, 1e-3, 5e-4, 50, 3, 8, 0, seed=seed)
        w.blobs = _blobs(rng, w.nx, w.ny, w.nz, 1, 6.0, 6.0, 0.0, 0.5, 0.5)
    elif name == "large":
        w = Workload(name, 256, 384, 256, 3.75e-4, 1.875e-4, 200, 5, 8, 1, seed=seed)
        w.blobs = _blobs(rng, w.nx, w.ny, w.nz, 5, 6.0, 10.0, 48.0, 0.3, 0.8)
    else:
        raise ValueError(f"unknown workload {name!r}")
    return w


def scaled(w: Workload, **kw) -> Workload:
    return replace(w, **kw)
