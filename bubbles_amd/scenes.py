"""Synthetic scenes for the SPH / PCISPH hot path.

The reference builds scenes from shapes and emitters (e.g. the double dam
break at src/tests/test_pcisph3.cpp:1785-1856: a reversed ``MakeBox``
container with identity transform, ``UtilBuildGridForDomain`` for the grid and
``Setup(WaterDensity, spacing, scale, grid, set)``).  These helpers build the
same kind of scene from simple-cubic lattices so every run is reproducible:

* positions come from a SplitMix64 stream (per-axis uniform jitter), so the
  host, the oracle and the GPU see identical inputs;
* the reference computes particle mass for a BCC lattice (number density
  2/s^3, src/core/particle.h:557-582).  A simple-cubic block at spacing ``s``
  therefore uses ``targetSpacing = s * 2**(1/3)`` and
  ``relativeRadius = h / targetSpacing`` so it starts near rest density with
  kernel radius ``h``; no reference formula changes.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` outputs of SplitMix64 (Steele/Lea/Flood 2014) as uint64."""
    state = np.uint64(seed & MASK64)
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = state + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def shuffle_permutation(n: int, seed: int) -> np.ndarray:
    """Fisher-Yates permutation driven by SplitMix64 (seed + 0xA5A5)."""
    r = uniform01(seed + 0xA5A5, max(n, 1))
    perm = np.arange(n, dtype=np.int64)
    for i in range(n - 1, 0, -1):
        j = int(r[i] * (i + 1))
        perm[i], perm[j] = perm[j], perm[i]
    return perm


@dataclasses.dataclass
class Scene:
    """Everything a solver Setup needs, in reference units (FP64)."""

    name: str
    positions: np.ndarray          # (n, 3) float64, particle id order
    velocities: np.ndarray         # (n, 3) float64
    target_density: float
    target_spacing: float
    relative_radius: float
    tank_center: tuple
    tank_size: tuple
    grid_spacing: float
    grid_scale: float
    viscosity: float = 0.04
    dt: float = 5e-4
    lattice_spacing: float = 0.0
    kernel_radius: float = 0.0

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    def shuffled(self, seed: int) -> "Scene":
        perm = shuffle_permutation(self.n, seed)
        return dataclasses.replace(self, name=self.name + "-shuffled",
                                   positions=self.positions[perm].copy(),
                                   velocities=self.velocities[perm].copy())


def lattice_block(counts, spacing, origin, seed, jitter=1e-4) -> np.ndarray:
    """Simple-cubic block, x fastest, with per-axis U(-jitter, jitter)."""
    nx, ny, nz = counts
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    pts = pts * spacing + np.asarray(origin, dtype=np.float64)
    if jitter:
        u = uniform01(seed, pts.size).reshape(pts.shape)
        pts = pts + (2.0 * u - 1.0) * jitter
    return pts


def _scene(name, blocks, spacing, h, tank_size, seed, viscosity, dt, jitter=1e-4):
    target_spacing = spacing * 2.0 ** (1.0 / 3.0)
    radius = target_spacing                   # SetTargetSpacing sets the collision radius
    lo = -0.5 * np.asarray(tank_size, dtype=np.float64)
    parts = []
    for b, (counts, corner) in enumerate(blocks):
        origin = lo + radius + np.asarray(corner, dtype=np.float64)
        parts.append(lattice_block(counts, spacing, origin, seed + 7919 * b, jitter))
    pos = np.concatenate(parts, axis=0)
    return Scene(name=name, positions=pos, velocities=np.zeros_like(pos),
                 target_density=1000.0, target_spacing=target_spacing,
                 relative_radius=h / target_spacing, tank_center=(0.0, 0.0, 0.0),
                 tank_size=tuple(float(t) for t in tank_size), grid_spacing=h,
                 grid_scale=1.0, viscosity=viscosity, dt=dt, lattice_spacing=spacing,
                 kernel_radius=h)


def dam_break(nx=128, ny=64, nz=128, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=2,
              dt=5e-4, viscosity=0.04, jitter=1e-4) -> Scene:
    """PCISPH dam break: an ``nx*ny*nz`` block in the low corner of the tank,
    kernel radius 2*spacing, grid cell = kernel radius (1M particles at the
    defaults: 128x64x128 at 0.01 in a 2x1x2 tank, dt 5e-4)."""
    return _scene(f"dam_break_{nx}x{ny}x{nz}", [((nx, ny, nz), (0.0, 0.0, 0.0))],
                  spacing, 2.0 * spacing, tank, seed, viscosity, dt, jitter)


def sph_dam_break(seed=1) -> Scene:
    """WCSPH dam break: 32x16x32 block at 0.02 in a unit tank, h = 0.04,
    Advance(1e-3) per frame."""
    return _scene("sph_dam_break_32x16x32", [((32, 16, 32), (0.0, 0.0, 0.0))],
                  0.02, 0.04, (1.0, 1.0, 1.0), seed, 0.04, 1e-3)


def block_drop(scale=1.0, seed=3, spacing=0.005, dt=2.5e-4) -> Scene:
    """PCISPH block drop: a cube falling onto a resting layer.  At scale 1 and
    spacing 0.005 the cube is 64^3 and the layer 512x32x496 (2^23 particles)
    in a 4x2x4 tank; ``scale`` shrinks every count, and the tank follows
    scale * spacing / 0.005 so the layer always fits."""
    c = max(2, int(round(64 * scale)))
    lx, ly, lz = max(2, int(round(512 * scale))), max(2, int(round(32 * scale))), max(2, int(round(496 * scale)))
    f = scale * spacing / 0.005
    tank = (4.0 * f, 2.0 * f, 4.0 * f)
    layer_h = ly * spacing
    drop_y = layer_h + 0.4 * tank[1]
    cx = 0.5 * tank[0] - 0.5 * c * spacing
    cz = 0.5 * tank[2] - 0.5 * c * spacing
    blocks = [((lx, ly, lz), (0.0, 0.0, 0.0)), ((c, c, c), (cx, drop_y, cz))]
    return _scene(f"block_drop_{c}c", blocks, spacing, 2.0 * spacing, tank, seed, 0.01, dt)


def scene_summary(s: Scene) -> dict:
    return {"name": s.name, "n": s.n, "spacing": s.lattice_spacing,
            "kernel_radius": s.kernel_radius, "tank": list(s.tank_size), "dt": s.dt}


def expected_grid_dims(s: Scene):
    """Cell counts UtilBuildGridForDomain gives for the scene (ceil(extent/len))."""
    return tuple(int(math.ceil(t * (1.0 / (s.grid_spacing * s.grid_scale)))) for t in s.tank_size)
