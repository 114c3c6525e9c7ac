"""Collider shapes on the host: affine transforms, the collider description
the C ABI takes (bbx_collider), SDF grids, and two closed meshes.

Mirrors the reference's shape constructors (MakeBox / MakeSphere / MakeMesh,
src/shapes/box.cpp:6-10, src/shapes/sphere.cpp:5-9, src/shapes/bvh.cpp:51-61)
and Transform helpers (Translate, Scale, RotateX/Y/Z, src/core/transform.h).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SHAPE_BOX, SHAPE_SPHERE, SHAPE_MESH, SHAPE_SDF = 0, 1, 2, 3


def identity() -> np.ndarray:
    return np.eye(4)


def translate(x, y=None, z=None) -> np.ndarray:
    if y is None:
        x, y, z = x
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def scale(sx, sy=None, sz=None) -> np.ndarray:
    if sy is None:
        sy = sz = sx
    return np.diag([sx, sy, sz, 1.0])


def _rot(axis: int, degrees: float) -> np.ndarray:
    t = math.radians(degrees)
    c, s = math.cos(t), math.sin(t)
    m = np.eye(4)
    a, b = [(1, 2), (2, 0), (0, 1)][axis]
    m[a, a], m[a, b], m[b, a], m[b, b] = c, -s, s, c
    return m


def rotate_x(deg):
    return _rot(0, deg)


def rotate_y(deg):
    return _rot(1, deg)


def rotate_z(deg):
    return _rot(2, deg)


@dataclasses.dataclass
class SdfGrid:
    """FieldGrid3f, vertex centred: values[k, j, i] at origin + spacing * (i, j, k)."""

    nodes: tuple
    origin: tuple
    spacing: tuple
    values: np.ndarray  # float64, flat, x fastest

    def flat(self) -> np.ndarray:
        return np.ascontiguousarray(self.values, dtype=np.float64).reshape(-1)


@dataclasses.dataclass
class Collider:
    kind: int
    to_world: np.ndarray = dataclasses.field(default_factory=identity)
    reverse_orientation: bool = False
    active: bool = True
    friction: float = 0.0
    size: tuple = (0.0, 0.0, 0.0)
    radius: float = 0.0
    linear_velocity: tuple = (0.0, 0.0, 0.0)
    angular_velocity: tuple = (0.0, 0.0, 0.0)
    bounds_min: tuple = (0.0, 0.0, 0.0)
    bounds_max: tuple = (0.0, 0.0, 0.0)
    sdf: object = None  # SdfGrid, or a device SDF (bubbles_amd.abi.Sdf)

    @property
    def to_object(self) -> np.ndarray:
        return np.linalg.inv(self.to_world)


def box(to_world, size, reverse_orientation=False, friction=0.0) -> Collider:
    return Collider(SHAPE_BOX, np.asarray(to_world, dtype=np.float64), reverse_orientation,
                    size=tuple(float(s) for s in size), friction=friction)


def sphere(to_world, radius, reverse_orientation=False, friction=0.0) -> Collider:
    return Collider(SHAPE_SPHERE, np.asarray(to_world, dtype=np.float64), reverse_orientation,
                    radius=float(radius), friction=friction)


def mesh(vertices, sdf, reverse_orientation=False, friction=0.0) -> Collider:
    """A mesh collider (vertices already in world space, as MakeMesh bakes its
    transform into them) colliding through its SDF."""
    v = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    return Collider(SHAPE_MESH, identity(), reverse_orientation, friction=friction,
                    bounds_min=tuple(v.min(axis=0)), bounds_max=tuple(v.max(axis=0)), sdf=sdf)


def sdf_shape(sdf, to_world=None, friction=0.0) -> Collider:
    return Collider(SHAPE_SDF, identity() if to_world is None else np.asarray(to_world, np.float64),
                    friction=friction, sdf=sdf)


def transform_points(m, pts) -> np.ndarray:
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    return p @ np.asarray(m)[:3, :3].T + np.asarray(m)[:3, 3]


def cube_mesh(size=1.0):
    """A closed, outward-oriented cube centred at the origin: 8 vertices, 12 triangles."""
    h = size / 2.0
    v = np.array([[x, y, z] for z in (-h, h) for y in (-h, h) for x in (-h, h)], dtype=np.float64)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tris = []
    for a, b, c, d in quads:
        tris += [(a, b, c), (a, c, d)]
    return v, np.array(tris, dtype=np.int32)


def icosphere(radius=1.0, subdivisions=2):
    """A closed, outward-oriented icosphere (20 * 4^subdivisions triangles)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(verts) * radius, np.array(f, dtype=np.int32)
