"""ctypes binding of ``libbubbles_amd.so`` (the C ABI in include/bubbles_amd.h).

This is plumbing for the test-suite and the benchmark: it loads the in-tree
gfx950 library and fails loudly when it is missing -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libbubbles_amd.so")
HEADER = os.path.join(ROOT, "include", "bubbles_amd.h")

OK = 0
PCISPH, SPH = 0, 1
MODE_REFERENCE_EXACT, MODE_TEXTBOOK = 0, 1
WARM_NONE, WARM_MLP, WARM_AUTO = 0, 1, 2
WARM_AUTO_MIN_ITERATIONS = 3
PROFILE_ALL_LAUNCHES = -1000
BUCKET = 100
NFEAT = 16

FIELD = {"position": 0, "velocity": 1, "force": 2, "density": 3, "pressure": 4,
         "density_predicted": 5, "pressure_force": 6, "temp_position": 7, "features": 8,
         "warm_pressure": 9}
FIELD_WIDTH = {0: 3, 1: 3, 2: 3, 3: 1, 4: 1, 5: 1, 6: 3, 7: 3, 8: 16, 9: 1}


class BbxError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("target_density", C.c_double), ("target_spacing", C.c_double),
                ("relative_radius", C.c_double), ("tank_center", C.c_double * 3),
                ("tank_size", C.c_double * 3), ("grid_spacing", C.c_double),
                ("grid_scale", C.c_double), ("eos_exponent", C.c_double),
                ("negative_pressure_scale", C.c_double), ("viscosity", C.c_double),
                ("pseudo_viscosity", C.c_double), ("sound_speed", C.c_double),
                ("drag_coefficient", C.c_double), ("gravity", C.c_double * 3)]


class GridDesc(C.Structure):
    _fields_ = [("usizes", C.c_uint * 3), ("total", C.c_uint), ("cells_len", C.c_double * 3),
                ("min_point", C.c_double * 3), ("bounds_min", C.c_double * 3),
                ("bounds_max", C.c_double * 3)]


class Info(C.Structure):
    _fields_ = [("count", C.c_int), ("mass", C.c_double), ("kernel_radius", C.c_double),
                ("radius", C.c_double), ("delta_denom", C.c_double), ("grid", GridDesc)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("full_redistribution", C.c_int),
                ("bucket_overflows", C.c_int), ("lost_particles", C.c_int),
                ("max_density_error", C.c_double), ("gpu_ms", C.c_double),
                ("warm_started", C.c_int), ("pad", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BbxCollider(C.Structure):
    _fields_ = [("kind", C.c_int), ("reverse_orientation", C.c_int), ("active", C.c_int), ("pad", C.c_int),
                ("friction", C.c_double), ("to_world", C.c_double * 12), ("to_object", C.c_double * 12),
                ("size", C.c_double * 3), ("radius", C.c_double), ("linear_velocity", C.c_double * 3),
                ("angular_velocity", C.c_double * 3), ("bounds_min", C.c_double * 3),
                ("bounds_max", C.c_double * 3), ("sdf", C.c_void_p)]


_lib = None


def header_symbols(path: str = HEADER) -> list:
    """Function names declared in the public header."""
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bbx_\w+)\s*\(", text, re.M)))


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BbxError(f"{LIB_PATH} is missing: run `python bubbles_amd/build.py` (no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, u = C.c_void_p, C.c_int, C.c_double, C.c_uint
    dp, ip, fp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_float)
    sig = {
        "bbx_last_error": (C.c_char_p, []),
        "bbx_version": (i, [ip, ip]),
        "bbx_device_count": (i, [ip]),
        "bbx_default_params": (i, [C.POINTER(Params)]),
        "bbx_solver_create": (i, [i, C.POINTER(Params), dp, dp, i, C.POINTER(vp)]),
        "bbx_solver_destroy": (i, [vp]),
        "bbx_solver_setup": (i, [vp]),
        "bbx_solver_info": (i, [vp, C.POINTER(Info)]),
        "bbx_solver_download": (i, [vp, i, dp]),
        "bbx_solver_upload": (i, [vp, i, dp]),
        "bbx_solver_get_flag": (i, [vp, ip]),
        "bbx_solver_set_flag": (i, [vp, i]),
        "bbx_grid_distribute": (i, [vp, C.POINTER(Stats)]),
        "bbx_grid_get_chains": (i, [vp, ip, ip]),
        "bbx_grid_set_chains": (i, [vp, ip, ip]),
        "bbx_grid_hash": (i, [vp, dp, i, ip]),
        "bbx_grid_cell_neighbors": (i, [vp, i, ip, ip]),
        "bbx_grid_describe": (i, [C.POINTER(Params), C.POINTER(GridDesc)]),
        "bbx_grid_hash_desc": (i, [C.POINTER(GridDesc), dp, i, ip, ip]),
        "bbx_neighbors_get": (i, [vp, ip, ip]),
        "bbx_stage_density": (i, [vp]),
        "bbx_stage_nonpressure": (i, [vp]),
        "bbx_pcisph_substep": (i, [vp, d, i, d, i, i, C.POINTER(Stats)]),
        "bbx_pcisph_advance": (i, [vp, d, i, d, i, i, ip]),
        "bbx_sph_substep": (i, [vp, d, C.POINTER(Stats)]),
        "bbx_sph_advance": (i, [vp, d, ip]),
        "bbx_pcisph_delta": (i, [vp, d, dp]),
        "bbx_number_of_time_steps": (i, [vp, d, d, C.POINTER(u)]),
        "bbx_mlp_create": (i, [i, C.POINTER(vp)]),
        "bbx_mlp_destroy": (i, [vp]),
        "bbx_mlp_param_count": (i, [ip]),
        "bbx_mlp_get_params": (i, [vp, fp]),
        "bbx_mlp_set_params": (i, [vp, fp]),
        "bbx_mlp_infer": (i, [vp, fp, i, fp]),
        "bbx_mlp_set_expectile": (i, [vp, C.c_float]),
        "bbx_mlp_gradients": (i, [vp, fp, fp, i, fp, fp]),
        "bbx_mlp_train_step": (i, [vp, fp, fp, i, C.c_float, fp]),
        "bbx_mlp_train": (i, [vp, fp, fp, i, i, i, C.c_float, i, fp]),
        "bbx_solver_keep_features": (i, [vp, i]),
        "bbx_pcisph_run": (i, [vp, d, i, i, d, i, i, C.POINTER(Stats)]),
        "bbx_solver_attach_mlp": (i, [vp, vp]),
        "bbx_pcisph_harvest": (i, [vp, d, i, d, fp, fp, C.POINTER(Stats)]),
        "bbx_profile_enable": (i, [vp, i]),
        "bbx_profile_reset": (i, [vp]),
        "bbx_profile_select": (i, [vp, C.c_char_p, i]),
        "bbx_profile_get": (i, [vp, C.c_char_p, i, dp, ip, i, ip]),
        "bbx_solver_synchronize": (i, [vp]),
        "bbx_stream_copy": (i, [d, i, dp, dp]),
        "bbx_serializer_write": (i, [C.c_char_p, i, d, i, dp, dp, dp, d, ip]),
        "bbx_solver_save": (i, [vp, C.c_char_p, i]),
        "bbx_solver_reserve": (i, [vp, i]),
        "bbx_solver_capacity": (i, [vp, ip]),
        "bbx_solver_append": (i, [vp, dp, dp, i]),
        "bbx_serializer_read": (i, [C.c_char_p, i, ip, ip, dp, dp, dp, dp, ip]),
        "bbx_serializer_parse_number": (i, [C.c_char_p, dp]),
        "bbx_sdf_create": (i, [ip, dp, dp, dp, C.POINTER(vp)]),
        "bbx_sdf_from_mesh": (i, [dp, i, ip, i, d, d, C.POINTER(vp)]),
        "bbx_sdf_describe": (i, [vp, ip, dp, dp]),
        "bbx_sdf_download": (i, [vp, dp]),
        "bbx_sdf_sample": (i, [vp, dp, i, dp, dp]),
        "bbx_sdf_destroy": (None, [vp]),
        "bbx_solver_set_colliders": (i, [vp, C.POINTER(BbxCollider), i]),
        "bbx_colliders_resolve": (i, [C.POINTER(BbxCollider), i, d, d, dp, dp, i, ip]),
        "bbx_colliders_resolve_noalloc": (i, [C.POINTER(BbxCollider), i, d, d, dp, dp, i, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, what: str = ""):
    if rc != OK:
        msg = load().bbx_last_error().decode(errors="replace")
        raise BbxError(f"{what} failed ({rc}): {msg}")


def device_count() -> int:
    n = C.c_int(0)
    check(load().bbx_device_count(C.byref(n)), "bbx_device_count")
    return n.value


def stream_copy(nbytes: float, iters: int = 20):
    """(GB/s, avg ms) of a device copy of nbytes (read + write)."""
    g, ms = C.c_double(), C.c_double()
    check(load().bbx_stream_copy(float(nbytes), iters, C.byref(g), C.byref(ms)), "bbx_stream_copy")
    return g.value, ms.value


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# serializer flags (SERIALIZER_* of the reference's serializer.h)
SER_POSITION, SER_VELOCITY, SER_DENSITY, SER_BOUNDARY = 0x01, 0x02, 0x04, 0x08
SER_NORMAL, SER_MASS, SER_LAYERS, SER_XYZ, SER_BOUNDARY_EXCLUSIVE = 0x10, 0x20, 0x40, 0x80, 0x100


def serializer_write(path, flags, spacing, positions=None, velocities=None, densities=None,
                     mass=0.0, boundary=None):
    """Append one FluidBegin block (the reference's SaveSphParticleSet format)."""
    arrs = [positions, velocities, densities]
    n = next((len(a) for a in arrs + [boundary] if a is not None), 0)
    pos = np.ascontiguousarray(positions if positions is not None else np.zeros((n, 3)), dtype=np.float64)
    vel = np.ascontiguousarray(velocities if velocities is not None else np.zeros((n, 3)), dtype=np.float64)
    den = np.ascontiguousarray(densities if densities is not None else np.zeros(n), dtype=np.float64)
    bnd = None if boundary is None else np.ascontiguousarray(boundary, dtype=np.int32)
    check(load().bbx_serializer_write(os.fsencode(path), flags, float(spacing), n, _dp(pos), _dp(vel),
                                      _dp(den), float(mass), _ip(bnd) if bnd is not None else None),
          "bbx_serializer_write")


def serializer_read(path, default_flags=SER_POSITION):
    """First FluidBegin block: dict of flags and per-particle arrays."""
    n, fl = C.c_int(), C.c_int(default_flags)
    check(load().bbx_serializer_read(os.fsencode(path), -1, C.byref(n), C.byref(fl), None, None, None, None,
                                     None), "bbx_serializer_read")
    cap = n.value
    pos, vel = np.zeros((cap, 3)), np.zeros((cap, 3))
    den, mas = np.zeros(cap), np.zeros(cap)
    bnd = np.zeros(cap, dtype=np.int32)
    fl = C.c_int(default_flags)
    check(load().bbx_serializer_read(os.fsencode(path), cap, C.byref(n), C.byref(fl), _dp(pos), _dp(vel),
                                     _dp(den), _dp(mas), _ip(bnd)), "bbx_serializer_read")
    return {"flags": fl.value, "position": pos, "velocity": vel, "density": den, "mass": mas,
            "boundary": bnd}


def serializer_parse_number(text: str) -> float:
    """The reference's ParseFloat of one token."""
    v = C.c_double()
    check(load().bbx_serializer_parse_number(text.encode(), C.byref(v)), "bbx_serializer_parse_number")
    return v.value


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Sdf:
    """A FieldGrid3f on the device (bbx_sdf)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_grid(cls, grid):
        """Upload a shapes.SdfGrid."""
        h = C.c_void_p()
        nodes = np.ascontiguousarray(grid.nodes, dtype=np.int32)
        org = np.ascontiguousarray(grid.origin, dtype=np.float64)
        spc = np.ascontiguousarray(grid.spacing, dtype=np.float64)
        vals = grid.flat()
        check(load().bbx_sdf_create(_ip(nodes), _dp(org), _dp(spc), _dp(vals), C.byref(h)), "bbx_sdf_create")
        return cls(h)

    @classmethod
    def from_mesh(cls, vertices, triangles, dx=0.01, margin=0.1):
        """GenerateShapeSDF of a closed triangle mesh, built on the device."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
        h = C.c_void_p()
        check(load().bbx_sdf_from_mesh(_dp(v), v.shape[0], _ip(t), t.shape[0], float(dx), float(margin),
                                       C.byref(h)), "bbx_sdf_from_mesh")
        return cls(h)

    def describe(self):
        nodes = np.zeros(3, dtype=np.int32)
        org, spc = np.zeros(3), np.zeros(3)
        check(load().bbx_sdf_describe(self.h, _ip(nodes), _dp(org), _dp(spc)), "bbx_sdf_describe")
        return tuple(int(x) for x in nodes), tuple(org), tuple(spc)

    def download(self):
        from bubbles_amd.shapes import SdfGrid

        nodes, org, spc = self.describe()
        vals = np.zeros(int(np.prod(nodes)))
        check(load().bbx_sdf_download(self.h, _dp(vals)), "bbx_sdf_download")
        return SdfGrid(nodes, org, spc, vals)

    def sample(self, points):
        """(values, gradients) of FieldGrid::Sample / ::Gradient on the device."""
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        vals, grads = np.zeros(p.shape[0]), np.zeros((p.shape[0], 3))
        check(load().bbx_sdf_sample(self.h, _dp(p), p.shape[0], _dp(vals), _dp(grads)), "bbx_sdf_sample")
        return vals, grads

    def close(self):
        if self.h:
            load().bbx_sdf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def collider_structs(colliders):
    """bbx_collider array for a list of shapes.Collider; host SDF grids are
    uploaded (the returned keep-alive list holds the device copies)."""
    arr = (BbxCollider * max(1, len(colliders)))()
    keep = []
    for k, c in enumerate(colliders):
        b = arr[k]
        b.kind, b.reverse_orientation, b.active = int(c.kind), int(bool(c.reverse_orientation)), int(bool(c.active))
        b.friction = float(c.friction)
        tw = np.asarray(c.to_world, dtype=np.float64)
        to = np.linalg.inv(tw)
        b.to_world[:] = list(tw[:3, :].reshape(-1))
        b.to_object[:] = list(to[:3, :].reshape(-1))
        b.size[:] = list(c.size)
        b.radius = float(c.radius)
        b.linear_velocity[:] = list(c.linear_velocity)
        b.angular_velocity[:] = list(c.angular_velocity)
        b.bounds_min[:] = list(c.bounds_min)
        b.bounds_max[:] = list(c.bounds_max)
        if c.sdf is not None:
            dev = c.sdf if isinstance(c.sdf, Sdf) else Sdf.from_grid(c.sdf)
            keep.append(dev)
            b.sdf = dev.h
    return arr, keep


def colliders_resolve(colliders, radius, restitution, positions, velocities):
    """ColliderSet3::ResolveCollision on the device for every point:
    (positions, velocities, collided)."""
    arr, keep = collider_structs(colliders)
    p = np.array(positions, dtype=np.float64).reshape(-1, 3)
    v = np.array(velocities, dtype=np.float64).reshape(-1, 3)
    hit = np.zeros(p.shape[0], dtype=np.int32)
    check(load().bbx_colliders_resolve(arr, len(colliders), float(radius), float(restitution), _dp(p), _dp(v),
                                       p.shape[0], _ip(hit)), "bbx_colliders_resolve")
    del keep
    return p, v, hit


NOALLOC_MAX_POINTS = 65536  # BBX_NOALLOC_MAX_POINTS


def colliders_resolve_noalloc(colliders, radius, restitution, positions, velocities):
    """The same as colliders_resolve through the entry point that allocates
    no device memory (at most NOALLOC_MAX_POINTS points per call)."""
    arr, keep = collider_structs(colliders)
    p = np.array(positions, dtype=np.float64).reshape(-1, 3)
    v = np.array(velocities, dtype=np.float64).reshape(-1, 3)
    hit = np.zeros(p.shape[0], dtype=np.int32)
    check(load().bbx_colliders_resolve_noalloc(arr, len(colliders), float(radius), float(restitution), _dp(p),
                                               _dp(v), p.shape[0], _ip(hit)), "bbx_colliders_resolve_noalloc")
    del keep
    return p, v, hit


def params_for(scene, **over) -> Params:
    p = Params()
    check(load().bbx_default_params(C.byref(p)), "bbx_default_params")
    p.target_density = scene.target_density
    p.target_spacing = scene.target_spacing
    p.relative_radius = scene.relative_radius
    p.tank_center[:] = list(scene.tank_center)
    p.tank_size[:] = list(scene.tank_size)
    p.grid_spacing = scene.grid_spacing
    p.grid_scale = scene.grid_scale
    p.viscosity = scene.viscosity
    for k, v in over.items():
        if k == "gravity":
            p.gravity[:] = list(v)
        else:
            setattr(p, k, v)
    return p


class Solver:
    """One device solver (PciSphSolver3 or SphSolver3 over the C ABI)."""

    def __init__(self, scene, kind=PCISPH, setup=True, **over):
        L = load()
        self.scene = scene
        self.params = params_for(scene, **over)
        pos = np.ascontiguousarray(scene.positions, dtype=np.float64)
        vel = np.ascontiguousarray(scene.velocities, dtype=np.float64)
        self.n = int(pos.shape[0])
        h = C.c_void_p()
        check(L.bbx_solver_create(kind, C.byref(self.params), _dp(pos), _dp(vel), self.n,
                                  C.byref(h)), "bbx_solver_create")
        self.h = h
        self.kind = kind
        if setup:
            self.setup()

    def close(self):
        if getattr(self, "h", None):
            load().bbx_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self):
        check(load().bbx_solver_setup(self.h), "bbx_solver_setup")

    def info(self) -> dict:
        inf = Info()
        check(load().bbx_solver_info(self.h, C.byref(inf)), "bbx_solver_info")
        return {"count": inf.count, "mass": inf.mass, "kernel_radius": inf.kernel_radius,
                "radius": inf.radius, "delta_denom": inf.delta_denom, "grid": inf.grid}

    def reserve(self, capacity: int):
        check(load().bbx_solver_reserve(self.h, int(capacity)), "bbx_solver_reserve")

    def capacity(self) -> int:
        c = C.c_int()
        check(load().bbx_solver_capacity(self.h, C.byref(c)), "bbx_solver_capacity")
        return c.value

    def append(self, positions, velocities=None):
        """Add particles (ids n, n+1, ...) at the tails of their cells' chains."""
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        v = None if velocities is None else np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1, 3)
        check(load().bbx_solver_append(self.h, _dp(p), _dp(v) if v is not None else None, p.shape[0]),
              "bbx_solver_append")
        self.n += p.shape[0]

    def set_colliders(self, colliders):
        """Replace the tank by a collider list (empty: back to the tank)."""
        arr, keep = collider_structs(colliders)
        check(load().bbx_solver_set_colliders(self.h, arr, len(colliders)), "bbx_solver_set_colliders")
        self._collider_keep = keep  # device SDFs must outlive the list

    def save(self, path, flags=SER_POSITION | SER_VELOCITY | SER_DENSITY):
        """Append this solver's particles as one serializer block."""
        check(load().bbx_solver_save(self.h, os.fsencode(path), flags), "bbx_solver_save")

    def get(self, field: str) -> np.ndarray:
        f = FIELD[field]
        w = FIELD_WIDTH[f]
        out = np.zeros((self.n, w) if w > 1 else (self.n,), dtype=np.float64)
        check(load().bbx_solver_download(self.h, f, _dp(out)), "bbx_solver_download")
        return out

    def set(self, field: str, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        check(load().bbx_solver_upload(self.h, FIELD[field], _dp(a)), "bbx_solver_upload")

    @property
    def flag(self) -> int:
        v = C.c_int()
        check(load().bbx_solver_get_flag(self.h, C.byref(v)), "bbx_solver_get_flag")
        return v.value

    @flag.setter
    def flag(self, v):
        check(load().bbx_solver_set_flag(self.h, int(v)), "bbx_solver_set_flag")

    def distribute(self) -> Stats:
        st = Stats()
        check(load().bbx_grid_distribute(self.h, C.byref(st)), "bbx_grid_distribute")
        return st

    def chains(self):
        total = self.info()["grid"].total
        start = np.zeros(total + 1, dtype=np.int32)
        members = np.zeros(max(self.n, 1), dtype=np.int32)
        check(load().bbx_grid_get_chains(self.h, _ip(start), _ip(members)), "bbx_grid_get_chains")
        return start, members[: start[-1]]

    def set_chains(self, start, members):
        s = np.ascontiguousarray(start, dtype=np.int32)
        m = np.ascontiguousarray(members, dtype=np.int32)
        if m.size == 0:
            m = np.zeros(1, dtype=np.int32)
        check(load().bbx_grid_set_chains(self.h, _ip(s), _ip(m)), "bbx_grid_set_chains")

    def bucket_counts(self) -> np.ndarray:
        counts = np.zeros(self.n, dtype=np.int32)
        check(load().bbx_neighbors_get(self.h, _ip(counts), None), "bbx_neighbors_get")
        return counts

    def buckets(self):
        counts = np.zeros(self.n, dtype=np.int32)
        ids = np.zeros((self.n, BUCKET), dtype=np.int32)
        check(load().bbx_neighbors_get(self.h, _ip(counts), _ip(ids)), "bbx_neighbors_get")
        return counts, ids

    def hash(self, points) -> np.ndarray:
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(p.shape[0], dtype=np.int32)
        check(load().bbx_grid_hash(self.h, _dp(p), p.shape[0], _ip(out)), "bbx_grid_hash")
        return out

    def cell_neighbors(self, cell: int):
        out = (C.c_int * 27)()
        cnt = C.c_int()
        check(load().bbx_grid_cell_neighbors(self.h, cell, out, C.byref(cnt)), "bbx_grid_cell_neighbors")
        return list(out[: cnt.value])

    def stage_density(self):
        check(load().bbx_stage_density(self.h), "bbx_stage_density")

    def stage_nonpressure(self):
        check(load().bbx_stage_nonpressure(self.h), "bbx_stage_nonpressure")

    def delta(self, dt: float) -> float:
        v = C.c_double()
        check(load().bbx_pcisph_delta(self.h, dt, C.byref(v)), "bbx_pcisph_delta")
        return v.value

    def number_of_time_steps(self, interval, scale) -> int:
        v = C.c_uint()
        check(load().bbx_number_of_time_steps(self.h, interval, scale, C.byref(v)),
              "bbx_number_of_time_steps")
        return v.value

    def pcisph_substep(self, dt, max_it=5, tol=0.01, mode=MODE_REFERENCE_EXACT, warm=WARM_NONE):
        st = Stats()
        check(load().bbx_pcisph_substep(self.h, dt, max_it, tol, mode, warm, C.byref(st)),
              "bbx_pcisph_substep")
        return st

    def pcisph_run(self, dt, count, max_it=5, tol=0.01, mode=MODE_REFERENCE_EXACT, warm=WARM_NONE):
        """`count` substeps of fixed dt in one call: a list of Stats."""
        st = (Stats * max(count, 1))()
        check(load().bbx_pcisph_run(self.h, dt, count, max_it, tol, mode, warm, st), "bbx_pcisph_run")
        return list(st[:count])

    def pcisph_advance(self, interval, max_it=5, tol=0.01, mode=MODE_REFERENCE_EXACT, warm=WARM_NONE):
        n = C.c_int()
        check(load().bbx_pcisph_advance(self.h, interval, max_it, tol, mode, warm, C.byref(n)),
              "bbx_pcisph_advance")
        return n.value

    def sph_substep(self, dt):
        st = Stats()
        check(load().bbx_sph_substep(self.h, dt, C.byref(st)), "bbx_sph_substep")
        return st

    def sph_advance(self, interval):
        n = C.c_int()
        check(load().bbx_sph_advance(self.h, interval, C.byref(n)), "bbx_sph_advance")
        return n.value

    def attach_mlp(self, mlp):
        check(load().bbx_solver_attach_mlp(self.h, mlp.h if mlp else None), "bbx_solver_attach_mlp")
        self._mlp = mlp  # the network object must outlive its use by this solver

    def keep_features(self, keep=True):
        """Store the features of warm-started substeps (get("features"))."""
        check(load().bbx_solver_keep_features(self.h, 1 if keep else 0), "bbx_solver_keep_features")

    def harvest(self, dt, max_it=8, tol=0.01):
        feats = np.zeros((self.n, NFEAT), dtype=np.float32)
        targ = np.zeros(self.n, dtype=np.float32)
        st = Stats()
        check(load().bbx_pcisph_harvest(self.h, dt, max_it, tol, _fp(feats), _fp(targ), C.byref(st)),
              "bbx_pcisph_harvest")
        return feats, targ, st

    def profile(self, enable=True):
        check(load().bbx_profile_enable(self.h, 1 if enable else 0), "bbx_profile_enable")

    def profile_select(self, kernel=None, iteration=None):
        """Time only `kernel` (None: all) and only its launch of `iteration`
        (None: every launch; -1: the warm-start launch)."""
        it = PROFILE_ALL_LAUNCHES if iteration is None else int(iteration)
        check(load().bbx_profile_select(self.h, kernel.encode() if kernel else None, it),
              "bbx_profile_select")

    def profile_reset(self):
        check(load().bbx_profile_reset(self.h), "bbx_profile_reset")

    def profile_get(self) -> dict:
        cap = 64
        names = C.create_string_buffer(8192)
        ms = np.zeros(cap, dtype=np.float64)
        cnt = np.zeros(cap, dtype=np.int32)
        k = C.c_int()
        check(load().bbx_profile_get(self.h, names, 8192, _dp(ms), _ip(cnt), cap, C.byref(k)),
              "bbx_profile_get")
        nm = names.value.decode().split("\n")[: k.value]
        return {n: {"total_ms": float(ms[i]), "launches": int(cnt[i])}
                for i, n in enumerate(nm) if cnt[i] > 0}

    def synchronize(self):
        check(load().bbx_solver_synchronize(self.h), "bbx_solver_synchronize")


class Mlp:
    """Warm-start network (16 -> 64 -> 64 -> 1, GELU, BF16 MFMA)."""

    def __init__(self, seed=0):
        h = C.c_void_p()
        check(load().bbx_mlp_create(seed, C.byref(h)), "bbx_mlp_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().bbx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def param_count() -> int:
        n = C.c_int()
        check(load().bbx_mlp_param_count(C.byref(n)), "bbx_mlp_param_count")
        return n.value

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.param_count(), dtype=np.float32)
        check(load().bbx_mlp_get_params(self.h, _fp(out)), "bbx_mlp_get_params")
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        check(load().bbx_mlp_set_params(self.h, _fp(a)), "bbx_mlp_set_params")

    def set_expectile(self, tau: float):
        """Training loss: expectile tau (0.5 = mean squared error)."""
        check(load().bbx_mlp_set_expectile(self.h, float(tau)), "bbx_mlp_set_expectile")

    def infer(self, x) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(a.shape[0], dtype=np.float32)
        check(load().bbx_mlp_infer(self.h, _fp(a), a.shape[0], _fp(y)), "bbx_mlp_infer")
        return y

    def gradients(self, x, t):
        a = np.ascontiguousarray(x, dtype=np.float32)
        b = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.param_count(), dtype=np.float32)
        loss = C.c_float()
        check(load().bbx_mlp_gradients(self.h, _fp(a), _fp(b), a.shape[0], _fp(g), C.byref(loss)),
              "bbx_mlp_gradients")
        return g, loss.value

    def train_step(self, x, t, lr=1e-3) -> float:
        a = np.ascontiguousarray(x, dtype=np.float32)
        b = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        check(load().bbx_mlp_train_step(self.h, _fp(a), _fp(b), a.shape[0], lr, C.byref(loss)),
              "bbx_mlp_train_step")
        return loss.value

    def train(self, x, t, steps, batch=4096, lr=1e-3, seed=0) -> float:
        a = np.ascontiguousarray(x, dtype=np.float32)
        b = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        check(load().bbx_mlp_train(self.h, _fp(a), _fp(b), a.shape[0], steps, batch, lr, seed,
                                   C.byref(loss)), "bbx_mlp_train")
        return loss.value
