"""The C++ facade (include/bubbles/bubbles.h) drives the same device path as
the C ABI: a scene written in the reference's style gives bit-identical
results to the same scene run through the ABI, and its Setup scalars and
grid equal the oracle's."""
import dataclasses
import json
import os
import subprocess

import numpy as np
import pytest

from bubbles_amd import build as hipbuild
from bubbles_amd import scenes, shapes


def facade_scene():
    return scenes.dam_break(16, 8, 16, spacing=0.01, tank=(0.5, 0.25, 0.5), jitter=0.0)


def test_facade_example_is_built():
    assert os.path.exists(hipbuild.FACADE_BIN), "run python bubbles_amd/build.py"


@pytest.mark.gpu
def test_facade_matches_abi_and_oracle(bbx, oracle):
    frame = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"bbx_facade_frame_{os.getpid()}.txt")
    res = subprocess.run([hipbuild.FACADE_BIN, "3", frame], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    sc = facade_scene()
    assert out["n"] == sc.n
    g = bbx.Solver(sc)
    n_sub = 0
    for _ in range(3):
        n_sub += g.pcisph_advance(1.0 / 240.0, 8, 0.01, bbx.MODE_TEXTBOOK)
    assert out["substeps"] == n_sub
    inf = g.info()
    assert out["mass"] == inf["mass"]
    assert out["delta_denom"] == inf["delta_denom"]
    assert out["grid"] == list(inf["grid"].usizes)
    pos = g.get("position")
    # sequential sums in id order, like the C++ loop
    assert out["sum_y"] == float(np.cumsum(pos[:, 1])[-1])
    assert out["sum_density"] == float(np.cumsum(g.get("density"))[-1])
    counts, ids = g.buckets()
    assert out["bucket0"] == ids[0, :counts[0]].tolist()
    o = oracle.Sim(sc)
    oi = o.info()
    assert out["mass"] == oi["mass"] and out["delta_denom"] == oi["delta_denom"]
    assert out["hash0"] == oracle.grid_hash(oi["grid"], pos[0])[0]
    # the frame SaveSphParticleSet wrote, read back through the library's loader
    assert out["reread"] == sc.n
    r = bbx.serializer_read(frame)
    os.remove(frame)
    assert r["flags"] == bbx.SER_POSITION | bbx.SER_VELOCITY | bbx.SER_DENSITY
    printed = np.array([[float("%g" % v) for v in row] for row in pos])
    np.testing.assert_allclose(r["position"], printed, rtol=1e-15, atol=0)
    np.testing.assert_allclose(r["density"], [float("%g" % v) for v in g.get("density")], rtol=1e-15)


def _emit_scene():
    """The emitter block of tests/cpp/facade_emit.cpp, in its AddParticle
    order, with the velocity field (0.5 x, -1, 0)."""
    s = 0.01
    pts = [(-0.03 + x * s, 0.05 + y * s, -0.03 + z * s)
           for z in range(6) for y in range(4) for x in range(6)]
    pos = np.array(pts, dtype=np.float64)
    return dataclasses.replace(facade_scene(), positions=pos, velocities=_emit_velocity(pos))


def _emit_velocity(p):
    p = np.asarray(p, dtype=np.float64).reshape(-1, 3)
    return np.stack([0.5 * p[:, 0], -np.ones(len(p)), np.zeros(len(p))], axis=1)


def _oracle_mirror(oracle, sc, g):
    """An oracle simulation holding the device solver's positions and chains
    (its MapGrid and refill rule then read the same state)."""
    pos = g.get("position")
    o = oracle.Sim(dataclasses.replace(sc, positions=pos, velocities=np.zeros_like(pos)), setup=False)
    o.set_chains(*g.chains())
    return o


@pytest.mark.gpu
def test_facade_continuous_emission_matches_abi(bbx, oracle):
    """The facade's MapGrid / MapGridEmit / MapGridEmitToOther scene against
    the same frames run through the ABI, with every refill decided by the
    oracle's emission rule (pinned in tests/test_oracle_emit.py) on the
    device's own state.  MapGridEmit returns at entry once the builder has
    warned about its capacity; MapGridEmitToOther (odd frames) does not."""
    frames, cap = 16, 576
    out_path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"bbx_emit_{os.getpid()}.bin")
    res = subprocess.run([hipbuild.EMIT_BIN, str(frames), str(cap), out_path],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    fpos = np.fromfile(out_path, dtype=np.float64).reshape(-1, 3)
    os.remove(out_path)
    sc = _emit_scene()
    g = bbx.Solver(sc)
    # The scene maps its cells before Setup, from the builder's FP64 host
    # positions (ascending-id chains of a full distribution).  Mapping the
    # device's FP32 copies instead moves lattice points that sit exactly one
    # spacing d from a chain member across the Distance < d test.
    mapped = oracle.Sim(sc).map_grid()
    assert out["mapped"] == len(mapped[0])
    warned, counts = False, []
    for f in range(frames):
        g.pcisph_advance(1.0 / 240.0, 8, 0.01, bbx.MODE_TEXTBOOK)
        if not (warned and f % 2 == 0):
            new, hit = _oracle_mirror(oracle, sc, g).map_grid_emit(mapped, 0.01, cap)
            warned = warned or hit
            if len(new):
                g.append(new, _emit_velocity(new))
        counts.append(g.n)
    assert out["counts"] == counts
    assert out["warned"] == int(warned)
    # the scene exercises emission and the capacity limit
    assert counts[-1] == cap and len(set(counts)) >= 3, counts
    np.testing.assert_array_equal(fpos, g.get("position"))


def _write_cube_obj(path):
    """The obstacle mesh of tests/cpp/facade_obstacles.cpp as an OBJ file
    (shortest round-trip decimals, so both sides read the same doubles)."""
    v, t = shapes.cube_mesh(0.06)
    v = v + np.array([0.06, -0.16, 0.06])
    with open(path, "w") as f:
        f.write("# cube obstacle\n")
        for p in v:
            f.write("v %r %r %r\n" % (float(p[0]), float(p[1]), float(p[2])))
        for a, b, c in t + 1:
            f.write("f %d %d %d\n" % (a, b, c))
    return v, t


@pytest.mark.gpu
def test_facade_obstacles_match_abi(bbx, tmp_path):
    obj, out_path = tmp_path / "block.obj", tmp_path / "pos.bin"
    v, t = _write_cube_obj(obj)
    frames = 12
    res = subprocess.run([hipbuild.OBSTACLES_BIN, str(frames), str(obj), str(out_path)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    fpos = np.fromfile(out_path, dtype=np.float64).reshape(-1, 3)
    assert out["mesh_triangles"] == 12
    np.testing.assert_array_equal(out["mesh_bounds"], list(v.min(axis=0)) + list(v.max(axis=0)))
    s = 0.01
    pos = np.array([(x * s + -0.08, y * s + -0.02, z * s + -0.08)
                    for z in range(16) for y in range(12) for x in range(16)])
    sc = dataclasses.replace(scenes.dam_break(16, 12, 16, spacing=0.01, tank=(0.5, 0.4, 0.5), jitter=0.0),
                             positions=pos, velocities=np.tile([0.0, -1.5, 0.0], (len(pos), 1)))
    assert out["n"] == sc.n
    ball = shapes.sphere(shapes.translate(0.0, -0.12, 0.0), 0.05)
    block = shapes.mesh(v, bbx.Sdf.from_mesh(v, t, dx=0.01, margin=0.1))  # GenerateShapeSDF defaults
    tank = shapes.box(shapes.identity(), (0.5, 0.4, 0.5), reverse_orientation=True)
    g = bbx.Solver(sc)
    subs = 0
    for f in range(frames):
        paddle = shapes.box(shapes.translate(-0.06, -0.15, -0.06) @ shapes.rotate_z(25 + 5 * f), (0.1, 0.03, 0.1),
                            friction=0.1)
        ball.active = f < 8
        g.set_colliders([ball, paddle, block, tank])
        subs += g.pcisph_advance(1.0 / 240.0, 8, 0.01, bbx.MODE_TEXTBOOK)
    assert out["substeps"] == subs
    # the facade composes and inverts transforms in its own order: the FP64
    # matrices may differ in the last bit before they are rounded to FP32
    np.testing.assert_allclose(fpos, g.get("position"), atol=1e-6)
    assert out["min_y"] < -0.09  # the column went down among the obstacles
