"""Colliders on the device against the FP64 oracle: mesh SDF generation,
SDF sampling, the collider-set rule on scattered points, and PCISPH substeps
with obstacles.

Tolerances: SDF values are built in FP64 and stored in FP32, so a node
equals the oracle's value rounded to FP32, within one FP32 ulp (a node on
the surface may flip sign: the winding number there is 1/2 up to
round-off).  Collisions are FP32 arithmetic on
FP32 positions: 2e-5 absolute on positions and 2e-4 on velocities; points
whose FP64 distance to a collider is within 1e-5 of a decision threshold
(the particle radius, the inside test, a tie between colliders) may take the
other branch and are compared only on the rest.
"""
import dataclasses

import numpy as np
import pytest

from bubbles_amd import scenes, shapes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh", ["cube", "icosphere"])
def test_sdf_from_mesh_matches_oracle(bbx, oracle, mesh):
    if mesh == "cube":
        v, t = shapes.cube_mesh(0.4)
        v = shapes.transform_points(shapes.rotate_y(25) @ shapes.rotate_x(10), v)
    else:
        v, t = shapes.icosphere(0.25, subdivisions=3)
    dev = bbx.Sdf.from_mesh(v, t, dx=0.02, margin=0.15)
    g = dev.download()
    o = oracle.sdf_from_mesh(v, t, dx=0.02, margin=0.15, threads=16)
    assert g.nodes == o.nodes
    np.testing.assert_allclose(g.origin, o.origin, rtol=0, atol=0)
    np.testing.assert_allclose(g.spacing, o.spacing, rtol=0, atol=0)
    want = o.values.astype(np.float32).astype(np.float64)
    same_sign = np.sign(g.values) == np.sign(want)
    assert same_sign.mean() > 0.999 and np.all(np.abs(want[~same_sign]) <= 2e-5)
    # FP64 on both sides; the device's square root may differ in the last bit
    np.testing.assert_allclose(np.abs(g.values), np.abs(want), rtol=2.5e-7, atol=0)


def test_sdf_sample_and_gradient_match_oracle(bbx, oracle):
    v, t = shapes.icosphere(0.25, subdivisions=2)
    dev = bbx.Sdf.from_mesh(v, t, dx=0.02, margin=0.15)
    g = dev.download()
    rng = np.random.default_rng(0)
    pts = rng.uniform(-0.35, 0.35, size=(4000, 3))
    gv, gg = dev.sample(pts)
    ov, og = oracle.sdf_sample(g, pts.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(gv, ov, atol=2e-6)
    np.testing.assert_allclose(gg, og, atol=2e-4)


def _collider_set(bbx, mesh_sdf):
    tank = shapes.box(shapes.identity(), (1.0, 0.6, 1.0), reverse_orientation=True)
    obstacle = shapes.box(shapes.translate(-0.2, -0.1, 0.15) @ shapes.rotate_z(30) @ shapes.rotate_y(15),
                          (0.2, 0.1, 0.15), friction=0.2)
    ball = shapes.sphere(shapes.translate(0.2, 0.0, -0.2), 0.1)
    ball.linear_velocity = (0.0, 0.3, 0.0)
    ball.angular_velocity = (0.0, 0.0, 4.0)
    cv, ct = shapes.cube_mesh(0.16)
    cv = shapes.transform_points(shapes.translate(0.2, -0.1, 0.2) @ shapes.rotate_y(40), cv)
    mesh = shapes.mesh(cv, mesh_sdf(cv, ct))
    return [obstacle, ball, mesh, tank]


def test_collider_set_matches_oracle(bbx, oracle):
    devs = {}

    def dev_sdf(v, t):
        devs["sdf"] = bbx.Sdf.from_mesh(v, t, dx=0.01, margin=0.2)
        return devs["sdf"]

    gset = _collider_set(bbx, dev_sdf)
    host_grid = devs["sdf"].download()  # the oracle collides through the same FP32 field
    oset = [dataclasses.replace(c, sdf=host_grid) if c.kind == shapes.SHAPE_MESH else c for c in gset]
    rng = np.random.default_rng(1)
    n = 6000
    pos = rng.uniform([-0.52, -0.32, -0.52], [0.52, 0.32, 0.52], size=(n, 3)).astype(np.float32).astype(np.float64)
    vel = rng.normal(size=(n, 3)).astype(np.float32).astype(np.float64)
    radius, e = 0.012, 0.6
    gp, gv, gh = bbx.colliders_resolve(gset, radius, e, pos, vel)
    op, ov, oh = oracle.colliders_resolve(oset, radius, e, pos, vel)
    # FP64 decision margins: nearest-collider ties, the radius test, the inside test
    d = np.abs(np.array([[oracle.collider_query(c, p)["closest_distance"] for p in pos] for c in oset]))
    act = np.where(np.array([c.kind == shapes.SHAPE_MESH for c in oset])[:, None]
                   & ~np.all((pos >= np.array(oset[2].bounds_min)) & (pos <= np.array(oset[2].bounds_max)), axis=1)[None, :],
                   np.inf, d)
    srt = np.sort(act, axis=0)
    clear = (np.abs(srt[1] - srt[0]) > 1e-5) & (np.abs(srt[0] - radius) > 1e-5) & (srt[0] > 1e-5)
    ok = (gh == oh) & np.all(np.abs(gp - op) <= 2e-5, axis=1) & np.all(np.abs(gv - ov) <= 2e-4, axis=1)
    assert ok[clear].all(), np.nonzero(clear & ~ok)[0][:10]
    assert ok.mean() > 0.998 and clear.mean() > 0.99
    assert oh.sum() > 1000  # the set collides a good share of the points
    for kind in (shapes.SHAPE_BOX, shapes.SHAPE_SPHERE, shapes.SHAPE_MESH):
        sel = [i for i, c in enumerate(oset) if c.kind == kind]
        handled = np.isin(np.argmin(act, axis=0), sel) & (oh == 1)
        assert handled.sum() > 20, kind  # every kind handled some points


def _obstacle_scene():
    """A falling column (2 m/s down) over three obstacles and the tank."""
    sc = scenes.dam_break(24, 16, 24, spacing=0.01, tank=(0.6, 0.4, 0.6), seed=5)
    pos = sc.positions + np.array([0.12, 0.12, 0.12])
    vel = np.tile([0.0, -2.0, 0.0], (sc.n, 1))
    return dataclasses.replace(sc, positions=pos, velocities=vel)


def test_tank_as_a_collider_list_matches_the_tank_path(bbx):
    sc = _obstacle_scene()
    a = bbx.Solver(sc)
    b = bbx.Solver(sc)
    b.set_colliders([shapes.box(shapes.translate(*sc.tank_center), sc.tank_size, reverse_orientation=True)])
    for _ in range(10):
        sa = a.pcisph_substep(sc.dt, 5, 0.01, bbx.MODE_REFERENCE_EXACT)
        sb = b.pcisph_substep(sc.dt, 5, 0.01, bbx.MODE_REFERENCE_EXACT)
        assert sa.iterations == sb.iterations
    np.testing.assert_allclose(a.get("position"), b.get("position"), atol=1e-6)
    b.set_colliders([])  # back to the tank


def test_pcisph_with_obstacles_matches_oracle(bbx, oracle):
    sc = _obstacle_scene()
    ball = shapes.sphere(shapes.translate(0.0, -0.13, 0.0), 0.05)
    slab = shapes.box(shapes.translate(-0.1, -0.12, -0.1) @ shapes.rotate_z(20), (0.12, 0.03, 0.12), friction=0.1)
    tank = shapes.box(shapes.translate(*sc.tank_center), sc.tank_size, reverse_orientation=True)
    cv, ct = shapes.cube_mesh(0.06)
    cv = shapes.transform_points(shapes.translate(0.0, -0.14, -0.12) @ shapes.rotate_y(30), cv)
    sdf = bbx.Sdf.from_mesh(cv, ct, dx=0.004, margin=0.3)
    gset = [ball, slab, shapes.mesh(cv, sdf), tank]
    oset = [ball, slab, shapes.mesh(cv, sdf.download()), tank]
    g = bbx.Solver(sc)
    g.set_colliders(gset)
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")), threads=16)
    o.set_colliders(oset)
    deepest = 0.0
    for k in range(60):
        g.pcisph_substep(sc.dt, 5, 0.01, bbx.MODE_REFERENCE_EXACT)
        o.pcisph_substep(sc.dt, 5, 0.01, oracle.MODE_REFERENCE_EXACT)
        gp, op = g.get("position"), o.get("position")
        gv, ov = g.get("velocity"), o.get("velocity")
        perr = np.abs(gp - op).max(axis=1)
        verr = np.abs(gv - ov).max(axis=1)
        # a particle within round-off of a collision threshold may take the
        # other branch; it moves by at most about the particle radius
        assert (perr > 5e-5).sum() <= 2 and perr.max() < 0.02, (k, np.sort(perr)[-3:])
        assert (verr > 5e-3).sum() <= 2, (k, np.sort(verr)[-3:])
        o.set("position", gp)  # every substep is compared from the same state
        o.set("velocity", gv)
        deepest = min(deepest, float(gp[:, 1].min()))
    assert deepest < -0.12  # the column went down among the obstacles


def test_sph_with_obstacles_matches_oracle(bbx, oracle):
    """WCSPH (SphSolver3) with a sphere, a tilted box and the tank: the same
    collider rule in the integration pass."""
    sc = scenes.sph_dam_break()
    keep = sc.positions[:, 1] > -0.4  # the column's upper part, falling onto the obstacles
    sc = dataclasses.replace(sc, positions=sc.positions[keep].copy(), velocities=sc.velocities[keep].copy())
    sc = dataclasses.replace(sc, positions=sc.positions[:4096].copy(), velocities=np.tile([0.0, -2.0, 0.0], (4096, 1)))
    ball = shapes.sphere(shapes.translate(-0.3, -0.42, -0.3), 0.06)
    slab = shapes.box(shapes.translate(-0.1, -0.44, -0.1) @ shapes.rotate_x(15), (0.2, 0.04, 0.2), friction=0.2)
    tank = shapes.box(shapes.translate(*sc.tank_center), sc.tank_size, reverse_orientation=True)
    g = bbx.Solver(sc, kind=bbx.SPH)
    g.set_colliders([ball, slab, tank])
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")))
    o.set_colliders([ball, slab, tank])
    dt = 1.6e-4
    for k in range(40):
        g.sph_substep(dt)
        o.sph_substep(dt, oracle.SPH_TWO_PHASE)
        gp, op = g.get("position"), o.get("position")
        perr = np.abs(gp - op).max(axis=1)
        assert (perr > 5e-5).sum() <= 2 and perr.max() < 0.02, (k, np.sort(perr)[-3:])
        o.set("position", gp)
        o.set("velocity", g.get("velocity"))
    assert g.get("position")[:, 1].min() < -0.4


def test_append_into_a_scene_with_obstacles(bbx, oracle):
    """A growing set among obstacles: appended particles fall onto a sphere."""
    sc = _obstacle_scene()
    ball = shapes.sphere(shapes.translate(0.0, -0.13, 0.0), 0.05)
    tank = shapes.box(shapes.translate(*sc.tank_center), sc.tank_size, reverse_orientation=True)
    g = bbx.Solver(sc)
    g.set_colliders([ball, tank])
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")), threads=16)
    o.set_colliders([ball, tank])
    rng = np.random.default_rng(9)
    for k in range(30):
        if k % 10 == 0:
            p = (rng.uniform([-0.05, 0.1, -0.05], [0.05, 0.15, 0.05], size=(200, 3))
                 .astype(np.float32).astype(np.float64))
            v = np.tile([0.0, -2.0, 0.0], (200, 1))
            g.append(p, v)
            o.append(p, v)
        g.pcisph_substep(sc.dt, 5, 0.01, bbx.MODE_REFERENCE_EXACT)
        o.pcisph_substep(sc.dt, 5, 0.01, oracle.MODE_REFERENCE_EXACT)
        gp, op = g.get("position"), o.get("position")
        perr = np.abs(gp - op).max(axis=1)
        assert (perr > 5e-5).sum() <= 2 and perr.max() < 0.02, (k, np.sort(perr)[-3:])
        o.set("position", gp)
        o.set("velocity", g.get("velocity"))
    assert g.n == sc.n + 600
