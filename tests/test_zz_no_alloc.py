"""bbx_colliders_resolve_noalloc (staging in a static device array, no
device allocation) against bbx_colliders_resolve on the same inputs:
bit-identical positions, velocities and collision flags, on two calls in a
row with the second one smaller, then at the size limit, and a clear range
error beyond it."""
import numpy as np
import pytest

from bubbles_amd import shapes

pytestmark = pytest.mark.gpu


def _colliders(bbx):
    tank = shapes.box(shapes.identity(), (1.0, 0.6, 1.0), reverse_orientation=True)
    obstacle = shapes.box(shapes.translate(-0.2, -0.1, 0.15) @ shapes.rotate_z(30) @ shapes.rotate_y(15),
                          (0.2, 0.1, 0.15), friction=0.2)
    ball = shapes.sphere(shapes.translate(0.2, 0.0, -0.2), 0.1)
    ball.linear_velocity = (0.0, 0.3, 0.0)
    ball.angular_velocity = (0.0, 0.0, 4.0)
    cv, ct = shapes.cube_mesh(0.16)
    cv = shapes.transform_points(shapes.translate(0.2, -0.1, 0.2) @ shapes.rotate_y(40), cv)
    sdf = bbx.Sdf.from_mesh(cv, ct, dx=0.01, margin=0.2)
    return [obstacle, ball, shapes.mesh(cv, sdf), tank], sdf


def _points(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform([-0.52, -0.32, -0.52], [0.52, 0.32, 0.52], size=(n, 3))
    vel = rng.normal(size=(n, 3))
    return pos, vel


def test_noalloc_resolve_matches_allocating_entry_point(bbx):
    colliders, sdf = _colliders(bbx)
    for n, seed in ((50_000, 1), (3_000, 2), (bbx.NOALLOC_MAX_POINTS, 3)):
        print(f"no-alloc collider resolve: n={n} seed={seed}", flush=True)
        pos, vel = _points(n, seed)
        ep, ev, eh = bbx.colliders_resolve(colliders, 0.012, 0.6, pos, vel)
        gp, gv, gh = bbx.colliders_resolve_noalloc(colliders, 0.012, 0.6, pos, vel)
        assert eh.sum() > n // 10  # a good share of the points collide
        np.testing.assert_array_equal(gh, eh)
        np.testing.assert_array_equal(gp, ep)
        np.testing.assert_array_equal(gv, ev)
    del sdf


def test_noalloc_resolve_rejects_sizes_beyond_its_limit(bbx):
    colliders, sdf = _colliders(bbx)
    pos, vel = _points(bbx.NOALLOC_MAX_POINTS + 1, 4)
    with pytest.raises(bbx.BbxError, match="at most 65536 points"):
        bbx.colliders_resolve_noalloc(colliders, 0.012, 0.6, pos, vel)
    # an empty call is a no-op
    p, v, h = bbx.colliders_resolve_noalloc(colliders, 0.012, 0.6, np.zeros((0, 3)), np.zeros((0, 3)))
    assert p.shape == (0, 3) and h.size == 0
    del sdf
