"""Growing particle sets (ContinuousParticleSetBuilder3::Commit with a mapped
grid, src/core/grid.h:1410-1425 + DistributeByParticleList :358-387): new
particles take ids n.. and join the tail of their cell's chain in the order
given.  The GPU path is compared with the oracle's restatement: chains
bit-exact right after the append and, with buckets, over the substeps that
follow; physics within the usual FP32 tolerances."""
import dataclasses

import numpy as np
import pytest

from bubbles_amd import scenes

pytestmark = pytest.mark.gpu


def _same_chains(g, o):
    gs, gm = g.chains()
    os_, om = o.chains()
    np.testing.assert_array_equal(gs, os_)
    np.testing.assert_array_equal(gm, om)


def _same_buckets(g, o):
    gc, gi = g.buckets()
    oc, oi = o.buckets()
    np.testing.assert_array_equal(gc, oc)
    mask = np.arange(gi.shape[1])[None, :] < gc[:, None]
    np.testing.assert_array_equal(np.where(mask, gi, -1), np.where(mask, oi, -1))


def _new_block(sc, nx, ny, nz, seed):
    """A lattice block just above the fluid column, starting in the cells of
    its top layer."""
    rng = np.random.default_rng(seed)
    s = sc.lattice_spacing
    lo = sc.positions.min(axis=0)
    corner = np.array([lo[0] + 3 * s, sc.positions[:, 1].max() + 1.2 * s, lo[2] + 3 * s])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    p = g * s + corner + rng.uniform(-1e-4, 1e-4, size=g.shape)
    rng.shuffle(p)  # insertion order is not cell order
    v = np.tile([0.0, -1.0, 0.0], (p.shape[0], 1))
    # FP32-representable, so both sides hash the same coordinates
    return p.astype(np.float32).astype(np.float64), v


def _sync(g, o):
    o.set("position", g.get("position"))
    o.set("velocity", g.get("velocity"))


def test_append_matches_oracle(bbx, oracle):
    sc = scenes.dam_break(16, 8, 16, spacing=0.01, tank=(0.5, 0.25, 0.5), seed=7)
    g = bbx.Solver(sc)
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")))
    for _ in range(5):
        g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
        o.pcisph_substep(sc.dt, 8, 0.01, oracle.MODE_TEXTBOOK)
        _same_chains(g, o)
        _sync(g, o)
    n0 = sc.n
    p, v = _new_block(sc, 6, 3, 6, seed=3)
    cap0 = g.capacity()
    g.append(p, v)
    o.append(p, v)
    assert g.n == o.n == n0 + p.shape[0]
    assert g.capacity() >= g.n and (g.capacity() > cap0 or n0 + p.shape[0] <= cap0)
    _same_chains(g, o)
    # the new particles' state arrived
    np.testing.assert_array_equal(g.get("position")[n0:], p)
    np.testing.assert_array_equal(g.get("velocity")[n0:], v)
    for k in range(5):
        a = g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
        b = o.pcisph_substep(sc.dt, 8, 0.01, oracle.MODE_TEXTBOOK)
        _same_chains(g, o)
        _same_buckets(g, o)
        assert abs(a.iterations - b.iterations) <= 1
        if a.iterations != b.iterations:
            break
        np.testing.assert_allclose(g.get("position"), o.get("position"), atol=2e-5)
        _sync(g, o)


def test_reserve_keeps_state(bbx):
    sc = scenes.dam_break(8, 8, 8, spacing=0.01, tank=(0.3, 0.2, 0.3), seed=4)
    g = bbx.Solver(sc)
    for _ in range(3):
        g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
    before = {f: g.get(f) for f in ("position", "velocity", "density", "pressure", "pressure_force")}
    chains = g.chains()
    g.reserve(10 * sc.n)
    assert g.capacity() >= 10 * sc.n
    for f, a in before.items():
        np.testing.assert_array_equal(g.get(f), a, err_msg=f)
    for a, b in zip(g.chains(), chains):
        np.testing.assert_array_equal(a, b)
    # and the solver keeps stepping exactly as a solver that never grew
    h = bbx.Solver(sc)
    for _ in range(3):
        h.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
    for _ in range(3):
        a = g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
        b = h.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
        assert a.iterations == b.iterations
    np.testing.assert_array_equal(g.get("position"), h.get("position"))


def test_append_outside_the_grid_is_in_no_chain(bbx):
    sc = scenes.dam_break(8, 8, 8, spacing=0.01, tank=(0.3, 0.2, 0.3), seed=5)
    g = bbx.Solver(sc)
    g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
    g.append(np.array([[5.0, 5.0, 5.0], [0.1, 0.1, 0.1]]))
    start, members = g.chains()
    assert sc.n + 1 in members and sc.n not in members


def test_growing_from_an_empty_set_equals_starting_full(bbx):
    """A solver created empty and filled by one append steps exactly like a
    solver created with the same particles: both chains are in ascending id
    and the state is the same."""
    sc = scenes.dam_break(12, 8, 12, spacing=0.01, tank=(0.3, 0.2, 0.3), seed=6)
    empty = dataclasses.replace(sc, positions=np.zeros((0, 3)), velocities=np.zeros((0, 3)))
    a = bbx.Solver(empty)
    st = a.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)  # nothing to do
    a.append(sc.positions, sc.velocities)
    b = bbx.Solver(sc)
    _same_chains(a, b)
    for _ in range(5):
        sa = a.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
        sb = b.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
        assert sa.iterations == sb.iterations
    np.testing.assert_array_equal(a.get("position"), b.get("position"))
    np.testing.assert_array_equal(a.get("velocity"), b.get("velocity"))
