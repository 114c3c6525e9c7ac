"""Pin the oracle's growing particle set (ContinuousParticleSetBuilder3::
Commit with a mapped grid, src/core/grid.h:1410-1425, which hands the new
ids to DistributeByParticleList, src/core/grid.h:358-387): the old chains are
untouched and each new id is added at the tail of the chain of the cell its
position hashes into, in the order the particles were added; a position
outside the grid joins no chain.  The expected chains are built here from the
old chains and the grid hash alone."""
import numpy as np

from bubbles_amd import scenes


def _chains_by_cell(start, members):
    return {c: list(members[start[c]:start[c + 1]]) for c in range(len(start) - 1)
            if start[c + 1] > start[c]}


def test_append_goes_to_chain_tails_in_order(oracle):
    sc = scenes.dam_break(8, 6, 8, spacing=0.01, tank=(0.3, 0.2, 0.3), seed=11)
    o = oracle.Sim(sc)
    o.pcisph_substep(sc.dt, 5, 0.01, oracle.MODE_TEXTBOOK)
    start, members = o.chains()
    old = _chains_by_cell(start, members)
    g = o.info()["grid"]
    rng = np.random.default_rng(3)
    pts = rng.uniform(0.0, 0.12, size=(60, 3)) + [0.02, 0.02, 0.02]
    pts[7] = [5.0, 5.0, 5.0]      # outside the grid
    pts[20] = pts[19]             # two new particles in one cell, in order
    n0 = o.n
    o.append(pts, np.zeros_like(pts))
    assert o.n == n0 + pts.shape[0]
    exp = {c: list(m) for c, m in old.items()}
    for k, p in enumerate(pts):
        cid, _ = oracle.grid_hash(g, p)
        if cid >= 0:
            exp.setdefault(cid, []).append(n0 + k)
    got = _chains_by_cell(*o.chains())
    assert got == exp
    assert all(n0 + 7 not in m for m in got.values())
    np.testing.assert_array_equal(o.get("position")[n0:], pts)
    np.testing.assert_array_equal(o.get("velocity")[n0:], 0.0)


def test_appended_particles_get_brute_force_buckets(oracle):
    sc = scenes.dam_break(6, 6, 6, spacing=0.01, tank=(0.3, 0.2, 0.3), seed=12)
    o = oracle.Sim(sc)
    o.pcisph_substep(sc.dt, 5, 0.01, oracle.MODE_TEXTBOOK)
    rng = np.random.default_rng(4)
    pts = rng.uniform(0.02, 0.08, size=(40, 3))
    o.append(pts)
    o.distribute()
    pos = o.get("position")
    h = o.info()["kernel_radius"]
    counts, ids = o.buckets()
    for i in range(o.n - pts.shape[0], o.n):
        d = pos - pos[i]
        dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        of = dist * dist - h * h  # the FP64 IsWithinStd rule
        want = set(np.nonzero(~((np.abs(of) < 1e-8) | (of > 0)))[0].tolist())
        got = set(ids[i, :counts[i]].tolist())
        assert got == want, i
