"""GPU parity: neighbour grid (cell ids, chain order, buckets) bit-exact
against the FP64 oracle on the same FP32 positions.

Reference semantics under test: full distribution = ascending id per cell
(src/core/grid.h:389-407); incremental = concatenation of the 27 neighbour
cells' old chains in (y,x,z) order (src/core/grid.h:449-494); buckets = the
first 100 chain members within the FP64 cutoff (src/core/grid.h:422-447).
"""
import numpy as np
import pytest

from bubbles_amd import scenes

pytestmark = pytest.mark.gpu


def small_dam(order="sorted", seed=2):
    sc = scenes.dam_break(32, 16, 32, spacing=0.01, tank=(0.5, 0.25, 0.5), seed=seed)
    if order == "shuffled":
        sc = sc.shuffled(seed)
    return sc


def gpu_positions_scene(sc, pos):
    """Scene whose positions are the FP32 device positions (widened)."""
    import dataclasses

    return dataclasses.replace(sc, positions=np.asarray(pos, dtype=np.float64).copy())


def assert_chains_equal(a, b):
    sa, ma = a
    sb, mb = b
    assert np.array_equal(sa, sb), "cell_start differs"
    assert np.array_equal(ma, mb), "chain members / order differ"


def assert_buckets_equal(a, b):
    ca, ia = a
    cb, ib = b
    assert np.array_equal(ca, cb), f"bucket counts differ at {np.nonzero(ca != cb)[0][:10]}"
    n, k = ia.shape
    mask = np.arange(k)[None, :] < ca[:, None]
    assert np.array_equal(np.where(mask, ia, -1), np.where(mask, ib, -1)), "bucket ids/order differ"


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_setup_scalars_and_grid(bbx, oracle, order):
    sc = small_dam(order)
    g = bbx.Solver(sc)
    o = oracle.Sim(sc)
    gi, oi = g.info(), o.info()
    # FP64 host scalars are computed by the same method: exact
    assert gi["mass"] == oi["mass"]
    assert gi["kernel_radius"] == oi["kernel_radius"]
    assert gi["radius"] == oi["radius"]
    assert gi["delta_denom"] == oi["delta_denom"]
    gg, og = gi["grid"], oi["grid"]
    assert list(gg.usizes) == list(og.usizes) and gg.total == og.total
    for f in ("cells_len", "min_point", "bounds_min", "bounds_max"):
        assert list(getattr(gg, f)) == list(getattr(og, f))


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_setup_full_distribution_exact(bbx, oracle, order):
    sc = small_dam(order)
    g = bbx.Solver(sc)
    o = oracle.Sim(gpu_positions_scene(sc, g.get("position")))
    assert_chains_equal(g.chains(), o.chains())


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_incremental_distribution_exact_over_steps(bbx, oracle, order):
    """Run GPU PCISPH substeps; before each one, replay the GPU's own chain
    state and positions through the oracle's distribution and compare."""
    sc = small_dam(order)
    g = bbx.Solver(sc)
    o = oracle.Sim(gpu_positions_scene(sc, g.get("position")))
    for step in range(12):
        pre_chain = g.chains()
        pre_pos = g.get("position")
        flag = g.flag
        gst = g.distribute()
        o.set_chains(*pre_chain)
        o.set("position", pre_pos)
        o.flag = flag
        ost = o.distribute()
        assert gst.full_redistribution == ost.full_redistribution
        assert gst.lost_particles == ost.lost_particles
        assert gst.bucket_overflows == ost.bucket_overflows
        assert_chains_equal(g.chains(), o.chains())
        assert_buckets_equal(g.buckets(), o.buckets())
        g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)


def test_forced_full_redistribution_exact(bbx, oracle):
    """A velocity kick above 0.9 cell / dt sets requiresHigherLevelUpdate in
    integration; the next distribution is a full one (ascending ids)."""
    sc = small_dam("shuffled", seed=5)
    g = bbx.Solver(sc)
    g.pcisph_substep(sc.dt, 3, 0.01, bbx.MODE_TEXTBOOK)
    # a uniform kick (no relative velocity, so viscosity does not damp it)
    v = g.get("velocity")
    v[:, 0] += 0.95 * 0.02 / sc.dt * 1.2
    g.set("velocity", v)
    g.pcisph_substep(sc.dt, 3, 0.01, bbx.MODE_TEXTBOOK)
    assert g.flag == 1
    pre_chain = g.chains()
    pre_pos = g.get("position")
    gst = g.distribute()
    assert gst.full_redistribution == 1
    o = oracle.Sim(gpu_positions_scene(sc, pre_pos))
    o.set_chains(*pre_chain)
    o.flag = 1
    ost = o.distribute()
    assert ost.full_redistribution == 1
    assert_chains_equal(g.chains(), o.chains())
    assert_buckets_equal(g.buckets(), o.buckets())
    # full distribution is ascending id inside each cell
    start, members = g.chains()
    for c in np.nonzero(np.diff(start) > 1)[0][:200]:
        seg = members[start[c]:start[c + 1]]
        assert np.all(np.diff(seg) > 0)


def test_bucket_cutoff_edge_cases(bbx, oracle):
    """Pairs placed within a few ulps of the kernel radius exercise the FP64
    branch of the bucket rule (IsWithinStd, src/core/kernel.cpp:229-234)."""
    import dataclasses

    base = small_dam("sorted")
    h = base.kernel_radius
    rng = np.random.default_rng(11)
    pts = []
    for k in range(400):
        c = rng.uniform(-0.09, 0.09, size=3)  # pairs stay inside the grid
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = h * (1.0 + rng.choice([-1, 1]) * rng.choice([1e-9, 1e-8, 3e-7, 2e-6, 1e-5]))
        r = r - 1e-8 / (2 * h) * rng.choice([0.0, 1.0, 0.5])  # around h^2 - 1e-8
        pts.append(c)
        pts.append(c + r * d)
    pos = np.asarray(pts, dtype=np.float32).astype(np.float64)
    sc = dataclasses.replace(base, positions=pos, velocities=np.zeros_like(pos))
    g = bbx.Solver(sc)
    o = oracle.Sim(gpu_positions_scene(sc, g.get("position")))
    g.distribute()
    o.distribute()
    assert_buckets_equal(g.buckets(), o.buckets())


def test_host_hash_matches_oracle(bbx, oracle):
    sc = small_dam("sorted")
    g = bbx.Solver(sc)
    gd = g.info()["grid"]
    rng = np.random.default_rng(3)
    lo, hi = np.array(gd.bounds_min), np.array(gd.bounds_max)
    pts = rng.uniform(lo, hi, size=(5000, 3))
    # points exactly on and within 1e-8 of the bounds take the Epsilon branch
    pts[:50, 0] = lo[0]
    pts[50:100, 1] = hi[1]
    pts[100:150, 2] = hi[2] - 5e-9
    ids = g.hash(pts)
    og = oracle.Sim(sc).info()["grid"]
    ref = np.array([oracle.grid_hash(og, p)[0] for p in pts])
    assert np.array_equal(ids, ref)


def test_cell_neighbors_match_oracle(bbx, oracle):
    sc = small_dam("sorted")
    g = bbx.Solver(sc)
    og = oracle.Sim(sc).info()["grid"]
    total = og.total
    for c in [0, 1, total - 1, total // 2, 25, 25 * 13, 137]:
        assert g.cell_neighbors(c) == oracle.grid_neighbors(og, c)


@pytest.mark.parametrize("n", [0, 1, 2])
def test_tiny_particle_counts(bbx, oracle, n):
    import dataclasses

    base = small_dam("sorted")
    pos = base.positions[:n].copy()
    sc = dataclasses.replace(base, positions=pos, velocities=np.zeros_like(pos))
    g = bbx.Solver(sc)
    st = g.pcisph_substep(sc.dt, 3, 0.01, bbx.MODE_TEXTBOOK)
    assert st.iterations >= 1
    if n:
        o = oracle.Sim(gpu_positions_scene(sc, pos.astype(np.float32).astype(np.float64)))
        o.pcisph_substep(sc.dt, 3, 0.01, oracle.MODE_TEXTBOOK)
        np.testing.assert_allclose(g.get("position"), o.get("position"), rtol=0, atol=1e-6)


def test_bucket_cap_truncates_at_100_like_the_oracle(bbx, oracle):
    """150 particles within the radius of each other: every bucket keeps the
    first 100 in cell/chain order and the overflow count is 150 * 50
    (src/core/particle.h:45-51)."""
    import dataclasses

    rng = np.random.default_rng(5)
    pts = rng.uniform(-0.004, 0.004, size=(150, 3)).astype(np.float32).astype(np.float64)
    base = scenes.dam_break(2, 2, 2)
    sc = dataclasses.replace(base, positions=pts, velocities=np.zeros_like(pts))
    g = bbx.Solver(sc)
    o = oracle.Sim(sc)
    gst, ost = g.distribute(), o.distribute()
    assert gst.bucket_overflows == ost.bucket_overflows == 150 * 50
    gc, gi = g.buckets()
    oc, oi = o.buckets()
    assert np.all(gc == 100) and np.array_equal(gc, oc)
    np.testing.assert_array_equal(gi[:, :100], oi[:, :100])


def test_incremental_chain_order_hand_derived_on_device(bbx):
    """The oracle's hand-derived three-particle case on the device: a cell's
    new chain visits its neighbours in (y, x, z) order, a two-cell jump
    without the flag loses the particle, and the flag restores ascending
    ids (src/core/grid.h:449-494)."""
    import dataclasses

    base = scenes.dam_break(2, 2, 2)
    # 0.09 rather than the oracle test's 0.1: x = 0.1 lies on a cell face,
    # where the FP32 position (0.100000001) hashes one cell further
    sc = dataclasses.replace(base, positions=np.array([[-0.2, 0.0, 0.0], [0.09, 0.0, 0.0], [0.05, 0.0, 0.0]]),
                             velocities=np.zeros((3, 3)), target_spacing=0.1, relative_radius=2.0,
                             tank_size=(0.6, 0.2, 0.2), grid_spacing=0.2, grid_scale=1.0)
    g = bbx.Solver(sc)
    assert list(g.info()["grid"].usizes) == [3, 1, 1]
    start, members = g.chains()
    assert start.tolist() == [0, 1, 3, 3] and members.tolist() == [0, 1, 2]
    g.set("position", np.array([[-0.05, 0.0, 0.0], [-0.15, 0.0, 0.0], [0.05, 0.0, 0.0]]))
    g.distribute()
    start, members = g.chains()
    assert start.tolist() == [0, 1, 3, 3] and members.tolist() == [1, 0, 2]
    g.set("position", np.array([[0.0, 0.0, 0.0], [-0.01, 0.0, 0.0], [0.25, 0.0, 0.0]]))
    g.distribute()
    start, members = g.chains()
    assert start.tolist() == [0, 0, 2, 3] and members.tolist() == [1, 0, 2]
    g.set("position", np.array([[0.0, 0.0, 0.0], [-0.01, 0.0, 0.0], [-0.25, 0.0, 0.0]]))
    st = g.distribute()
    assert st.lost_particles == 1
    start, members = g.chains()
    assert members.tolist() == [1, 0]
    g.flag = 1
    st = g.distribute()
    assert st.full_redistribution == 1 and st.lost_particles == 0
    start, members = g.chains()
    assert start.tolist() == [0, 1, 3, 3] and members.tolist() == [2, 0, 1]


def test_crowded_cell_order_matches_oracle(bbx, oracle):
    """6,000 particles inside one cell, in shuffled id order, next to a
    lattice block: the counting sort ranks a crowded cell exactly (workgroups
    wholly inside it rank through LDS chunks, the two at its edges through
    global memory), so the full distribution (ascending id) and an
    incremental one (F7 concatenation of the 27 neighbours' old chains,
    src/core/grid.h:389-407,449-494) match the oracle at that occupancy."""
    import dataclasses

    base = small_dam("sorted")
    h = base.kernel_radius
    gd = bbx.Solver(base).info()["grid"]
    u = np.array(gd.usizes) - 4  # the far corner, away from the lattice block
    centre = np.array(gd.bounds_min) + (u + 0.5) * np.array(gd.cells_len)
    rng = np.random.default_rng(17)
    crowd = centre + rng.uniform(-0.45 * h, 0.45 * h, size=(6000, 3))
    pos = np.concatenate([base.positions, crowd]).astype(np.float32).astype(np.float64)
    perm = rng.permutation(pos.shape[0])
    pos = pos[perm]
    sc = dataclasses.replace(base, positions=pos, velocities=np.zeros_like(pos))
    g = bbx.Solver(sc)
    o = oracle.Sim(gpu_positions_scene(sc, g.get("position")))
    start, members = g.chains()
    assert np.diff(start).max() >= 6000
    assert_chains_equal((start, members), o.chains())
    # every particle moves by less than a third of a cell: crowd members
    # spread into the neighbouring cells and the rank order is rebuilt
    pre_chain = (start, members)
    moved = pos + rng.uniform(-0.3 * h, 0.3 * h, size=pos.shape)
    moved = moved.astype(np.float32).astype(np.float64)
    g.set("position", moved)
    assert g.flag == 0
    gst = g.distribute()
    o.set_chains(*pre_chain)
    o.set("position", g.get("position"))
    o.flag = 0
    ost = o.distribute()
    assert gst.full_redistribution == ost.full_redistribution == 0
    assert gst.lost_particles == ost.lost_particles
    assert_chains_equal(g.chains(), o.chains())
    assert_buckets_equal(g.buckets(), o.buckets())


@pytest.mark.timeout(900)
@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_dam_break_1m_grid_gate(bbx, oracle, order):
    """The 1,048,576-particle dam break (128x64x128 at 0.01, 2x1x2 tank):
    cell chains and buckets bit-exact against the oracle at substeps 0, 1
    and 50, in sorted and in shuffled particle order.  At each checkpoint
    the oracle replays the device's pre-distribution chain state and
    positions through its own distribution and bucket build."""
    sc = scenes.dam_break(128, 64, 128, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=2)
    if order == "shuffled":
        sc = sc.shuffled(2)
    g = bbx.Solver(sc)
    o = oracle.Sim(gpu_positions_scene(sc, g.get("position")), threads=16)
    assert_chains_equal(g.chains(), o.chains())  # Setup: ascending ids
    step = 0
    for target in (0, 1, 50):
        while step < target:
            g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
            step += 1
        pre_chain, pre_pos, flag = g.chains(), g.get("position"), g.flag
        gst = g.distribute()
        o.set_chains(*pre_chain)
        o.set("position", pre_pos)
        o.flag = flag
        ost = o.distribute()
        assert (gst.full_redistribution, gst.lost_particles, gst.bucket_overflows) == \
            (ost.full_redistribution, ost.lost_particles, ost.bucket_overflows), target
        assert_chains_equal(g.chains(), o.chains())
        assert_buckets_equal(g.buckets(), o.buckets())
        print(f"1M {order} substep {target}: chains and buckets exact", flush=True)


def test_device_hash_on_grid_bounds(bbx, oracle):
    """Particles whose FP32 coordinate lies within 1e-8 of a grid bound take
    the device hash's ExtremeEpsilon branch (src/core/grid.h:258-298): on the
    upper bound they are pulled into the last cell instead of falling
    outside the grid.  One FP32 ulp further out (more than 1e-8 away) they
    hash outside and join no chain.  The solver's setup distribution (device
    hash) is compared with the oracle's FP64 hash and chains."""
    import dataclasses

    base = small_dam("sorted")
    gd = bbx.Solver(base).info()["grid"]
    lo, hi = np.array(gd.bounds_min), np.array(gd.bounds_max)
    rng = np.random.default_rng(23)
    pts, expect_lost = [], 0
    for a in range(3):
        for bound in (lo[a], hi[a]):
            f = np.float32(bound)
            assert abs(float(f) - bound) < 1e-8  # this tank's bounds have such FP32 neighbours
            for k, coord in enumerate((f, np.nextafter(f, np.float32(-np.inf)), np.nextafter(f, np.float32(np.inf)))):
                for _ in range(16):
                    p = rng.uniform(lo + 0.05, hi - 0.05)
                    p[a] = float(coord)
                    pts.append(p)
                    outside = float(coord) < lo[a] or float(coord) > hi[a]
                    if outside and abs(float(coord) - bound) >= 1e-8:
                        expect_lost += 1
    pos = np.asarray(pts, dtype=np.float32).astype(np.float64)
    sc = dataclasses.replace(base, positions=pos, velocities=np.zeros_like(pos))
    g = bbx.Solver(sc)
    o = oracle.Sim(gpu_positions_scene(sc, g.get("position")))
    assert_chains_equal(g.chains(), o.chains())
    start, members = g.chains()
    assert members.size == pos.shape[0] - expect_lost and expect_lost == 2 * 3 * 16
    # every particle exactly on an upper bound is in the last cell of that axis
    og = o.info()["grid"]
    on_hi = [i for i in range(pos.shape[0]) if any(pos[i, a] == float(np.float32(hi[a])) for a in range(3))]
    assert len(on_hi) == 3 * 16
    for i in on_hi:
        cid, u = oracle.grid_hash(og, pos[i])
        a = int(np.argmax([pos[i, b] == float(np.float32(hi[b])) for b in range(3)]))
        assert cid >= 0 and u[a] == og.usizes[a] - 1
        cell = int(np.searchsorted(start, np.nonzero(members == i)[0][0], side="right") - 1)
        assert cell == cid
