"""Pin the oracle's continuous-emission rule (ContinuousParticleSetBuilder3::
MapGrid and MapGridEmit, src/core/grid.h:1286-1408) on a 3x1x1-cell strip
(cell 0.2 wide along x: cell 0 = [-0.3, -0.1), cell 1 = [-0.1, 0.1), cell
2 = [0.1, 0.3]) with hand-derived outcomes: which mapped positions a call
refills, in which order, the distance rejection, the skip of a cell whose
chain holds 100 particles, and the early stop at the builder's capacity."""
import dataclasses

import numpy as np

from bubbles_amd import scenes

A, B, C = [-0.25, 0.0, 0.0], [-0.15, 0.0, 0.0], [0.0, 0.0, 0.0]


def strip(points):
    base = scenes.dam_break(2, 2, 2)
    pos = np.asarray(points, dtype=np.float64)
    return dataclasses.replace(base, positions=pos, velocities=np.zeros_like(pos), target_spacing=0.1,
                               relative_radius=2.0, tank_size=(0.6, 0.2, 0.2), grid_spacing=0.2,
                               grid_scale=1.0)


def moved(sim, points):
    """Move the particles and redistribute from scratch (the flag: ascending
    ids), as a scene would between emissions."""
    sim.set("position", np.asarray(points, dtype=np.float64))
    sim.flag = 1
    sim.distribute()


def commit(sim, new):
    if len(new):
        sim.append(new)


def test_map_grid_records_chains_by_cell(oracle):
    s = oracle.Sim(strip([A, B, C]))
    assert list(s.info()["grid"].usizes) == [3, 1, 1]
    cells, offsets, pos = s.map_grid()
    # cells 0 and 1 hold particles; cell 0's chain is ids 0, 1 (ascending)
    assert cells.tolist() == [0, 1] and offsets.tolist() == [0, 2, 3]
    np.testing.assert_array_equal(pos, [A, B, C])


def test_refill_keeps_only_positions_clear_of_the_chain(oracle):
    s = oracle.Sim(strip([A, B, C]))
    mapped = s.map_grid()
    # B leaves for cell 2: cell 0 keeps A only.  A's mapped position is 0
    # from A (rejected), B's is 0.1 from A (kept); C's cell still holds C.
    moved(s, [A, [0.2, 0.0, 0.0], C])
    new, hit = s.map_grid_emit(mapped, 0.02, 1000)
    assert not hit
    np.testing.assert_array_equal(new, [B])
    commit(s, new)
    start, members = s.chains()
    assert start.tolist() == [0, 2, 3, 4] and members.tolist() == [0, 3, 2, 1]
    # nothing is left to refill: both mapped positions of cell 0 are taken
    new, hit = s.map_grid_emit(mapped, 0.02, 1000)
    assert len(new) == 0 and not hit
    # a rejection distance above the spacing of A and B rejects B as well
    moved(s, [A, [0.2, 0.0, 0.0], C, [0.25, 0.0, 0.0]])
    new, _ = s.map_grid_emit(mapped, 0.11, 1000)
    assert len(new) == 0


def test_refill_order_and_capacity_stop(oracle):
    s = oracle.Sim(strip([A, B, C]))
    mapped = s.map_grid()
    # A and B both leave: cell 0 is empty, so both of its mapped positions
    # come back, in chain order, before cell 1 (which still holds C)
    moved(s, [[0.2, 0.0, 0.0], [0.25, 0.0, 0.0], C])
    new, hit = s.map_grid_emit(mapped, 0.02, 1000)
    np.testing.assert_array_equal(new, [A, B])
    assert not hit
    # with room for one more particle only, the first is kept and the call
    # stops at the second (AddParticle fails: the kept one is committed)
    new, hit = s.map_grid_emit(mapped, 0.02, 4)
    np.testing.assert_array_equal(new, [A])
    assert hit
    new, hit = s.map_grid_emit(mapped, 0.02, 3)
    assert len(new) == 0 and hit


def test_full_chain_is_skipped(oracle):
    # 100 particles in a corner of cell 1, mapped there, then shifted inside
    # cell 1 far from every mapped position: a chain of 100 is passed over,
    # a chain of 99 takes exactly one (its first mapped position)
    rng = np.random.default_rng(1)
    cluster = np.c_[rng.uniform(-0.099, -0.09, 100), rng.uniform(-0.05, 0.05, 100), rng.uniform(-0.05, 0.05, 100)]
    s = oracle.Sim(strip(cluster))
    mapped = s.map_grid()
    assert mapped[0].tolist() == [1] and mapped[1].tolist() == [0, 100]
    shifted = cluster + [0.15, 0.0, 0.0]
    moved(s, shifted)
    new, hit = s.map_grid_emit(mapped, 0.02, 10000)
    assert len(new) == 0 and not hit
    shifted[57] = [0.25, 0.0, 0.0]  # one particle leaves for cell 2
    moved(s, shifted)
    new, hit = s.map_grid_emit(mapped, 0.02, 10000)
    np.testing.assert_array_equal(new, cluster[:1])
