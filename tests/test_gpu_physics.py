"""GPU parity: SPH / PCISPH physics (FP32 + Kahan) against the FP64 oracle.

Tolerances are stated per quantity; the oracle runs the reference method in
IEEE double on the same (FP32-rounded) inputs.
"""
import dataclasses
import time

import numpy as np
import pytest

from bubbles_amd import scenes

pytestmark = pytest.mark.gpu

# stated tolerances (FP32 device vs FP64 oracle)
RTOL_DENSITY = 2e-6        # relative, per particle (Kahan-summed poly6)
RTOL_FORCE = 2e-5          # relative to max |f|
RTOL_PRESSURE = 2e-3       # relative to max |p|; p = delta * (rho* - rho0) cancels
ATOL_POS_STEP = 2e-7       # absolute, metres, one substep (tank ~ 1 m, FP32 ulp 6e-8)


def small_dam(seed=2, **kw):
    return scenes.dam_break(32, 16, 32, spacing=0.01, tank=(0.5, 0.25, 0.5), seed=seed, **kw)


def paired(bbx, oracle, sc, **over):
    g = bbx.Solver(sc, **over)
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")), **over)
    return g, o


def test_density_matches_oracle(bbx, oracle):
    sc = small_dam()
    g, o = paired(bbx, oracle, sc)
    g.distribute()
    o.distribute()
    o.stage("interaction")
    o.stage("density", 1)
    gd, od = g.get("density"), o.get("density")
    np.testing.assert_allclose(gd, od, rtol=RTOL_DENSITY, atol=0)


def test_isolated_particle_density_closed_form(bbx, oracle):
    """One particle: rho = m * 315 / (64 pi h^3) (poly6 at r = 0)."""
    base = small_dam()
    pos = np.array([[0.01, 0.02, -0.03]])
    sc = dataclasses.replace(base, positions=pos, velocities=np.zeros_like(pos))
    g = bbx.Solver(sc)
    g.distribute()
    inf = g.info()
    h, m = inf["kernel_radius"], inf["mass"]
    expect = m * 315.0 / (64.0 * np.pi * h ** 3)
    assert abs(g.get("density")[0] - expect) <= 1e-6 * expect


def test_nonpressure_force_matches_oracle(bbx, oracle):
    sc = small_dam()
    g, o = paired(bbx, oracle, sc)
    # give the block some motion first so viscosity is non-zero
    for _ in range(5):
        g.pcisph_substep(sc.dt, 4, 0.01, bbx.MODE_TEXTBOOK)
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")))
    o.set("velocity", g.get("velocity"))
    o.set_chains(*g.chains())
    o.flag = g.flag
    g.distribute()
    g.stage_nonpressure()
    o.distribute()
    o.stage("interaction")
    o.stage("density", 1)
    o.stage("nonpressure")
    gf, of = g.get("force"), o.get("force")
    scale = np.abs(of).max()
    assert np.abs(gf - of).max() <= RTOL_FORCE * scale


@pytest.mark.parametrize("mode", ["exact", "textbook"])
def test_one_pcisph_substep_matches_oracle(bbx, oracle, mode):
    sc = small_dam()
    g, o = paired(bbx, oracle, sc)
    for _ in range(20):
        g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")))
    o.set("velocity", g.get("velocity"))
    o.set_chains(*g.chains())
    o.flag = g.flag
    m = bbx.MODE_REFERENCE_EXACT if mode == "exact" else bbx.MODE_TEXTBOOK
    gs = g.pcisph_substep(sc.dt, 8, 0.01, m)
    os_ = o.pcisph_substep(sc.dt, 8, 0.01, m)
    if mode == "exact":
        assert gs.iterations == 1 and os_.iterations == 1
    else:
        assert abs(gs.iterations - os_.iterations) <= 1
    assert abs(gs.max_density_error - os_.max_density_error) <= 1e-4
    if gs.iterations == os_.iterations:
        gp, op = g.get("pressure"), o.get("pressure")
        assert np.abs(gp - op).max() <= RTOL_PRESSURE * max(np.abs(op).max(), 1.0)
        np.testing.assert_allclose(g.get("position"), o.get("position"), rtol=0, atol=ATOL_POS_STEP)
        dv = np.abs(g.get("velocity") - o.get("velocity")).max()
        assert dv <= ATOL_POS_STEP / sc.dt


def test_pcisph_multistep_drift_reference_exact(bbx, oracle):
    """60 REFERENCE_EXACT substeps (fixed iteration count) from the same
    start: FP32 + Kahan drifts from FP64 by less than 1e-3 h."""
    sc = small_dam()
    g, o = paired(bbx, oracle, sc)
    for _ in range(60):
        g.pcisph_substep(sc.dt, 5, 0.01, bbx.MODE_REFERENCE_EXACT)
        o.pcisph_substep(sc.dt, 5, 0.01, oracle.MODE_REFERENCE_EXACT)
    dx = np.abs(g.get("position") - o.get("position")).max()
    assert dx <= 1e-3 * sc.kernel_radius, dx
    rel = np.abs(g.get("density") - o.get("density")) / o.get("density")
    assert np.median(rel) <= 1e-5 and rel.max() <= 1e-3, (np.median(rel), rel.max())


@pytest.mark.parametrize("warm", ["none", "auto"])
def test_pcisph_run_equals_repeated_substeps(bbx, warm):
    """bbx_pcisph_run(count) is `count` calls of bbx_pcisph_substep: the same
    state bit for bit and the same per-substep statistics, with and without
    the network's AUTO start (40 TEXTBOOK substeps into the collapse)."""
    from bubbles_amd import learned

    sc = small_dam()
    w = bbx.WARM_AUTO if warm == "auto" else bbx.WARM_NONE
    a, b = bbx.Solver(sc), bbx.Solver(sc)
    if warm == "auto":
        a.attach_mlp(learned.load())
        b.attach_mlp(learned.load())
    one = [a.pcisph_substep(sc.dt, 16, 0.01, bbx.MODE_TEXTBOOK, w) for _ in range(40)]
    b.synchronize()
    t0 = time.perf_counter()
    run = b.pcisph_run(sc.dt, 40, 16, 0.01, bbx.MODE_TEXTBOOK, w)
    wall_ms = (time.perf_counter() - t0) * 1e3
    assert len(run) == 40
    # each substep's device time, between the marks at its boundaries: the
    # substeps follow one another, so the times add up to at most the wall
    # time of the run
    assert all(x.gpu_ms > 0 for x in run)
    assert sum(x.gpu_ms for x in run) <= wall_ms * 1.01 + 0.1
    assert [s.iterations for s in one] == [s.iterations for s in run]
    assert [s.warm_started for s in one] == [s.warm_started for s in run]
    assert [s.max_density_error for s in one] == [s.max_density_error for s in run]
    for f in ("position", "velocity", "pressure", "density"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert b.pcisph_run(sc.dt, 0, 16, 0.01, bbx.MODE_TEXTBOOK, w) == []
    with pytest.raises(bbx.BbxError):
        b.pcisph_run(sc.dt, -1, 16, 0.01, bbx.MODE_TEXTBOOK, w)


def test_loop_decisions_without_the_published_signal(bbx, monkeypatch):
    """The host learns each loop decision, and with the integrating one the
    substep's statistics, from the pinned block the pressure-force kernel
    stores them in; with that block switched off (BBX_LOOP_SIGNAL=0) it reads
    the done flags and the statistics back through copies instead.  Both
    give the same substeps bit for bit and the same statistics, plain and
    warm-started, and the cap's range is 1..256."""
    from bubbles_amd import learned

    sc = small_dam(seed=5)

    def run():
        out = []
        for w in (bbx.WARM_NONE, bbx.WARM_AUTO):
            s = bbx.Solver(sc)
            s.attach_mlp(learned.load())
            st = s.pcisph_run(sc.dt, 30, 256, 0.01, bbx.MODE_TEXTBOOK, w)
            st.append(s.pcisph_substep(sc.dt, 256, 0.01, bbx.MODE_REFERENCE_EXACT, w))
            keys = ("iterations", "max_density_error", "full_redistribution", "bucket_overflows",
                    "lost_particles", "warm_started")
            out.append(([x.iterations for x in st], [tuple(getattr(x, k) for k in keys) for x in st],
                        s.get("position"), s.get("velocity")))
            s.close()
        return out

    ref = run()
    monkeypatch.setenv("BBX_LOOP_SIGNAL", "0")
    got = run()
    for a, b in zip(ref, got):
        assert a[0] == b[0] and a[1] == b[1]
        np.testing.assert_array_equal(a[2], b[2])
        np.testing.assert_array_equal(a[3], b[3])
    assert max(ref[0][0]) > 1  # the loop did go on at least once
    s = bbx.Solver(sc)
    with pytest.raises(bbx.BbxError):
        s.pcisph_substep(sc.dt, 257, 0.01, bbx.MODE_TEXTBOOK)
    with pytest.raises(bbx.BbxError):
        s.pcisph_substep(sc.dt, 0, 0.01, bbx.MODE_TEXTBOOK)
    s.close()


def test_pcisph_textbook_decisions_from_synchronised_states(bbx, oracle):
    """Every 4th of 48 GPU TEXTBOOK substeps, the oracle is reset to the GPU
    state and both run the next substep: the iteration count (the FP64
    max-error-vs-1% decision) and the reported max error agree."""
    sc = small_dam()
    g = bbx.Solver(sc)
    same, total = 0, 0
    for step in range(48):
        if step % 4 == 0:
            o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")))
            o.set("velocity", g.get("velocity"))
            o.set_chains(*g.chains())
            o.flag = g.flag
            a = g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
            b = o.pcisph_substep(sc.dt, 8, 0.01, oracle.MODE_TEXTBOOK)
            total += 1
            same += a.iterations == b.iterations
            if a.iterations == b.iterations:
                assert abs(a.max_density_error - b.max_density_error) <= 2e-5
        else:
            g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
    assert same >= total - 1, (same, total)


def test_pcisph_textbook_statistics_free_running(bbx, oracle):
    """60 free-running TEXTBOOK substeps: trajectories decorrelate (FP32 vs
    FP64 plus borderline decisions), the iteration statistics do not."""
    sc = small_dam()
    g, o = paired(bbx, oracle, sc)
    iters_g, iters_o = [], []
    for _ in range(60):
        iters_g.append(g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK).iterations)
        iters_o.append(o.pcisph_substep(sc.dt, 8, 0.01, oracle.MODE_TEXTBOOK).iterations)
    assert abs(np.mean(iters_g) - np.mean(iters_o)) <= 0.5
    dx = np.abs(g.get("position") - o.get("position")).max()
    assert dx <= 1e-2 * sc.kernel_radius, dx


def test_free_fall_closed_form(bbx):
    """Single particle, drag 0: symplectic Euler gives
    y_n = y0 + g dt^2 n (n + 1) / 2 (src/equations/sph_equations3.cpp:283-339)."""
    base = small_dam()
    pos = np.array([[0.0, 0.05, 0.0]])
    sc = dataclasses.replace(base, positions=pos, velocities=np.zeros_like(pos))
    g = bbx.Solver(sc, drag_coefficient=0.0)
    dt, n = 5e-4, 40
    for _ in range(n):
        g.pcisph_substep(dt, 3, 0.01, bbx.MODE_TEXTBOOK)
    gy = float(np.float32(-9.8))
    y = g.get("position")[0, 1]
    expect = 0.05 + gy * dt * dt * n * (n + 1) / 2
    assert abs(y - expect) < 2e-6
    vy = g.get("velocity")[0, 1]
    assert abs(vy - gy * dt * n) < 1e-5


def test_wall_bounce_restitution(bbx, oracle):
    """A particle crossing the floor is put at floor + radius with its normal
    velocity reflected by restitution 0.6 (src/core/collider.cpp:3-44)."""
    base = small_dam()
    r = base.target_spacing
    floor = -0.125
    pos = np.array([[0.0, floor + r + 1e-4, 0.0]])
    vel = np.array([[0.3, -2.0, 0.0]])
    sc = dataclasses.replace(base, positions=pos, velocities=vel)
    g = bbx.Solver(sc, drag_coefficient=0.0, gravity=(0.0, 0.0, 0.0))
    g.pcisph_substep(5e-4, 3, 0.01, bbx.MODE_TEXTBOOK)
    p, v = g.get("position")[0], g.get("velocity")[0]
    assert abs(p[1] - (floor + r)) < 1e-6
    assert abs(v[1] - 0.6 * 2.0) < 1e-5
    assert abs(v[0] - 0.3) < 1e-6
    o = oracle.Sim(dataclasses.replace(sc, positions=pos.astype(np.float32).astype(np.float64)),
                   drag_coefficient=0.0, gravity=(0.0, 0.0, 0.0))
    o.pcisph_substep(5e-4, 3, 0.01, oracle.MODE_TEXTBOOK)
    np.testing.assert_allclose(p, o.get("position")[0], atol=1e-6)
    np.testing.assert_allclose(v, o.get("velocity")[0], atol=1e-5)


def test_sph_substep_matches_oracle(bbx, oracle):
    sc = scenes.sph_dam_break()
    sc = dataclasses.replace(sc, positions=sc.positions[:4096].copy(),
                             velocities=sc.velocities[:4096].copy())
    g = bbx.Solver(sc, kind=bbx.SPH)
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")))
    dt = 1.6e-4
    for _ in range(3):
        g.sph_substep(dt)
        o.sph_substep(dt, oracle.SPH_TWO_PHASE)
    np.testing.assert_allclose(g.get("position"), o.get("position"), rtol=0, atol=5e-6)


def test_timestep_rule_matches_oracle(bbx, oracle):
    sc = small_dam()
    g, o = paired(bbx, oracle, sc)
    # before any force: speed limit only
    assert g.number_of_time_steps(1e-3, 5.0) == o.number_of_time_steps(1e-3, 5.0)
    assert g.number_of_time_steps(1e-3, 1.0) == o.number_of_time_steps(1e-3, 1.0)
    # PciSphSolver3::Advance reproduces the substep count of the oracle
    ng = g.pcisph_advance(1.0 / 240.0, 5, 0.01, bbx.MODE_REFERENCE_EXACT)
    no = o.pcisph_advance(1.0 / 240.0, 5, 0.01, oracle.MODE_REFERENCE_EXACT)
    assert ng == no


def _staging_run(bbx, sc_p, sc_s):
    g = bbx.Solver(sc_p)
    for _ in range(4):
        g.pcisph_substep(sc_p.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
    out = [g.get("position"), g.get("velocity"), g.get("pressure")]
    g.close()
    h = bbx.Solver(sc_s, kind=bbx.SPH, pseudo_viscosity=1000.0)  # XSPH active
    for _ in range(3):
        h.sph_substep(1.6e-4)
    out += [h.get("position"), h.get("velocity")]
    h.close()
    return out


@pytest.mark.parametrize("env", [{"BBX_TILE_CAP": "0"}, {"BBX_TILE_CAP2": "0"},
                                 {"BBX_TILE_CAP": "1500", "BBX_TILE_CAP2": "900"},
                                 {"BBX_BUCKET_ROW_MAX": "0"}, {"BBX_BUCKET_ROW_MAX": "20"}])
def test_staging_paths_bit_identical(bbx, monkeypatch, env):
    """Neighbours read from LDS tiles, through the LDS-index -> storage map
    (tiles over a kernel's LDS budget) or from global memory with storage
    indices (unstaged tiles) are the same values in the same order, and the
    bucket build's row-mask and cell-by-cell paths emit the same buckets, so
    every path gives bit-identical PCISPH and WCSPH (+ XSPH) results."""
    sc_p = small_dam(seed=7)
    s = scenes.sph_dam_break()
    sc_s = dataclasses.replace(s, positions=s.positions[:8192].copy(), velocities=s.velocities[:8192].copy())
    ref = _staging_run(bbx, sc_p, sc_s)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    got = _staging_run(bbx, sc_p, sc_s)
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a, b)


def test_published_counters_match_the_copied_ones(bbx, monkeypatch):
    """The counters a PCISPH substep reports with its integrating decision
    (LoopSignal) equal those the copy path reads back (BBX_LOOP_SIGNAL=0):
    150 particles within the radius of each other drop 150 * 50 bucket
    entries in the substep's bucket build (src/core/particle.h:45-51)."""
    rng = np.random.default_rng(5)
    pts = rng.uniform(-0.004, 0.004, size=(150, 3)).astype(np.float32).astype(np.float64)
    sc = dataclasses.replace(scenes.dam_break(2, 2, 2), positions=pts, velocities=np.zeros_like(pts))
    res = []
    for sig in ("1", "0"):
        monkeypatch.setenv("BBX_LOOP_SIGNAL", sig)
        s = bbx.Solver(sc)
        st = s.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK)
        res.append((st.bucket_overflows, st.lost_particles, st.full_redistribution, st.iterations,
                    st.max_density_error))
        s.close()
    assert res[0] == res[1]
    assert res[0][0] == 150 * 50
