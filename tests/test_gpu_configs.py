"""The learned warm start on the benchmark scene against the FP64 oracle, and
the training harvest against the oracle's features and converged pressures.

* The 1,048,576-particle dam break (128x64x128 at 0.01, h 0.02, 2x1x2 tank,
  dt 5e-4) with the shipped network, 20 REFERENCE_EXACT substeps: every
  substep the oracle starts its loop from the device's own warm pressures
  (the network is not part of the oracle), so the comparison isolates the
  FP32 solver on the learned path
  (src/equations/pcisph_equations3.cpp:3-256).  Then one TEXTBOOK substep
  from a synchronised state iterates to the 1% tolerance in the same number
  of corrections, to within one.
* bbx_pcisph_harvest: the (features, converged pressure / P) pairs the
  network trains on, by particle id, equal the oracle's features of the same
  state and the pressures its own TEXTBOOK loop converges to.
"""
import dataclasses

import numpy as np
import pytest

from bubbles_amd import scenes

pytestmark = pytest.mark.gpu


def _mirror(oracle, sc, g, threads=16):
    """An oracle simulation holding the device's state: positions,
    velocities, pressures (the previous substep's, a feature), chains and the
    0.9-cell flag."""
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")), threads=threads)
    o.set("velocity", g.get("velocity"))
    o.set("pressure", g.get("pressure"))
    o.set_chains(*g.chains())
    o.flag = g.flag
    return o


def _errors(g, o, h):
    rg, ro = g.get("density"), o.get("density")
    drel = np.abs(rg - ro) / ro
    perr = np.abs(g.get("position") - o.get("position")).max(axis=1) / h
    return {"density_rel_max": float(drel.max()), "density_rel_mean": float(drel.mean()),
            "pos_over_h_max": float(perr.max()), "pos_over_h_mean": float(perr.mean())}


@pytest.mark.timeout(600)
def test_learned_dam_break_1m_against_fp64(bbx, oracle):
    from bubbles_amd import learned

    sc = scenes.dam_break(128, 64, 128, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=2)
    g = bbx.Solver(sc)
    g.attach_mlp(learned.load())
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position")), threads=16)
    table = {}
    for k in range(1, 21):
        g.pcisph_substep(sc.dt, 5, 0.01, bbx.MODE_REFERENCE_EXACT, bbx.WARM_MLP)
        pw = g.get("warm_pressure")
        assert pw.max() > 0  # the network starts the loop
        o.pcisph_substep(sc.dt, 5, 0.01, oracle.MODE_REFERENCE_EXACT, warm=pw.astype(np.float64))
        if k in (1, 10, 20):
            table[k] = _errors(g, o, sc.kernel_radius)
            print("learned 1M", k, table[k], flush=True)
    # bounds: about 3x the measured values at substep 20 (density 6.4e-4 max,
    # 1.8e-6 mean; position 1.1e-3 h max, 2.9e-6 h mean; the maximum is one
    # particle of a million at the splash front); one substep is arithmetic
    # alone (measured 1.6e-7 and 1.6e-6 h)
    assert table[1]["density_rel_max"] <= 1e-6 and table[1]["pos_over_h_max"] <= 1e-5, table[1]
    last = table[20]
    assert last["density_rel_max"] <= 2e-3 and last["density_rel_mean"] <= 1e-5, last
    assert last["pos_over_h_max"] <= 3.5e-3 and last["pos_over_h_mean"] <= 1e-5, last
    # TEXTBOOK from a synchronised state, both loops from the same warm start
    o = _mirror(oracle, sc, g)
    st = g.pcisph_substep(sc.dt, 64, 0.01, bbx.MODE_TEXTBOOK, bbx.WARM_MLP)
    ost = o.pcisph_substep(sc.dt, 64, 0.01, oracle.MODE_TEXTBOOK,
                           warm=g.get("warm_pressure").astype(np.float64))
    print("learned 1M TEXTBOOK iterations", st.iterations, ost.iterations, st.max_density_error,
          ost.max_density_error, flush=True)
    assert abs(st.iterations - ost.iterations) <= 1
    assert st.max_density_error < 0.01 and ost.max_density_error < 0.01


@pytest.mark.timeout(300)
def test_harvest_pairs_match_oracle(bbx, oracle):
    sc = scenes.dam_break(32, 16, 32, spacing=0.01, tank=(0.5, 0.25, 0.5), seed=3)
    g = bbx.Solver(sc)
    # step into the collapse: the first state whose substep needed three or
    # more corrections (the targets then come from a loop that iterates)
    for k in range(200):
        if g.pcisph_substep(sc.dt, 32, 0.01, bbx.MODE_TEXTBOOK).iterations >= 3:
            break
    assert k < 199
    o = _mirror(oracle, sc, g, threads=8)
    f = _mirror(oracle, sc, g, threads=8)
    feats, targ, st = g.harvest(sc.dt, 32, 0.01)
    # features at the start of the substep, in FP64
    f.distribute()
    f.stage("interaction")
    f.stage("density", 0)  # keep the previous substep's pressure (a feature)
    f.stage("nonpressure")
    of = f.features(sc.dt)
    scale = np.abs(of).max(axis=0) + 1e-12
    ferr = (np.abs(feats - of) / scale).max(axis=0)
    print("harvest feature errors / column scale", np.array2string(ferr, precision=1), flush=True)
    assert ferr.max() <= 1e-3
    # targets: the converged pressure of the same TEXTBOOK substep / P
    ost = o.pcisph_substep(sc.dt, 32, 0.01, oracle.MODE_TEXTBOOK)
    P = o.delta(sc.dt) * sc.target_density
    ot = o.get("pressure") / P
    terr = np.abs(targ - ot).max() / max(np.abs(ot).max(), 1e-12)
    print("harvest iterations", st.iterations, ost.iterations, "target error / max", terr, flush=True)
    assert st.iterations >= 2 and ot.max() > 0  # the loop did work and produced pressures
    assert abs(st.iterations - ost.iterations) <= 1
    assert terr <= (2e-3 if st.iterations == ost.iterations else 5e-2)
    # the pairs are by particle id: permuting the storage changes nothing
    assert feats.shape == (sc.n, 16) and targ.shape == (sc.n,)


def _auto_rule(history):
    """The AUTO decision for the next substep from (iterations, warm) of the
    previous ones (include/bubbles_amd.h, BBX_WARM_AUTO)."""
    if not history:
        return False
    if any(k >= 4 for k, _ in history):
        return True  # the flow has become active: the network stays on
    it, warm = history[-1]
    if not warm:
        return it >= 3
    ones = 0
    for k, w in reversed(history):
        if not (w and k == 1):
            break
        ones += 1
    return ones < 8


@pytest.mark.timeout(300)
def test_auto_warm_start_policy_and_gain_at_1m(bbx):
    """BBX_WARM_AUTO on the 1M dam break to tolerance (cap 64), 150 substeps
    from rest: each substep's start (plain or network) follows the stated
    rule; until the first warm start the run is bit-identical to plain; and
    the device time over the window is lower than plain's (measured 1.15x
    over these first 150 substeps, 1.43x over the whole 500), with about as
    many unconverged substeps and no larger density error."""
    from bubbles_amd import learned

    sc = scenes.dam_break(128, 64, 128, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=2)
    p = bbx.Solver(sc)
    a = bbx.Solver(sc)
    a.attach_mlp(learned.load())
    hist, pt, at, perr, aerr = [], 0.0, 0.0, [], []
    same = True
    for k in range(150):
        sp = p.pcisph_substep(sc.dt, 64, 0.01, bbx.MODE_TEXTBOOK)
        sa = a.pcisph_substep(sc.dt, 64, 0.01, bbx.MODE_TEXTBOOK, bbx.WARM_AUTO)
        assert bool(sa.warm_started) == _auto_rule(hist), (k, hist[-3:])
        hist.append((sa.iterations, bool(sa.warm_started)))
        if same and not sa.warm_started:
            assert sa.iterations == sp.iterations
            if k % 10 == 0:
                assert np.array_equal(a.get("position"), p.get("position")), k
        same = same and not sa.warm_started
        pt += sp.gpu_ms
        at += sa.gpu_ms
        perr.append(sp.max_density_error)
        aerr.append(sa.max_density_error)
    warm = sum(w for _, w in hist)
    print(f"AUTO 1M: warm {warm}/150, device ms plain {pt:.1f} auto {at:.1f} ratio {pt / at:.3f}; "
          f"unconverged plain {np.sum(np.array(perr) >= 0.01)} auto {np.sum(np.array(aerr) >= 0.01)}", flush=True)
    assert not hist[0][1] and 20 <= warm < 150
    assert pt / at >= 1.05
    # both loops stop at the same tolerance; a stalled substep (error above
    # 1% after 64 corrections) is a property of the method at this dt, and
    # the learned run may stall in a few different substeps than plain
    assert np.sum(np.array(aerr) >= 0.01) <= np.sum(np.array(perr) >= 0.01) + 3
    assert max(aerr) <= 1.25 * max(perr) + 2e-3


@pytest.mark.timeout(300)
def test_whole_run_to_tolerance_at_1m(bbx):
    """The 1M dam break's whole 500 substeps from rest, TEXTBOOK to 1% with
    the benchmark's cap of 128: the learned start (AUTO) meets the tolerance
    in (nearly) every substep -- 500 of 500 when measured -- and in no fewer
    substeps than plain, with about half of plain's corrections and at least
    1.3x its device throughput (measured 1.43x; plain misses 1% in 1% of the
    substeps, src/equations/pcisph_equations3.cpp:218-256)."""
    from bubbles_amd import learned

    sc = scenes.dam_break(128, 64, 128, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=2)
    res = {}
    for label, warm in (("plain", bbx.WARM_NONE), ("auto", bbx.WARM_AUTO)):
        s = bbx.Solver(sc)
        if warm != bbx.WARM_NONE:
            s.attach_mlp(learned.load())
        st = s.pcisph_run(sc.dt, 500, 128, 0.01, bbx.MODE_TEXTBOOK, warm)
        s.close()
        err = np.array([x.max_density_error for x in st])
        res[label] = {"it": float(np.mean([x.iterations for x in st])),
                      "ms": float(np.sum([x.gpu_ms for x in st])),
                      "unconverged": int(np.sum(err >= 0.01)), "max_err": float(err.max())}
    print("whole run at cap 128:", res, flush=True)
    p, a = res["plain"], res["auto"]
    assert a["unconverged"] <= 2 and a["unconverged"] <= p["unconverged"]
    assert a["max_err"] <= max(p["max_err"], 0.0125)
    assert a["it"] <= 0.7 * p["it"]
    assert p["ms"] / a["ms"] >= 1.3
