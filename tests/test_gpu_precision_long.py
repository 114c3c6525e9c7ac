"""Long-horizon precision study on the benchmark geometry (dam break at
spacing 0.01, h 0.02, 2x1x2 tank): FP32 + Kahan on the GPU against the FP64
oracle from the same FP32 start.

The 4M-20 and learned 262k-200 cases run by default; the 1000-substep
horizons and the drift study are opt-in (several minutes of oracle time on
16 host threads each): set BBX_LONG_PRECISION=1, and
BBX_PRECISION_OUT=<file> to write the tables as JSON.  Cases:

* plain REFERENCE_EXACT (one iteration per substep, so both trajectories
  differ by arithmetic alone): 262,144, 1,048,576 and 4,194,304 particles
  (256x64x256, which needs a 4x1x4 tank) over 1000 substeps, and the 4M
  case over 20;
* learned warm start, REFERENCE_EXACT, 262,144 over 200 substeps and
  1,048,576 over 1000: every
  substep the oracle starts from the GPU's own warm pressures (the network is
  not part of the oracle), so the comparison isolates the FP32 solver;
* TEXTBOOK iteration drift, 262,144 over 200 substeps: how often the FP32 and
  FP64 runs stop after a different number of corrections (each run on its own
  trajectory).

Measured on MI355X at 262k, REFERENCE_EXACT (density relative, pressure
force relative to max |f_p|, position in units of h; max / mean):

  substep   density           f_p               position
  1         1.5e-7 / 2.7e-8   1.4e-4 / 9.6e-7   1.5e-6 / 3.3e-7
  100       1.5e-3 / 8.6e-6   3.4e-3 / 3.0e-5   1.5e-3 / 1.6e-5
  300       1.6e-2 / 5.3e-6   6.5e-2 / 4.9e-5   1.7e-2 / 2.3e-5
  1000      1.6e-1 / 1.0e-4   6.3e-1 / 5.4e-4   5.3e-1 / 7.6e-4

After 1000 substeps (max / mean):

  N           density           f_p               position
  262,144     1.7e-1 / 1.1e-4   6.3e-1 / 5.4e-4   5.3e-1 / 8.2e-4 h
  1,048,576   2.2e-1 / 7.5e-5   1.4e0  / 3.4e-4   1.0e0  / 5.9e-4 h
  4,194,304   9.0e-2 / 2.6e-5   5.0e-1 / 1.0e-4   4.5e-1 / 2.1e-4 h
  1M learned  3.4e-1 / 1.2e-4   4.9e-1 / 1.1e-4   1.3e0  / 7.5e-4 h

4,194,304 particles, 20 substeps: density 3.4e-4 / 9.7e-6, f_p 7.4e-3 /
5.9e-5, position 3.0e-4 / 9.7e-6 h (the 1M case of the short study reads
3.2e-4, 4.3e-3 and 3.0e-4 h at the same substep).

The means stay at round-off while the maxima grow about 10x per 300
substeps: the splash is chaotic, and particles whose trajectories have
separated by a fraction of h differ in everything.  Bounds are about 3x the
measured values.  The learned warm start, fed the same warm pressures, stays
as close (200 substeps: density 9.0e-3 max, position 5.1e-3 h max).
TEXTBOOK over 200 substeps: 6.46 iterations (FP32) vs 6.47 (FP64), the
same count in 91% of the substeps, never more than one apart.
"""
import dataclasses
import json
import os

import numpy as np
import pytest

from bubbles_amd import scenes

pytestmark = [pytest.mark.gpu]
# the 1000-substep horizons and the drift study take several minutes of
# oracle time each; the 4M-20 and learned 262k-200 cases run by default
_long = pytest.mark.skipif(os.environ.get("BBX_LONG_PRECISION") != "1",
                           reason="long precision study: set BBX_LONG_PRECISION=1")


def _errors(g, o, h):
    rg, ro = g.get("density"), o.get("density")
    drel = np.abs(rg - ro) / ro
    fg, fo = g.get("pressure_force"), o.get("pressure_force")
    fscale = max(np.abs(fo).max(), 1e-30)
    ferr = np.abs(fg - fo).max(axis=1) / fscale
    perr = np.abs(g.get("position") - o.get("position")).max(axis=1) / h
    return {"density_rel_max": float(drel.max()), "density_rel_mean": float(drel.mean()),
            "fp_rel_max": float(ferr.max()), "fp_rel_mean": float(ferr.mean()),
            "pos_over_h_max": float(perr.max()), "pos_over_h_mean": float(perr.mean())}


def _record(key, value):
    print(json.dumps({key: value}), flush=True)
    out = os.environ.get("BBX_PRECISION_OUT")
    if out:
        rec = {}
        if os.path.exists(out):
            with open(out) as f:
                rec = json.load(f)
        rec[key] = value
        with open(out, "w") as f:
            json.dump(rec, f, indent=1)


def _pair(bbx, oracle, dims, tank=(2.0, 1.0, 2.0)):
    sc = scenes.dam_break(*dims, spacing=0.01, tank=tank, seed=2)
    g = bbx.Solver(sc)
    o = oracle.Sim(dataclasses.replace(sc, positions=g.get("position").astype(np.float64)), threads=16)
    return sc, g, o


_EVERY_100 = (1, 10) + tuple(range(100, 1001, 100))


@pytest.mark.timeout(1800)
@pytest.mark.parametrize("dims,tank,steps,checkpoints,bound", [
    pytest.param((64, 64, 64), (2.0, 1.0, 2.0), 1000, (1, 10, 100, 300, 1000),
                 {"density_rel_max": 0.5, "density_rel_mean": 3e-4, "fp_rel_max": 2.0, "fp_rel_mean": 2e-3,
                  "pos_over_h_max": 1.5, "pos_over_h_mean": 3e-3}, id="262k-1000", marks=_long),
    pytest.param((256, 64, 256), (4.0, 1.0, 4.0), 20, (1, 10, 20),
                 {"density_rel_max": 1e-3, "density_rel_mean": 3e-5, "fp_rel_max": 3e-2, "fp_rel_mean": 3e-4,
                  "pos_over_h_max": 1e-3, "pos_over_h_mean": 3e-5}, id="4M-20"),
    pytest.param((128, 64, 128), (2.0, 1.0, 2.0), 1000, _EVERY_100,
                 {"density_rel_max": 1.0, "density_rel_mean": 2e-4, "fp_rel_max": 4.0, "fp_rel_mean": 1e-3,
                  "pos_over_h_max": 3.0, "pos_over_h_mean": 1.5e-3}, id="1M-1000", marks=_long),
    pytest.param((256, 64, 256), (4.0, 1.0, 4.0), 1000, _EVERY_100,
                 {"density_rel_max": 1.0, "density_rel_mean": 1e-4, "fp_rel_max": 4.0, "fp_rel_mean": 5e-4,
                  "pos_over_h_max": 3.0, "pos_over_h_mean": 1e-3}, id="4M-1000", marks=_long),
])
def test_plain_exact_long(bbx, oracle, dims, tank, steps, checkpoints, bound):
    sc, g, o = _pair(bbx, oracle, dims, tank)
    table = {}
    for k in range(1, steps + 1):
        g.pcisph_substep(sc.dt, 5, 0.01, bbx.MODE_REFERENCE_EXACT)
        o.pcisph_substep(sc.dt, 5, 0.01, oracle.MODE_REFERENCE_EXACT)
        if k in checkpoints:
            table[k] = _errors(g, o, sc.kernel_radius)
            print(sc.n, k, table[k], flush=True)
    _record(f"plain_exact_{sc.n}_{steps}", table)
    last = table[steps]
    for k, b in bound.items():
        assert last[k] <= b, (k, last)
    # one substep: arithmetic alone (the f_p maximum is a tail statistic over
    # all particles and grows with the count: 1.4e-4 at 262k, 1.7e-4 at 4M)
    first = table[1]
    assert first["density_rel_max"] <= 1e-6 and first["pos_over_h_max"] <= 1e-5, first
    assert first["fp_rel_max"] <= 5e-4, first


@pytest.mark.timeout(1800)
@pytest.mark.parametrize("dims,steps", [pytest.param((64, 64, 64), 200, id="262k-200"),
                                        pytest.param((128, 64, 128), 1000, id="1M-1000", marks=_long)])
def test_learned_exact_against_fp64(bbx, oracle, dims, steps):
    from bubbles_amd import learned

    sc, g, o = _pair(bbx, oracle, dims)
    g.attach_mlp(learned.load())
    table = {}
    checkpoints = (1, 10, 100, 200) if steps == 200 else _EVERY_100
    for k in range(1, steps + 1):
        g.pcisph_substep(sc.dt, 5, 0.01, bbx.MODE_REFERENCE_EXACT, bbx.WARM_MLP)
        o.pcisph_substep(sc.dt, 5, 0.01, oracle.MODE_REFERENCE_EXACT,
                         warm=g.get("warm_pressure").astype(np.float64))
        if k in checkpoints:
            table[k] = _errors(g, o, sc.kernel_radius)
            print(sc.n, k, table[k], flush=True)
    _record(f"learned_exact_{sc.n}_{steps}", table)
    last = table[steps]
    if steps == 200:
        assert last["density_rel_max"] <= 2e-2 and last["density_rel_mean"] <= 2e-5, last
        assert last["pos_over_h_max"] <= 2e-2 and last["pos_over_h_mean"] <= 5e-5, last
    else:
        # the maximum is a splash particle: alone in one run and among
        # neighbours in the other, its density differs by a multiple of the
        # oracle's (0.34 to 3.3 measured); the means are what the horizon
        # bounds
        assert last["density_rel_max"] <= 10.0 and last["density_rel_mean"] <= 4e-4, last
        assert last["pos_over_h_max"] <= 4.5 and last["pos_over_h_mean"] <= 3e-3, last
    assert max(table[1]["density_rel_max"], table[1]["pos_over_h_max"]) <= 1.5e-4


@_long
@pytest.mark.timeout(1800)
def test_textbook_iteration_drift(bbx, oracle):
    sc, g, o = _pair(bbx, oracle, (64, 64, 64))
    its_g, its_o = [], []
    for _ in range(200):
        its_g.append(g.pcisph_substep(sc.dt, 8, 0.01, bbx.MODE_TEXTBOOK).iterations)
        its_o.append(o.pcisph_substep(sc.dt, 8, 0.01, oracle.MODE_TEXTBOOK).iterations)
    a, b = np.array(its_g), np.array(its_o)
    rec = {"substeps": 200, "mean_iterations_fp32": float(a.mean()), "mean_iterations_fp64": float(b.mean()),
           "equal_fraction": float(np.mean(a == b)), "mean_abs_diff": float(np.mean(np.abs(a - b))),
           "max_abs_diff": int(np.abs(a - b).max()),
           "final": _errors(g, o, sc.kernel_radius)}
    _record(f"textbook_drift_{sc.n}", rec)
    assert abs(a.mean() - b.mean()) <= 0.5 and rec["equal_fraction"] >= 0.7 and rec["max_abs_diff"] <= 2, rec
