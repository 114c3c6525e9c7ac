"""The C++ facade (include/bubbles/bubbles.h) drives the same device path as
the C ABI: a scene written in the reference's style gives bit-identical
results to the same scene run through the ABI, and its Setup scalars and
grid equal the oracle's."""
import json
import os
import subprocess

import numpy as np
import pytest

from bubbles_amd import build as hipbuild
from bubbles_amd import scenes


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
