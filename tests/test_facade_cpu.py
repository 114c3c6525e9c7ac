"""Host logic of the C++ facade (include/bubbles/bubbles.h) that needs no
device: the OBJ reader behind MakeMesh (LoadObj, src/third/obj_loader.cpp:434-441
and the face records of the Wavefront format) and MakeMesh's vertex
transform and bounds (src/shapes/bvh.cpp:50-61)."""
import json
import os
import subprocess

import pytest

from bubbles_amd import build as hipbuild


def _run(path):
    if not os.path.exists(hipbuild.LOADOBJ_BIN):
        pytest.fail("run python bubbles_amd/build.py")
    res = subprocess.run([hipbuild.LOADOBJ_BIN, str(path)], capture_output=True, text=True, timeout=30)
    return res.returncode, json.loads(res.stdout.strip().splitlines()[-1])


def test_loadobj_triangulates_and_resolves_indices(tmp_path):
    obj = tmp_path / "quad_and_tri.obj"
    obj.write_text(
        "# a square as a quad, then a triangle with negative indices\n"
        "o thing\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
        "vn 0 0 1\nvt 0 0\n"
        "f 1/1/1 2/1/1 3/1/1 4/1/1\n"
        "v 0 0 1\n"
        "f -1 -4 -3\n")
    rc, out = _run(obj)
    assert rc == 0, out
    assert out["name"] == "quad_and_tri.obj" and out["vertices"] == 5
    # fan triangulation of the quad, then the indices of v5, v2, v3
    assert out["triangles"] == [[0, 1, 2], [0, 2, 3], [4, 1, 2]]
    # MakeMesh applies Translate(1, 2, 3) * Scale(2) to the vertices
    assert out["bounds"] == [1.0, 2.0, 3.0, 3.0, 4.0, 5.0]


def test_loadobj_rejects_out_of_range_faces(tmp_path):
    obj = tmp_path / "bad.obj"
    obj.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 7\n")
    rc, out = _run(obj)
    assert rc == 3 and "out of range" in out["error"]


def test_loadobj_missing_file(tmp_path):
    rc, out = _run(tmp_path / "none.obj")
    assert rc == 3 and "cannot open" in out["error"]
