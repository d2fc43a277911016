"""Mesh files (pamo_amd/meshio.py) and the verifier's command line
(python -m pamo_amd.verify).  CPU only."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from pamo_amd import meshio, synth
from pamo_amd import verify as pv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_MESH = "/root/reference/mesh/Dino_B015CZP872_SmallOrangeTRex.obj"


@pytest.mark.parametrize("ext", [".obj", ".ply"])
def test_round_trip_is_exact_in_float32(tmp_path, ext):
    v, f = synth.displaced_icosphere(3, 2)
    v = v.astype(np.float32)
    p = str(tmp_path / ("m" + ext))
    meshio.save(p, v, f)
    v2, f2 = meshio.load(p)
    assert v2.shape == v.shape and f2.shape == f.shape
    assert np.array_equal(v2.astype(np.float32), v)
    assert np.array_equal(f2, f)


def test_obj_tokens_polygons_and_negative_indices(tmp_path):
    p = tmp_path / "poly.obj"
    p.write_text("# comment\nmtllib x.mtl\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0 1.0\n"
                 "vt 0 0\nvn 0 0 1\n"
                 "g quad\nf 1/1/1 2/1/1 3/1/1 4/1/1\n"
                 "usemtl m\nf -5//1 -4//1 -3//1 -1//1 -2//1\n"
                 "f 1 2 3\n")
    v, f = meshio.load(str(p))
    assert v.shape == (5, 3) and v[4].tolist() == [0.5, 1.5, 0.0]
    # quad -> (0,1,2),(2,3,0); pentagon -> fan from its first corner
    assert f.tolist() == [[0, 1, 2], [2, 3, 0],
                          [0, 1, 2], [0, 2, 4], [0, 4, 3],
                          [0, 1, 2]]


def test_ascii_ply_with_quads_and_extra_properties(tmp_path):
    p = tmp_path / "a.ply"
    p.write_text("ply\nformat ascii 1.0\ncomment made by hand\nelement vertex 4\n"
                 "property float x\nproperty float y\nproperty float z\nproperty uchar red\n"
                 "element face 2\nproperty list uchar int vertex_indices\nproperty int flags\nend_header\n"
                 "0 0 0 9\n1 0 0 9\n1 1 0 9\n0 1 0 9\n"
                 "4 0 1 2 3 7\n3 3 2 1 0\n")
    v, f = meshio.load(str(p))
    assert v.tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
    assert f.tolist() == [[0, 1, 2], [2, 3, 0], [3, 2, 1]]


def test_bad_index_and_format_raise(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    with pytest.raises(ValueError):
        meshio.load(str(p))
    with pytest.raises(ValueError):
        meshio.load(str(tmp_path / "x.stl"))


@pytest.mark.skipif(not os.path.exists(REF_MESH), reason="reference demo meshes not present")
def test_reads_reference_demo_mesh():
    """The Dino mesh that demo.sh runs: counts from its own v/f records."""
    v, f = meshio.load(REF_MESH)
    assert v.shape == (52433, 3) and f.shape == (97112, 3)
    assert f.min() == 0 and f.max() == 52432


def _cli(*args):
    return subprocess.run([sys.executable, "-m", "pamo_amd.verify", *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_verify_cli_exit_status_and_json(tmp_path):
    v, f = synth.displaced_icosphere(2, 1)
    good = str(tmp_path / "good.ply")
    meshio.save(good, v, f)
    r = _cli(good, "--json")
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["ok"] and out["watertight"] and out["intersecting_pairs"] == 0
    assert out["faces"] == len(f) and out["euler"] == 2
    holed = str(tmp_path / "holed.obj")
    meshio.save(holed, v, f[1:])
    r = _cli(holed)
    assert r.returncode == 1
    assert "boundary_edges" in r.stdout and "False" in r.stdout
    w = v.copy()
    w[0] = -w[0] * 1.5
    bad = str(tmp_path / "x.obj")
    meshio.save(bad, w, f)
    out = json.loads(_cli(bad, "--json", "--max-pairs", "4").stdout)
    assert out["intersecting_pairs"] > 0 and 1 <= len(out["pairs"]) <= 4
    assert pv.verify(w, f)["intersecting_pairs"] == out["intersecting_pairs"]
