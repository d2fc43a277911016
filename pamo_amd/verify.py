"""Exact mesh verifier: manifoldness, watertightness, degenerate faces and
self-intersections with exact predicates (libpamo_verify.so, built from
csrc/verify.cpp).  It stands in for the reference's offline CGAL checker
(simp_cuda/safe_project/scripts/cgal_intersect/main.cpp), which needs CGAL.

Command line:  python -m pamo_amd.verify mesh.obj [--max-pairs N] [--json]
Exit status 0 when the mesh is closed, manifold, free of degenerate faces
and free of self-intersections; 1 otherwise.
"""
import ctypes
import json
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libpamo_verify.so")
_vf = None


def _lib():
    global _vf
    if _vf is None:
        if not os.path.exists(_PATH):
            raise OSError("libpamo_verify.so not found at %s: build it with __graft_entry__.build()" % _PATH)
        _vf = ctypes.CDLL(_PATH)
        P = ctypes.c_void_p
        _vf.vf_check.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_longlong]
        _vf.vf_orient3d.argtypes = [P, P, P, P]
        _vf.vf_orient3d_exact.argtypes = [P, P, P, P]
        _vf.vf_tri_pair.argtypes = [P, P, P, P]
    return _vf


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


VERIFY_KEYS = ["faces", "vertices", "edges", "boundary_edges", "nonmanifold_edges", "misoriented_edges",
               "nonmanifold_vertices", "unreferenced_vertices", "degenerate_faces", "intersecting_pairs",
               "euler", "bad_indices"]


def verify(points, tris, max_pairs=64):
    """Exact check: manifold, watertight, degenerate faces, self-intersections."""
    points = np.ascontiguousarray(points, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    st = np.zeros(12, np.int64)
    pairs = np.zeros((max_pairs, 2), np.int32)
    _lib().vf_check(_ptr(points), _ptr(tris), points.shape[0], tris.shape[0], _ptr(st), _ptr(pairs), max_pairs)
    res = dict(zip(VERIFY_KEYS, [int(x) for x in st]))
    res["pairs"] = pairs[:min(max_pairs, res["intersecting_pairs"])]
    res["manifold"] = (res["nonmanifold_edges"] == 0 and res["nonmanifold_vertices"] == 0
                       and res["misoriented_edges"] == 0 and res["bad_indices"] == 0)
    res["watertight"] = res["boundary_edges"] == 0 and res["nonmanifold_edges"] == 0
    res["ok"] = (res["manifold"] and res["watertight"] and res["degenerate_faces"] == 0
                 and res["intersecting_pairs"] == 0)
    return res


def orient3d(a, b, c, d, exact_only=False):
    f = _lib().vf_orient3d_exact if exact_only else _lib().vf_orient3d
    arrs = [np.ascontiguousarray(x, np.float64).reshape(3) for x in (a, b, c, d)]
    return int(f(*[_ptr(x) for x in arrs]))


def tri_pair_intersects(a, b, ia, ib):
    a = np.ascontiguousarray(a, np.float32).reshape(9)
    b = np.ascontiguousarray(b, np.float32).reshape(9)
    ia = np.ascontiguousarray(ia, np.int32).reshape(3)
    ib = np.ascontiguousarray(ib, np.int32).reshape(3)
    return bool(_lib().vf_tri_pair(_ptr(a), _ptr(b), _ptr(ia), _ptr(ib)))




def main(argv=None):
    import argparse
    from . import meshio
    ap = argparse.ArgumentParser(prog="python -m pamo_amd.verify", description=__doc__.split("\n\n")[0])
    ap.add_argument("mesh", help=".obj or .ply")
    ap.add_argument("--max-pairs", type=int, default=16, help="intersecting pairs to list")
    ap.add_argument("--json", action="store_true", help="print the result as one JSON line")
    a = ap.parse_args(argv)
    v, f = meshio.load(a.mesh)
    r = verify(v, f, a.max_pairs)
    out = {k: (bool(x) if isinstance(x, (bool, np.bool_)) else int(x)) for k, x in r.items() if k != "pairs"}
    out["pairs"] = r["pairs"].tolist()
    if a.json:
        print(json.dumps(out))
    else:
        for k in VERIFY_KEYS + ["manifold", "watertight", "ok"]:
            print("%-22s %s" % (k, out[k]))
        for p in out["pairs"]:
            print("intersecting faces %d %d" % tuple(p))
    return 0 if r["ok"] else 1


if __name__ == "__main__":
    sys.exit(main())
