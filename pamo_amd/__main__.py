"""Command line mirroring the reference's example.py: load a mesh, run PaMO
(stage 1 remeshing, stage 2 simplification, stage 3 safe projection) on the
GPU, write the result.

    python -m pamo_amd -i in.obj -o out.obj -r 0.001 [-mv 0] [--disable_stage1] [--disable_stage3] [--verify]

Options follow example.py:13-21 (-i/--input, -o/--output, -r/--ratio,
-mv/--min-vertex, --disable_stage1, --disable_stage3): every stage runs
unless disabled, as there.  ``--verify`` runs the exact verifier on the
output.
"""
import argparse
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pamo_amd")
    ap.add_argument("-i", "--input", type=str, required=True, help=".obj or .ply")
    ap.add_argument("-o", "--output", type=str, required=True, help=".obj or .ply")
    ap.add_argument("-r", "--ratio", type=float, default=0.1)
    ap.add_argument("-mv", "--min-vertex", type=int, default=0)
    ap.add_argument("--disable_stage1", action="store_true", help="Disable remeshing")
    ap.add_argument("--disable_stage3", action="store_true", help="Disable safe projection")
    ap.add_argument("--threshold", type=float, default=1e-3)
    ap.add_argument("--tolerance", type=int, default=4)
    ap.add_argument("--verify", action="store_true", help="check the output with the exact verifier")
    a = ap.parse_args(argv)

    import torch
    from . import PaMO, meshio
    v, f = meshio.load(a.input)
    print("# of input verts : {}".format(len(v)))
    print("# of input faces : {}".format(len(f)))
    pamo = PaMO((v, f), use_stage1=not a.disable_stage1, use_stage3=not a.disable_stage3)
    points = torch.from_numpy(v).float().cuda()
    tris = torch.from_numpy(f).int().cuda()
    start = time.time()
    verts, faces = pamo.run(points, tris, min_verts=a.min_vertex, ratio=a.ratio, tolerance=a.tolerance,
                            threshold=a.threshold)
    end = time.time()
    print("Total time: ", end - start)
    print("# of output verts : {}".format(len(verts)))
    print("# of output faces : {}".format(len(faces)))
    meshio.save(a.output, verts, faces)
    if a.verify:
        from .verify import verify
        r = verify(verts, faces)
        print("verify: manifold=%s watertight=%s degenerate_faces=%d intersecting_pairs=%d" %
              (r["manifold"], r["watertight"], r["degenerate_faces"], r["intersecting_pairs"]))
        return 0 if r["ok"] else 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
