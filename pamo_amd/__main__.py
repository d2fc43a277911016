"""Command line mirroring the reference's example.py: load a mesh, run
PaMO's stage 2 on the GPU, write the result.

    python -m pamo_amd -i in.obj -o out.obj -r 0.001 [-mv 0] [--verify]

Options follow example.py:13-21 (-i/--input, -o/--output, -r/--ratio,
-mv/--min-vertex, --disable_stage1, --disable_stage3).  Stages 1 and 3 are
not part of pamo_amd, so the run is always stage 2 alone, as example.py
with both --disable flags; the flags are accepted for compatibility.
``--verify`` runs the exact verifier on the output.
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
    ap.add_argument("--disable_stage1", action="store_true", help="accepted; stage 1 is not part of pamo_amd")
    ap.add_argument("--disable_stage3", action="store_true", help="accepted; stage 3 is not part of pamo_amd")
    ap.add_argument("--threshold", type=float, default=1e-3)
    ap.add_argument("--tolerance", type=int, default=4)
    ap.add_argument("--verify", action="store_true", help="check the output with the exact verifier")
    a = ap.parse_args(argv)

    import torch
    from . import PaMO, meshio
    if not (a.disable_stage1 and a.disable_stage3):
        print("note: stages 1 and 3 are not part of pamo_amd; running stage 2 only")
    v, f = meshio.load(a.input)
    print("# of input verts : {}".format(len(v)))
    print("# of input faces : {}".format(len(f)))
    pamo = PaMO((v, f), use_stage1=False, use_stage3=False)
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
