"""Stage-3 inputs shared by the CPU and GPU tests: a target mesh and its
stage-2 simplification (the reference's stage-3 input, pamo/__init__.py:160-176:
the stage-2 vertices with the centring undone, the input mesh as the target)."""
import numpy as np

import oracle
from pamo_amd import synth


def stage2_pair(level=4, seed=1, ratio=0.02):
    """(target V, target F, stage-2 V, stage-2 F) in the input frame, float64."""
    v, f = synth.displaced_icosphere(level, seed)
    verts, scale, mean = oracle.preprocess(v, f)
    target = max(int(ratio * f.shape[0]), 10)
    ov, of, _ = oracle.run(verts, f, scale, target)
    return (np.asarray(v, np.float64), np.asarray(f, np.int32), ov.astype(np.float64) + mean.astype(np.float64),
            of.astype(np.int32))


def tetra():
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    f = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.int32)
    return v, f


def close_spheres(level=3, gap=4e-4, overlap=0.01, r=0.25):
    """Two icospheres of radius r side by side along x, `gap` apart (the mesh
    to project), and the same two spheres overlapping by `overlap` (the
    target), already in the normalised frame: the target term pulls the
    spheres into each other, so the barrier (d_hat 1e-3) and the CCD hold
    them apart.  Returns (target V, target F, V, F), float64 / int32."""
    v, f = synth.icosphere(level)
    v = np.asarray(v, np.float64) * r
    f = np.asarray(f, np.int32)

    def pair(d):
        a = v - [r + 0.5 * d, 0.0, 0.0]
        b = v + [r + 0.5 * d, 0.0, 0.0]
        return np.concatenate([a, b]), np.concatenate([f, f + len(v)]).astype(np.int32)

    gv, gf = pair(-overlap)
    mv, mf = pair(gap)
    return gv, gf, mv, mf
