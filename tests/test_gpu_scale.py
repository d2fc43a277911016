"""GPU parity at the benchmark configurations' sizes and shapes, bit for bit
against the CPU oracle's run of the same loop (simp_cuda/pamo/__init__.py:
128-158).

The device switches paths by size: the LBVH instead of the exhaustive search
above 262,144 alive faces, the walk's lanes per query (8 up to 65,536 alive
faces, 4 up to 262,144, 1 above) and the scan tile shapes (1024 x 8 above
2^20 items).  These cases run each of those paths at its real size: the
2.0M-face union of the plain benchmark to the end, a 1.08M-face union with
thin slab pairs (the stress family) to the end, a 277k-face union for its
first rounds with each candidate search forced, a many-mesh batch drawn from
the batch workload's generator, and a Poisson stream through the server."""
import numpy as np
import pytest
import torch

import oracle
from pamo_amd import _lib, batch, simplify_stage2, synth

pytestmark = pytest.mark.gpu


def _same(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b).astype(a.dtype, copy=False)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        na, nb = np.isnan(a), np.isnan(b)
        if not np.array_equal(na, nb):
            return False
        a, b = np.where(na, 0, a), np.where(nb, 0, b)
    return a.tobytes() == b.tobytes()


def _check_run(cuda, v, f, ratio, iters=1000000, ctx=None, label=""):
    """simplify_stage2 against oracle.run: rounds, every round's face count,
    and the final vertices and faces, bit for bit."""
    verts, scale, _ = oracle.preprocess(v, f)
    target = int(ratio * f.shape[0])
    print("case", label, f.shape[0], "faces, iters", iters, flush=True)
    gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target,
                                  iters=iters, ctx=ctx)
    ov, of, st, orf = oracle.run(verts, f, scale, target, iters=iters, return_rounds=True)
    print("  rounds gpu %d oracle %d, final faces %d %d" % (gst["rounds"], st["rounds"], gf.shape[0], of.shape[0]),
          flush=True)
    assert gst["rounds"] == st["rounds"]
    assert [int(x) for x in gst["round_faces"]] == [int(x) for x in orf]
    assert gst["max_undo_passes"] == st["max_undo_passes"]
    assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)
    return gst


def _two_spheres(level=3, seed=3):
    v1, f1 = synth.displaced_icosphere(level, seed)
    v2, f2 = synth.displaced_icosphere(level, seed + 1)
    v2 = v2 + np.array([0.9, 0.05, 0.0], np.float32)
    return np.concatenate([v1, v2]).astype(np.float32), np.concatenate([f1, f2 + v1.shape[0]]).astype(np.int32)


@pytest.mark.parametrize("kshift", [0, 2])
def test_walk_split_forced_small(cuda, kshift):
    """The LBVH walk with 1 and 4 lanes per query (the splits of meshes
    above 262,144 and 65,536 faces) on small meshes, with and without
    self-intersections."""
    ctx = _lib.Context(0)
    ctx.debug_guard(1)
    ctx.debug_walk_split(kshift)
    for v, f in (_two_spheres(), synth.sdf_union(20, 64, 8)):
        _check_run(cuda, v, f, 0.01, ctx=ctx, label="walk split %d" % kshift)


@pytest.mark.parametrize("guard", [0, 1, 2])
def test_277k_union_first_rounds(cuda, guard):
    """A 277,372-face union, rounds 1 to 3 (each run from the start and
    compared whole): by size (0) the first check is the LBVH with one lane
    per query; (1) the LBVH throughout, four lanes from the second round on;
    (2) the exhaustive search at that size."""
    v, f = synth.sdf_union(40, 190, 11)
    assert f.shape[0] > 262144
    ctx = _lib.Context(0)
    ctx.debug_guard(guard)
    for iters in (1, 2, 3):
        _check_run(cuda, v, f, 0.001, iters=iters, ctx=ctx, label="277k guard %d" % guard)


def test_277k_union_full_run(cuda):
    v, f = synth.sdf_union(40, 190, 11)
    _check_run(cuda, v, f, 0.001, label="277k full")


def test_sdf2m_full_run(cuda):
    """The plain benchmark's mesh (smooth union of 200 spheres, seed 2, 300^3
    lattice: 2,017,532 faces, ratio 0.001) run to its end: 210 rounds,
    through every size class of the device's paths."""
    v, f = synth.sdf_union(200, 300, seed=2, k=0.02)
    assert f.shape[0] > 2_000_000
    st = _check_run(cuda, v, f, 0.001, label="sdf2m")
    assert st["rounds"] > 100


def test_slab_union_1m_full_run(cuda):
    """The stress family (600 spheres and 64 thin slab pairs, seed 3) on a
    230^3 lattice: 1,076,568 faces, above 2^20, so its first rounds' face
    scans take the 1024 x 8 tiles; run to the end."""
    v, f = synth.sdf_union(600, 230, seed=3, k=0.02, n_slab_pairs=64)
    assert f.shape[0] > (1 << 20)
    _check_run(cuda, v, f, 0.001, label="slab 1.08M")


def _draws(n, seed_base):
    specs = synth.batch_specs(n, seed_base=seed_base, fmax=60_000)
    return synth.batch_meshes(specs)


def test_batch_draws_equal_alone(cuda):
    """32 meshes drawn from the batch workload's generator (families mixed in
    thirds, log-uniform sizes, ratios 0.01 / 0.005 / 0.001; here up to 60k
    faces) in one packed engine: each result equals the oracle's run of that
    mesh alone."""
    meshes = _draws(32, 4_000_000)
    res = batch.simplify_batch([(v, f) for v, f, _ in meshes], [r for _, _, r in meshes])
    for i, ((v, f, r), (bv, bf, info)) in enumerate(zip(meshes, res)):
        verts, scale, mean = oracle.preprocess(v, f)
        ov, of, st = oracle.run(verts, f, scale, int(r * f.shape[0]))
        assert info["rounds"] == st["rounds"], i
        assert _same(bv, ov + mean) and _same(bf, of), i


def test_server_poisson_stream_equals_alone(cuda):
    """A 32-job Poisson stream (200 jobs/s) of the serving workload's
    generator through the continuous-batching server: jobs join while others
    are mid-run, and each result equals the oracle's run of that job alone."""
    meshes = _draws(32, 5_000_000)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / 200, size=len(meshes)))
    jobs = [dict(points=v, faces=f, ratio=r, arrival=float(t)) for (v, f, r), t in zip(meshes, arrivals)]
    out = {}
    srv = batch.Server()
    recs = srv.run(jobs, on_result=lambda i, v, f, info: out.__setitem__(i, (v, f)))
    assert all(r is not None for r in recs)
    assert srv.stats["max_live"] > 1
    for i, (v, f, r) in enumerate(meshes):
        verts, scale, mean = oracle.preprocess(v, f)
        ov, of, _ = oracle.run(verts, f, scale, int(r * f.shape[0]))
        assert _same(out[i][0], ov + mean) and _same(out[i][1], of), i
