"""GPU tests of the batched engine: every mesh's result, alone or packed with
others (admitted at once or mid-stream), equals the CPU oracle's run of that
mesh alone, bit for bit."""
import numpy as np
import pytest
import torch

import oracle
from pamo_amd import batch, simplify_stage2, synth

pytestmark = pytest.mark.gpu


def _same(a, b):
    """Bit-exact equality; for float arrays NaN must sit at the same places,
    with any payload (the NaN that 0/0 yields differs between the host and
    the GPU, and no result depends on it)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b).astype(a.dtype, copy=False)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        na, nb = np.isnan(a), np.isnan(b)
        if not np.array_equal(na, nb):
            return False
        a = np.where(na, 0, a)
        b = np.where(nb, 0, b)
    return a.tobytes() == b.tobytes()


def _two_spheres(level=3, seed=3):
    v1, f1 = synth.displaced_icosphere(level, seed)
    v2, f2 = synth.displaced_icosphere(level, seed + 1)
    v2 = v2 + np.array([0.9, 0.05, 0.0], np.float32)
    return np.concatenate([v1, v2]).astype(np.float32), np.concatenate([f1, f2 + v1.shape[0]]).astype(np.int32)


def _meshes():
    return [synth.displaced_icosphere(4, 11), synth.noisy_torus(48, 20, 6), synth.sdf_union(10, 50, 7),
            _two_spheres(), synth.displaced_icosphere(3, 12), synth.sdf_union(20, 64, 8)]


def _oracle_result(v, f, ratio):
    verts, scale, mean = oracle.preprocess(v, f)
    target = int(ratio * f.shape[0])
    ov, of, st = oracle.run(verts, f, scale, target)
    return ov + mean, of, st


def test_engine_single_mesh_matches_oracle(cuda):
    for v, f in _meshes()[:4]:
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(0.01 * f.shape[0])
        ov, of, st = oracle.run(verts, f, scale, target)
        gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target)
        print("mesh", f.shape, st, {k: gst[k] for k in ("rounds", "exit_reason")}, flush=True)
        assert gst["rounds"] == st["rounds"]
        assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)


def test_batch_equals_alone(cuda):
    meshes = _meshes()
    ratios = [0.01, 0.005, 0.02, 0.01, 0.05, 0.01]
    res = batch.simplify_batch(meshes, ratios)
    for (v, f), r, (bv, bf, info) in zip(meshes, ratios, res):
        ov, of, st = _oracle_result(v, f, r)
        print("batch mesh", f.shape, info["rounds"], st["rounds"], flush=True)
        assert info["rounds"] == st["rounds"]
        assert _same(bv, ov) and _same(bf, of)


def test_admission_mid_stream(cuda):
    """Jobs submitted while others are several rounds in."""
    meshes = _meshes()
    eng = batch.BatchEngine()
    dev = cuda
    handles = {}
    for i in (0, 1):
        v, t, s, m = batch.prepare(*meshes[i], dev)
        handles[eng.submit(v, t, s, int(0.01 * len(meshes[i][1])), job=i)] = (i, m)
    eng.step(3)
    for i in (2, 3):
        v, t, s, m = batch.prepare(*meshes[i], dev)
        handles[eng.submit(v, t, s, int(0.01 * len(meshes[i][1])), job=i)] = (i, m)
    eng.step(5)
    for i in (4, 5):
        v, t, s, m = batch.prepare(*meshes[i], dev)
        handles[eng.submit(v, t, s, int(0.01 * len(meshes[i][1])), job=i)] = (i, m)
    got = {}
    for _ in range(2000):
        _, active = eng.step(1)
        for h, _ in eng.poll():
            v, f, info = eng.take(h)
            i, m = handles[h]
            got[i] = (v.cpu().numpy() + m, f.cpu().numpy())
        if active == 0:
            break
    assert sorted(got) == list(range(6))
    for i, (v, f) in enumerate(meshes):
        ov, of, _ = _oracle_result(v, f, 0.01)
        assert _same(got[i][0], ov) and _same(got[i][1], of), i


@pytest.mark.parametrize("rounds_per_step", [1, 3])
def test_server_results_match(cuda, rounds_per_step):
    """Served results equal the oracle's, with one or several rounds per
    loop iteration; the loop records the engine's peak device bytes, which
    cover at least the packed positions and faces of the admitted meshes."""
    meshes = _meshes()[:4]
    jobs = [dict(points=v, faces=f, ratio=0.01, arrival=0.002 * i) for i, (v, f) in enumerate(meshes)]
    out = {}
    srv = batch.Server(rounds_per_step=rounds_per_step)
    recs = srv.run(jobs, on_result=lambda i, v, f, info: out.__setitem__(i, (v, f)))
    assert all(r is not None and r["latency"] >= 0 for r in recs)
    for i, (v, f) in enumerate(meshes):
        ov, of, _ = _oracle_result(v, f, 0.01)
        assert _same(out[i][0], ov) and _same(out[i][1], of), i
    st = srv.stats
    assert st["live_faces_peak"] >= max(len(f) for _, f in meshes)
    assert st["device_bytes_peak"] >= 24 * st["live_faces_peak"]
    assert srv.eng.ctx.device_bytes() >= st["device_bytes_peak"] // 2 > 0


@pytest.mark.parametrize("guard", [1, 2])
def test_guard_paths_match_oracle(cuda, guard):
    """The LBVH walk (1) and the exhaustive search (2) each give the
    oracle's result, on meshes with and without self-intersections."""
    from pamo_amd import _lib
    ctx = _lib.Context(0)
    ctx.debug_guard(guard)
    for v, f in (_meshes()[1], _two_spheres(), _meshes()[5]):
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(0.01 * f.shape[0])
        ov, of, st = oracle.run(verts, f, scale, target)
        gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target,
                                      ctx=ctx)
        print("guard", guard, "mesh", f.shape, gst["rounds"], st["rounds"], flush=True)
        assert gst["rounds"] == st["rounds"]
        assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)


@pytest.mark.parametrize("guard", [1, 2])
@pytest.mark.parametrize("cap", [0, 333])
def test_pair_overflow_fallback(cuda, cap, guard):
    """A tiny candidate-pair buffer sends most queries (cap 0: all with a
    candidate) to the fallback (the single-kernel walk, or the per-query
    exhaustive test); the result must not change."""
    from pamo_amd import _lib
    ctx = _lib.Context(0)
    ctx.debug_pair_cap(cap)
    ctx.debug_guard(guard)
    for v, f in (_meshes()[0], _two_spheres()):
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(0.01 * f.shape[0])
        ov, of, st = oracle.run(verts, f, scale, target)
        gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target,
                                      ctx=ctx)
        print("cap", cap, "mesh", f.shape, gst["rounds"], st["rounds"], flush=True)
        assert gst["rounds"] == st["rounds"]
        assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)


def test_pamo_simplify_host_abi(cuda):
    """pamo_simplify (the Blender add-on's host-pointer entry point) equals
    the oracle's loop on the raw vertices with the bounding-box scale, no
    size exit and tolerance 4."""
    from pamo_amd import simplify_host
    for v, f in (_meshes()[0], _two_spheres()):
        v = np.ascontiguousarray(v, np.float32)
        scale = float(np.max(v.max(0) - v.min(0)))
        ov, of, st = oracle.run(v, f, scale, -1, 1e-3, 4)
        hv, hf = simplify_host(v, f, 1e-3)
        print("pamo_simplify", f.shape, st, flush=True)
        assert st["exit_reason"] in (3, 4)
        assert _same(hv, ov) and _same(hf, of)


def test_region_packing_after_retirement(cuda):
    """A large mesh with a mild ratio retires after a few rounds while a
    small one runs on: the retired regions then exceed the kept ones, so the
    kept mesh's persistent regions (vertex map, invalid-vertex table, edge
    costs) are moved mid-run. Both results equal the oracle's; a third
    mesh admitted after the move also does."""
    big, small, late = synth.sdf_union(20, 80, 9), synth.displaced_icosphere(3, 12), synth.noisy_torus(32, 12, 4)
    assert big[1].shape[0] > 8 * small[1].shape[0]
    eng = batch.BatchEngine()
    meta = {}

    def submit(mesh, ratio, job):
        v, t, s, m = batch.prepare(*mesh, cuda)
        meta[eng.submit(v, t, s, int(ratio * mesh[1].shape[0]), job=job)] = (job, m)

    submit(big, 0.5, 0)
    submit(small, 0.01, 1)
    got = {}
    small_running = None
    for _ in range(2000):
        _, active = eng.step(1)
        for h, _ in eng.poll():
            v, f, info = eng.take(h)
            job, m = meta[h]
            got[job] = (v.cpu().numpy() + m, f.cpu().numpy(), info["rounds"])
        if 0 in got and small_running is None:
            small_running = 1 not in got
            submit(late, 0.01, 2)
        if active == 0 and len(got) == 3:
            break
    assert small_running, "the small mesh must still run when the large one retires"
    for job, (v, f), r in ((0, big, 0.5), (1, small, 0.01), (2, late, 0.01)):
        ov, of, st = _oracle_result(v, f, r)
        gv, gf, rounds = got[job]
        print("packing", job, f.shape, rounds, st["rounds"], flush=True)
        assert rounds == st["rounds"]
        assert _same(gv, ov) and _same(gf, of), job


def test_long_fans_match_oracle(cuda):
    """Vertices of degree above the fan kernel's LDS capacity (the poles of
    latitude-longitude spheres, 80 and 200 faces; the second keeps a
    187-face fan to the end) take the global-memory sort path; results equal
    the oracle's, alone and batched."""
    meshes = [synth.uv_sphere(80, 40, 21), synth.uv_sphere(200, 24, 22)]
    for v, f in meshes:
        assert np.bincount(f.ravel()).max() > 64
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(0.01 * f.shape[0])
        ov, of, st = oracle.run(verts, f, scale, target)
        gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target)
        print("long fans", f.shape, st, gst["rounds"], flush=True)
        assert gst["rounds"] == st["rounds"]
        assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)
    res = batch.simplify_batch(meshes + [synth.displaced_icosphere(3, 12)], [0.01, 0.01, 0.05])
    for (v, f), r, (bv, bf, info) in zip(meshes, [0.01, 0.01], res):
        ov, of, st = _oracle_result(v, f, r)
        assert info["rounds"] == st["rounds"]
        assert _same(bv, ov) and _same(bf, of)


@pytest.mark.parametrize("guard", [0, 1, 2])
def test_strict_guard_matches_oracle(cuda, guard):
    """The strict guard on the device equals the oracle's, bit for bit, with
    either candidate search; its output passes the exact verifier where the
    reference's guard keeps intersections."""
    from pamo_amd import _lib
    from pamo_amd.verify import verify
    cases = [synth.sdf_union(20, 80, 3, n_slab_pairs=6), synth.displaced_icosphere(4, 5)]
    for v, f in cases:
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(0.001 * f.shape[0])
        ov, of, st = oracle.run(verts, f, scale, target, strict=True)
        ctx = _lib.Context(0)
        ctx.debug_guard(guard)
        gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target,
                                      ctx=ctx, strict=True)
        print("strict", f.shape, guard, st, gst["rounds"], gst["max_undo_passes"], flush=True)
        assert gst["rounds"] == st["rounds"]
        assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)
        assert verify(ov, of)["ok"]


def test_strict_batch_equals_alone(cuda):
    """Strict guard, batched: a mesh whose input already intersects (two
    overlapping spheres) re-marks the same collapses on every undo pass up to
    the strict cap of 64, so its undo list grows far past what the
    reference's 5 passes write. Packed next to other meshes, every mesh's
    result equals its run alone and the oracle's, bit for bit (an undo region
    sized for fewer passes would spill into the next mesh's list)."""
    meshes = [synth.displaced_icosphere(4, 11), _two_spheres(), synth.sdf_union(10, 50, 7)]
    ratios = [0.01, 0.01, 0.01]
    res = batch.simplify_batch(meshes, ratios, engine=batch.BatchEngine(strict=True))
    for (v, f), r, (bv, bf, info) in zip(meshes, ratios, res):
        verts, scale, mean = oracle.preprocess(v, f)
        target = int(r * f.shape[0])
        ov, of, st = oracle.run(verts, f, scale, target, strict=True)
        gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target,
                                      strict=True)
        print("strict batch", f.shape, info["rounds"], info["max_undo_passes"], st, flush=True)
        assert info["rounds"] == st["rounds"] == gst["rounds"]
        assert info["max_undo_passes"] == st["max_undo_passes"] == gst["max_undo_passes"]
        assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)
        assert _same(bv, ov + mean) and _same(bf, of)
    assert res[1][2]["max_undo_passes"] == 64


def test_optimal_placement_matches_oracle(cuda):
    """The opt-in optimal placement on the device equals the oracle's, bit for
    bit, alone and batched; outputs stay closed, manifold and free of
    intersections."""
    from pamo_amd.verify import verify
    meshes = [synth.displaced_icosphere(4, 5), synth.sdf_union(10, 40, 7), _two_spheres(), synth.noisy_torus(48, 20, 6)]
    ratios = [0.02, 0.02, 0.01, 0.01]
    for (v, f), r in zip(meshes, ratios):
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(r * f.shape[0])
        ov, of, st = oracle.run(verts, f, scale, target, placement=1)
        gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target,
                                      placement="optimal")
        print("optimal placement", f.shape, st, gst["rounds"], flush=True)
        assert gst["rounds"] == st["rounds"]
        assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)
    res = batch.simplify_batch(meshes, ratios, engine=batch.BatchEngine(placement="optimal"))
    for (v, f), r, (bv, bf, info) in zip(meshes, ratios, res):
        verts, scale, mean = oracle.preprocess(v, f)
        ov, of, st = oracle.run(verts, f, scale, int(r * f.shape[0]), placement=1)
        assert info["rounds"] == st["rounds"]
        assert _same(bv, ov + mean) and _same(bf, of)
    v, f = meshes[0]
    verts, scale, _ = oracle.preprocess(v, f)
    ov, of, _ = oracle.run(verts, f, scale, int(0.02 * f.shape[0]), placement=1)
    r = verify(ov, of)
    assert r["manifold"] and r["watertight"] and r["intersecting_pairs"] == 0


@pytest.mark.parametrize("setting", [
    dict(ec_lanes=2, morton_bits=6, emit_groups=1),
    dict(ec_lanes=8, bf_max_faces=0, walk8_max=0, walk4_max=0, morton_bits=7),
    dict(bf_split_max_faces=0, bf_max_faces=1 << 30, emit_groups=65536),
    dict(pipeline=0, graph=0),
    dict(pipeline=1, graph=0),
    dict(pipeline=1, graph=1, graph_max_faces=1 << 30, bf_max_faces=1 << 30),
    dict(pipeline=0, graph=1, graph_max_faces=1 << 30, bf_max_faces=0),
])
def test_tuning_does_not_change_results(cuda, setting):
    """Launch shapes from the autotuner's range, including its extremes
    (2 and 8 lanes per edge, 6-bit Morton codes, the LBVH at every size or
    the exhaustive search at every size, one persistent workgroup), and the
    round schedules (rounds queued behind an unread round or not, recorded
    into graphs or launched one by one, with the exhaustive search or the
    LBVH inside the recording): every result equals the oracle's, alone and
    batched, the batch stepping many rounds per call so that queued rounds
    meet marks and finished meshes.  Each mesh runs twice on the same
    context, the second time from the recordings of the first.  The batched
    LBVH with 7-bit codes places each mesh's id at bit 21 of the key; the
    per-mesh leaf ranges must be read from there."""
    from pamo_amd import _lib
    meshes = [_meshes()[0], _two_spheres(), synth.sdf_union(20, 64, 8)]
    ctx = _lib.Context(0, tuning=False)
    for k, v in setting.items():
        ctx.set_tuning(k, v)
    assert all(ctx.tuning()[k] == v for k, v in setting.items())
    for v, f in meshes:
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(0.01 * f.shape[0])
        ov, of, st = oracle.run(verts, f, scale, target)
        checks = []
        for rep in range(2):
            gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale,
                                          target, ctx=ctx)
            print("tuning", setting, f.shape, rep, gst["rounds"], st["rounds"], gst["checks"], flush=True)
            assert gst["rounds"] == st["rounds"]
            assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)
            checks.append(gst["checks"])
        assert checks[0] == checks[1] >= st["rounds"]   # a check per round plus one per undo pass
    res = batch.simplify_batch(meshes, 0.01, engine=batch.BatchEngine(ctx=ctx), rounds_per_step=64)
    for (v, f), (bv, bf, info) in zip(meshes, res):
        ov, of, _ = _oracle_result(v, f, 0.01)
        assert _same(bv, ov) and _same(bf, of)
