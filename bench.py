#!/usr/bin/env python3
"""Stage-2 benchmark on one MI355X: input Mfaces/s to 0.1% of the faces.

One step = one full guarded simplification (the stage-2 loop of PaMO.run,
ratio 0.001, tolerance 4, threshold 1e-3, min_verts 0) of a seeded 2M-face
SDF sphere-union mesh whose vertices and faces are already resident in HBM.
value = input faces / seconds per step / 1e6.

Workloads (--workload): sdf2m (default, 2.0M faces), sdf7m (7.05M faces
with 64 thin slab pairs: the guard-heavy case), ico20k (20,480-face
displaced icosphere, ratio 0.01), batch512 (512 meshes in one packed
engine) and serve (Poisson arrivals into the continuous-batching server).

Also reported on the same line (plain workloads):
  roofline        the edge-cost kernel's algorithmic bytes per launch over
                  its average launch time (HIP events around every launch in
                  the timed steps), against the 8 TB/s HBM peak;
  kernel_classes  the same for the quadric (fan slots), edge cost, claim and
                  LBVH build classes, and the traversal's time, from one
                  extra untimed run with every class timed;
  cpu_baseline    the CPU oracle (the reference's stage-2 logic restated in
                  C++/OpenMP) on the host cores, timed on the first rounds of
                  the same workload and scaled by the work left (see sample),
                  and the same on one thread (one_thread);
  fp64_error      the first round's edge costs and collapse positions in FP32
                  (the reference's arithmetic) against the same expressions
                  in FP64: relative cost error, changed decisions (winners);
  verify          the exact FP64 check of the output (manifold, watertight,
                  self-intersections); rounds_match_gpu is the per-round face
                  counts' agreement with the oracle on the rounds it ran.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "input Mfaces/s to 0.1% (plain vs batched/served); p50/p99 job latency"
EXIT_NAMES = {1: "target", 2: "10 faces", 3: "tolerance", 4: "iteration cap"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="sdf2m", choices=["sdf2m", "sdf7m", "ico20k", "batch512", "serve", "stage3"])
    p.add_argument("--meshes", type=int, default=512, help="batch512: number of meshes")
    p.add_argument("--jobs", type=int, default=1000, help="serve: number of jobs")
    p.add_argument("--rate", default="200", help="serve: Poisson arrival rate(s) in jobs/s, comma-separated")
    p.add_argument("--seq-sample", type=int, default=8, help="batch/serve: plain-path timing on every n-th mesh")
    p.add_argument("--grid", type=int, default=0,
                   help="SDF lattice (default: 300 for sdf2m, ~2.0M faces; 584 for sdf7m, ~7.05M faces)")
    p.add_argument("--cpu1-seconds", type=float, default=10.0,
                   help="budget of the single-thread CPU sample (0: skip)")
    p.add_argument("--cpu-seconds", type=float, default=20.0,
                   help="budget of the all-cores CPU sample where the oracle does not run to completion (sdf7m)")
    p.add_argument("--fp64-seconds", type=float, default=150.0,
                   help="budget of the whole-run FP64 shadow (it stops after the round that passes it)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-full", action="store_true",
                   help="sdf7m: run the oracle's whole loop (one forward per round, progress lines) instead of a "
                        "sample, and compare its output with the GPU's bit for bit")
    p.add_argument("--no-verify", action="store_true")
    p.add_argument("--placement", default="midpoint", choices=["midpoint", "optimal"],
                   help="plain workloads: where a collapse places its vertex (optimal: not the reference's)")
    p.add_argument("--strict", action="store_true",
                   help="plain workloads: the strict guard (exact pair test, undo until clean; not the reference's)")
    p.add_argument("--save-mesh", default="",
                   help="plain workloads: write the output mesh (.obj or .ply) here, in the centred frame the "
                        "guard works in")
    p.add_argument("--no-classes", action="store_true",
                   help="skip the extra untimed run that times every kernel class (for kernel traces of one run)")
    p.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic.json"))
    p.add_argument("--no-legs", action="store_true",
                   help="sdf2m: skip the batched and served legs of the default line")
    p.add_argument("--leg-meshes", type=int, default=512, help="batched leg: meshes of the many-mesh mix")
    p.add_argument("--leg-jobs", type=int, default=1000, help="served leg: jobs of the Poisson stream")
    p.add_argument("--leg-rate", type=float, default=350.0, help="served leg: arrival rate in jobs/s")
    p.add_argument("--serve-rounds", type=int, default=1, help="server: rounds per loop iteration")
    p.add_argument("--copy-probe", type=int, default=0,
                   help="MiB copied device-to-device once before timing: a copy kernel of known size that "
                        "calibrates the FETCH_SIZE / WRITE_SIZE counters in a rocprofv3 --pmc run")
    return p.parse_args()


def make_mesh(args):
    from pamo_amd import synth
    t = time.time()
    if args.workload in ("sdf2m", "stage3"):
        grid = args.grid or 300
        v, f = synth.sdf_union(200, grid, seed=2, k=0.02)
        cfg = dict(workload="sdf_union_2M", generator="smooth-min union of 200 spheres (k=0.02), seed 2",
                   grid=grid, ratio=0.001)
    elif args.workload == "sdf7m":
        grid = args.grid or 584
        v, f = synth.sdf_union(600, grid, seed=3, k=0.02, n_slab_pairs=64)
        cfg = dict(workload="sdf_union_slabs_7M",
                   generator="smooth-min union of 600 spheres (k=0.02) and 64 thin slab pairs "
                             "(thickness U(1e-3,5e-3), gap about one lattice step), seed 3",
                   grid=grid, ratio=0.001)
    else:
        v, f = synth.displaced_icosphere(5, 1)
        cfg = dict(workload="displaced_icosphere_20k", generator="icosphere L5 + 16 Gaussian bumps, seed 1",
                   ratio=0.01)
    log("mesh %s: %d verts %d faces (%.1fs)" % (cfg["workload"], v.shape[0], f.shape[0], time.time() - t))
    return v, f, cfg


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (each GPU would simplify its own "
              "meshes independently); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch512":
        return run_batch(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "stage3":
        return run_stage3(args)
    return run_plain(args)


def run_plain(args):
    import numpy as np
    import torch

    from pamo_amd import _lib, preprocess_numpy, simplify_stage2

    v, f, cfg = make_mesh(args)
    F0 = int(f.shape[0])
    verts, scale, tris_mean = preprocess_numpy(v, f)
    ratio = cfg.pop("ratio")
    target = int(ratio * F0)
    dev = torch.device("cuda:0")
    torch.zeros(1, device=dev)
    torch.cuda.synchronize()
    t_up = time.perf_counter()
    dv = torch.from_numpy(verts).to(dev)
    df = torch.from_numpy(f).to(dev)
    torch.cuda.synchronize()
    h2d_ms = 1e3 * (time.perf_counter() - t_up)   # host buffers -> HBM (pageable), outside the timed region
    ctx = _lib.Context(0)
    ctx.set_strict(args.strict)
    ctx.set_placement(args.placement)
    if args.copy_probe:
        src = torch.ones(args.copy_probe << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: %d bytes" % src.numel())
        del src, dst
    for w in range(args.warmup):
        _, _, st = simplify_stage2(dv, df, scale, target, ctx=ctx)
        torch.cuda.synchronize()
        log("warmup %d: rounds %d" % (w, st["rounds"]))
    dom = _lib.K_EDGE_COST
    ctx.profile([dom])
    times = []
    for s in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ov, of, st = simplify_stage2(dv, df, scale, target, ctx=ctx)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        times.append(t1 - t0)
        log("step %d: %.1f ms, rounds %d, final faces %d" % (s, 1e3 * (t1 - t0), st["rounds"], of.shape[0]))
    prof = ctx.profile_read(dom)
    ctx.profile([])
    classes = None
    if not args.no_classes:
        classes = _class_profile(ctx, lambda: (simplify_stage2(dv, df, scale, target, ctx=ctx),
                                               torch.cuda.synchronize()))
    sec = float(np.mean(times))
    value = F0 / sec / 1e6
    avg_ms = prof["ms"] / max(prof["launches"], 1)
    avg_bytes = prof["bytes"] / max(prof["launches"], 1)
    achieved = (avg_bytes / (avg_ms / 1e3)) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    if os.path.exists(args.traffic_file):
        try:
            traffic = json.load(open(args.traffic_file)).get(_lib.KERNEL_NAMES[dom], {}).get(args.workload)
        except Exception:
            traffic = None
    roofline = dict(kernel=_lib.KERNEL_NAMES[dom], bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    launches=prof["launches"], avg_launch_us=round(1e3 * avg_ms, 3),
                    avg_algorithmic_bytes=round(avg_bytes))
    out_v = ov.cpu().numpy()
    out_f = of.cpu().numpy()
    if args.save_mesh:
        from pamo_amd import meshio
        meshio.save(args.save_mesh, out_v, out_f)
    round_faces = [int(x) for x in st["round_faces"]]
    result = dict(metric=METRIC, value=round(value, 3),
                  unit="Mfaces/s",
                  n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(1e3 * sec, 3),
                  higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
                  data="synthetic (seeded SDF mesh generated in-process)",
                  config=dict(cfg, faces=F0, vertices=int(v.shape[0]), ratio=ratio, tolerance=4, threshold=1e-3,
                              min_verts=0, target_faces=target, rounds=st["rounds"], final_faces=int(out_f.shape[0]),
                              exit_reason=st["exit_reason"],
                              stop="target %d faces (%g%%) or %d rounds without change (tolerance); this run "
                                   "ends by %s at %d faces (%.3f%%)" % (
                                       target, 100 * ratio, 4, EXIT_NAMES.get(st["exit_reason"], "?"),
                                       int(out_f.shape[0]), 100.0 * out_f.shape[0] / F0),
                              parallelism="single GPU", strict_guard=args.strict, placement=args.placement,
                              tuning=dict(ctx.tuning(), source=("PAMO_TUNING_FILE" if os.environ.get("PAMO_TUNING_FILE")
                                                               else "pamo_amd/tuning.json" if _lib.tuning_cache()
                                                               else "defaults")),
                              guard_checks=st["checks"], query_faces=st["query_faces"],
                              candidate_pairs=st["candidate_pairs"]),
                  latency_ms=dict(p50=round(1e3 * float(np.percentile(times, 50)), 3),
                                  p99=round(1e3 * float(np.percentile(times, 99)), 3)),
                  roofline=roofline, kernel_classes=classes,
                  h2d_ms=round(h2d_ms, 3))
    if not args.no_cpu:
        ncores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        if args.workload == "sdf7m" and not args.cpu_full:
            result["cpu_baseline"] = _cpu_plain(verts, f, scale, target, round_faces, ncores, args.cpu_seconds)
        else:
            result["cpu_baseline"] = _cpu_full(verts, f, scale, target, round_faces, ncores, out_v, out_f,
                                               args.placement, rounds_loop=args.workload == "sdf7m")
        if args.cpu1_seconds > 0:
            one = _cpu_plain(verts, f, scale, target, round_faces, 1, args.cpu1_seconds)
            result["cpu_baseline"]["one_thread"] = {k: one[k] for k in ("value", "unit", "cores", "sample",
                                                                        "rounds_match_gpu")}
        # FP32 (the reference's arithmetic, which the GPU matches bit for bit)
        # against FP64: every round's costs and winners, and the output's
        # positions against exact averaging (oracle.fp64_shadow)
        import oracle
        oracle.set_threads(ncores)
        budget = args.fp64_seconds
        log("fp64 shadow: %d threads, budget %.0fs" % (ncores, budget))
        sh = oracle.fp64_shadow(verts, f, scale, target, max_seconds=budget)
        result["fp64_error"] = dict(sh, scope="every round of the run" if sh["exit_reason"] != 5 else
                                    "the first %d rounds (time budget %gs)" % (sh["rounds"], budget),
                                    position_unit="fraction of the scale (largest extent)")
    if not args.no_verify:
        from pamo_amd.verify import verify
        r = verify(out_v, out_f)
        result["verify"] = {k: r[k] for k in ("manifold", "watertight", "degenerate_faces", "intersecting_pairs",
                                               "ok")}
        result["verify"]["max_undo_passes"] = st["max_undo_passes"]
    if args.workload == "sdf2m" and not args.no_legs and not args.strict and args.placement == "midpoint":
        result["placement_optimal"] = _placement_leg(args, ctx, dv, df, scale, target, verts, out_v, out_f, sec,
                                                     st["rounds"])
        result["stage1"] = _stage1_leg(args, v, f, ctx)
        result["stage3"] = _stage3_leg(args, v, f, out_v + tris_mean, out_f, ctx)
    if args.workload == "sdf2m" and not args.no_legs and not args.strict and args.placement == "midpoint":
        del dv, df, ov, of
        ctx.close()
        torch.cuda.empty_cache()
        result["batched"] = _batched_leg(args)
        torch.cuda.empty_cache()
        result["served"] = _served_leg(args)
    print(json.dumps(result), flush=True)


def surface_distance(points, out_v, out_f, n=20000, seed=0):
    """Distance from a seeded sample of n input points to the output surface
    (closest point on each triangle: the plane projection when it falls
    inside, else the nearest edge point), in float64 on the GPU.  Returns
    mean and max, relative to the input's bounding-box diagonal."""
    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(points), size=min(n, len(points)), replace=False)
    p = torch.from_numpy(points[idx].astype(np.float64)).to(dev)
    tri = torch.from_numpy(out_v.astype(np.float64)[out_f]).to(dev)   # [F,3,3]
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    ab, ac = b - a, c - a
    nrm = torch.cross(ab, ac, dim=1)
    nn = (nrm * nrm).sum(1)
    g00, g01, g11 = (ab * ab).sum(1), (ab * ac).sum(1), (ac * ac).sum(1)
    det = g00 * g11 - g01 * g01
    best = []
    for i in range(0, len(p), 1024):
        q = p[i:i + 1024, None, :]                                   # [n,1,3]
        ap = q - a[None]
        d1, d2 = (ap * ab[None]).sum(2), (ap * ac[None]).sum(2)
        s_ = (g11 * d1 - g01 * d2) / det
        t_ = (g00 * d2 - g01 * d1) / det
        inside = (s_ >= 0) & (t_ >= 0) & (s_ + t_ <= 1) & (nn > 0)
        dist = torch.where(inside, ((ap * nrm[None]).sum(2)).abs() / nn.sqrt(), torch.full_like(d1, float("inf")))
        for u, v in ((a, b), (b, c), (c, a)):
            e = v - u
            h = (((q - u[None]) * e[None]).sum(2) / (e * e).sum(1)).clamp(0, 1)
            dist = torch.minimum(dist, (q - (u[None] + h[..., None] * e[None])).norm(dim=2))
        best.append(dist.min(1).values)
    d = torch.cat(best)
    diag = float(np.linalg.norm(points.max(0) - points.min(0)))
    return dict(mean=round(float(d.mean()) / diag, 7), max=round(float(d.max()) / diag, 7))


def _placement_leg(args, ctx, dv, df, scale, target, verts, out_v, out_f, sec_mid, rounds_mid):
    """The opt-in optimal placement (not the reference's) on the same mesh and
    target: its time, rounds and faces beside the midpoint run's, and both
    outputs' distance to the input surface."""
    import numpy as np
    import torch
    from pamo_amd import simplify_stage2
    ctx.set_placement("optimal")
    simplify_stage2(dv, df, scale, target, ctx=ctx)
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ov, of, st = simplify_stage2(dv, df, scale, target, ctx=ctx)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ctx.set_placement("midpoint")
    sec = float(np.mean(times))
    o_v, o_f = ov.cpu().numpy(), of.cpu().numpy()
    out = dict(ms_per_step=round(1e3 * sec, 3), value=round(len(df) / sec / 1e6, 3), unit="Mfaces/s",
               rounds=st["rounds"], final_faces=int(o_f.shape[0]),
               midpoint=dict(ms_per_step=round(1e3 * sec_mid, 3), rounds=rounds_mid, final_faces=int(out_f.shape[0])),
               surface_distance=dict(optimal=surface_distance(verts, o_v, o_f),
                                     midpoint=surface_distance(verts, out_v, out_f),
                                     measure="input-vertex sample (20,000, seed 0) to output surface, "
                                             "/ bounding-box diagonal"))
    if not args.no_verify:
        from pamo_amd.verify import verify
        r = verify(o_v, o_f)
        out["verify"] = {k: r[k] for k in ("manifold", "watertight", "intersecting_pairs", "ok")}
    return out


def _stage1_leg(args, v, f, ctx):
    """Stage 1 (remeshing, pamo_amd.remesh) of the same input at the
    reference's lattice (R = 256, band 3/256): the signed distance and the
    extraction, timed apart, the shell checked closed and oriented; beside
    it the oracle's restatement on the host at R = 128 against the GPU at
    R = 128, outputs compared."""
    import numpy as np
    import torch
    from pamo_amd import remesh
    band = 3.0 / 256
    margin = 2 * band + 1
    tris, tmin, tmax, mean = remesh.normalize(v, f, band, margin)
    dt = torch.from_numpy(tris).to("cuda:0")
    vm = remesh.VolumeToMesh(ctx)

    def run(R, reps):
        ts, te = [], []
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            vol = remesh.get_sdf(dt, R, band, ctx=ctx)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            gv, gf = vm(vol - 0.9 / R)
            torch.cuda.synchronize()
            ts.append(t1 - t0)
            te.append(time.perf_counter() - t1)
        return float(np.mean(ts[1:])), float(np.mean(te[1:])), gv, gf

    s256, e256, gv, gf = run(256, max(args.steps, 1))
    out = dict(R=256, band=band, input_faces=int(len(f)), sdf_ms=round(1e3 * s256, 3), extract_ms=round(1e3 * e256, 3),
               value=round(len(f) / (s256 + e256) / 1e6, 3), unit="input Mfaces/s",
               shell_vertices=int(gv.shape[0]), shell_faces=int(gf.shape[0]))
    if not args.no_verify:
        from pamo_amd.verify import verify
        r = verify(gv.cpu().numpy(), gf.cpu().numpy())
        out["verify"] = {k: r[k] for k in ("manifold", "watertight", "misoriented_edges", "intersecting_pairs",
                                           "euler")}
    if not args.no_cpu:
        import oracle
        s128, e128, g128v, g128f = run(128, 1)
        t0 = time.perf_counter()
        vol = oracle.mesh_to_sdf(tris, 128, band)
        ov, of = oracle.volume_to_mesh(vol - np.float32(0.9 / 128))
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = dict(kind="port", cores=1, R=128, seconds=round(tc, 3),
                                   gpu_ms=round(1e3 * (s128 + e128), 3),
                                   value=round(len(f) / tc / 1e6, 4), unit="input Mfaces/s",
                                   outputs_equal=bool(np.array_equal(of, g128f.cpu().numpy()) and
                                                      np.array_equal(ov.view(np.uint32),
                                                                     g128v.cpu().numpy().view(np.uint32))),
                                   sample="the same mesh at R = 128 (the oracle is serial)")
    return out


def _stage3_leg(args, v, f, s2v, s2f, ctx):
    """Stage 3 (safe projection, pamo_amd.safe_project) as PaMO runs it after
    stage 2 (simp_cuda/pamo/__init__.py:160-176): 5 steps moving the stage-2
    output toward the input mesh, one system reused across calls.  Timed:
    registration (topology, rest state, the target's LBVH) and the steps;
    beside it the float64 oracle on the host (all cores) on the same call,
    outputs compared."""
    import numpy as np
    import torch
    from pamo_amd import safe_project as s3
    system = s3.Stage3System(ctx=ctx)
    s, t = s3.get_normalization_transform(v)
    gv = np.asarray(v, np.float64) * s + t
    mv = np.asarray(s2v, np.float64) * s + t
    dgv, dgf = torch.from_numpy(gv).cuda(), torch.from_numpy(np.asarray(f, np.int32)).cuda()
    dmv, dmf = torch.from_numpy(mv).cuda(), torch.from_numpy(np.asarray(s2f, np.int32)).cuda()
    reg, stp = [], []
    for _ in range(max(args.steps, 1) + max(args.warmup, 1)):
        system.clear()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        system.register_mesh(dgv, dgf, dmv, dmf)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        system.step(5)
        torch.cuda.synchronize()
        reg.append(t1 - t0)
        stp.append(time.perf_counter() - t1)
    out = (system.get_vertices() - t) / s
    sz = system.sizes()
    w = max(args.warmup, 1)
    reg, stp = reg[w:], stp[w:]
    sec = float(np.mean(reg)) + float(np.mean(stp))
    res = dict(steps=5, newton_per_step=10, target_faces=int(len(f)), mesh_vertices=int(sz["nV"]),
               mesh_faces=int(sz["nF"]), contacts_last=int(sz["contacts"]), dtype="f64",
               register_ms=round(1e3 * float(np.mean(reg)), 3), steps_ms=round(1e3 * float(np.mean(stp)), 3),
               ms=round(1e3 * sec, 3), value=round(len(s2f) / sec / 1e3, 3), unit="stage-2 kfaces/s")
    if not args.no_cpu:
        import oracle
        from oracle import stage3 as o3
        ncores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        oracle.set_threads(ncores)
        log("stage 3: oracle process on %d threads" % ncores)
        t0 = time.perf_counter()
        ov, _ = o3.process(v, f, s2v, s2f, 5)
        tc = time.perf_counter() - t0
        ext = float(np.max(np.max(v, 0) - np.min(v, 0)))
        res["cpu_baseline"] = dict(kind="port", cores=ncores, seconds=round(tc, 3),
                                   value=round(len(s2f) / tc / 1e3, 4), unit="stage-2 kfaces/s",
                                   max_err_over_extent=float(np.abs(out - ov).max() / ext),
                                   sample="the same call: 5 steps on the stage-2 output (the oracle finds "
                                          "nearest target points by scanning every target triangle)")
    return res


def run_stage3(args):
    """Stage 3 alone: the 2.0M-face SDF union through stage 2 (untimed), then
    the safe projection of its output toward the input, timed over --steps
    calls after --warmup (the stage3 leg of the default line)."""
    import torch
    from pamo_amd import _lib, preprocess_numpy, simplify_stage2
    v, f, cfg = make_mesh(args)
    verts, scale, tris_mean = preprocess_numpy(v, f)
    ctx = _lib.Context(0)
    dv, df = torch.from_numpy(verts).cuda(), torch.from_numpy(f).cuda()
    ov, of, st = simplify_stage2(dv, df, scale, int(cfg["ratio"] * len(f)), ctx=ctx)
    s2v, s2f = ov.cpu().numpy() + tris_mean, of.cpu().numpy()
    log("stage 2: %d -> %d faces" % (len(f), len(s2f)))
    del dv, df, ov, of
    leg = _stage3_leg(args, v, f, s2v, s2f, ctx)
    result = dict(metric="stage-3 safe projection of the stage-2 output (5 steps, PaMO's call)",
                  value=leg.pop("value"), unit=leg.pop("unit"), n_gpus=1, steps=args.steps, warmup=args.warmup,
                  ms_per_step=leg["ms"], higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
                  data="synthetic (seeded SDF mesh generated in-process)",
                  config=dict(workload="stage3_sdf_union_2M", input_faces=int(len(f)), stage2_faces=int(len(s2f))),
                  stage3=leg)
    if "cpu_baseline" in leg:
        result["cpu_baseline"] = leg.pop("cpu_baseline")
    print(json.dumps(result), flush=True)


def _plain_each(meshes, ctx, with_prep):
    """The plain path (pamo_run_stage2, one mesh at a time) on every mesh:
    per-mesh seconds and outputs.  with_prep: the timed region also holds
    the job's host preparation and upload (what a one-job-at-a-time server
    pays per job); otherwise the inputs are already on the device."""
    import torch
    from pamo_amd import batch, simplify_stage2
    dev = torch.device("cuda:0")
    out = []
    t_log = time.perf_counter()
    for n, (v, f, r) in enumerate(meshes):
        if not with_prep:
            dv, dt, sc, m = batch.prepare(v, f, dev)
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        if with_prep:
            dv, dt, sc, m = batch.prepare(v, f, dev)
        ov, of, st = simplify_stage2(dv, dt, sc, int(r * len(f)), ctx=ctx)
        torch.cuda.synchronize()
        out.append((time.perf_counter() - t0, ov, of, m))
        if time.perf_counter() - t_log > 10:
            t_log = time.perf_counter()
            log("plain path: %d of %d meshes" % (n + 1, len(meshes)))
    return out


def _batched_leg(args):
    """The many-mesh mix (512 meshes, log-uniform 2k-400k faces, three
    families, ratios 0.01/0.005/0.001, seeds 4000000+i) in one packed
    engine, inputs resident in HBM when timing starts; against the plain
    path on every mesh, one at a time, with the same resident inputs.  Every
    batched result must equal the plain one bit for bit."""
    import torch
    from pamo_amd import _lib, batch, synth
    meshes = _gen(synth.batch_specs(args.leg_meshes))
    F_total = sum(len(m[1]) for m in meshes)
    dev = torch.device("cuda:0")
    prepared = [batch.prepare(v, f, dev) for v, f, r in meshes]
    torch.cuda.synchronize()
    ctx = _lib.Context(0)
    eng = batch.BatchEngine(ctx=ctx)
    times, outs = [], {}
    for s in range(2):   # the first run sizes the engine's buffers
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        handles = {}
        for i, ((dv, dt, sc, m), (v, f, r)) in enumerate(zip(prepared, meshes)):
            handles[eng.submit(dv, dt, sc, int(r * len(f)), job=i)] = i
        pending = len(meshes)
        while pending:
            _, active = eng.step(1)
            for h, _ in eng.poll():
                outs[handles[h]] = eng.take(h)[:2]
                pending -= 1
            if active == 0 and pending:
                raise RuntimeError("engine idle with %d meshes pending" % pending)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log("batched leg run %d: %.1f ms" % (s, 1e3 * times[-1]))
    del eng, prepared
    plain = _plain_each(meshes, ctx, with_prep=False)
    t_plain = sum(p[0] for p in plain)
    same = sum(int(torch.equal(outs[i][0], plain[i][1]) and torch.equal(outs[i][1], plain[i][2]))
               for i in range(len(meshes)))
    sec = times[-1]
    return dict(value=round(F_total / sec / 1e6, 3), unit="Mfaces/s", meshes=len(meshes), faces=F_total,
                face_range=[int(min(len(m[1]) for m in meshes)), int(max(len(m[1]) for m in meshes))],
                seconds=round(sec, 4), workload="batch512: 512 meshes, families in thirds, seeds 4000000+i",
                plain=dict(value=round(F_total / t_plain / 1e6, 3), unit="Mfaces/s", seconds=round(t_plain, 3),
                           meshes=len(meshes)),
                speedup_over_plain=round(t_plain / sec, 2), results_equal_plain="%d of %d" % (same, len(meshes)))


def _served_leg(args):
    """A Poisson stream (arrival seed 5) of the serving workload's jobs
    (seeds 5000000+i) through the continuous-batching server, latency from
    arrival to result on the host.  The plain path serves the same stream
    one job at a time: every job's service time (host preparation, upload,
    pamo_run_stage2) is measured, and its latencies are those of a FIFO
    queue over the same arrivals with those times.  Every served result
    must equal the plain one bit for bit."""
    import numpy as np
    import torch
    from pamo_amd import _lib, batch, synth
    meshes = _gen(synth.batch_specs(args.leg_jobs, seed_base=5_000_000))
    F_total = sum(len(m[1]) for m in meshes)
    rate = args.leg_rate
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(meshes)))
    jobs = [dict(points=v, faces=f, ratio=r, arrival=float(t)) for (v, f, r), t in zip(meshes, arrivals)]
    ctx = _lib.Context(0)
    # warm-up: a few jobs through the plain path load every kernel
    _plain_each(meshes[:4], ctx, with_prep=True)
    got = {}
    srv = batch.Server(engine=batch.BatchEngine(ctx=ctx), rounds_per_step=args.serve_rounds)
    t0 = time.perf_counter()
    recs = srv.run(jobs, on_result=lambda i, v, f, info: got.__setitem__(i, (v, f)))
    log("served leg: %d jobs in %.1fs" % (len(jobs), time.perf_counter() - t0))
    lat = np.array([r["latency"] for r in recs])
    span = max(r["finished"] for r in recs) - float(arrivals[0])
    plain = _plain_each(meshes, ctx, with_prep=True)
    svc = np.array([p[0] for p in plain])
    done, seq_lat = 0.0, []
    for t_arr, s_i in zip(arrivals, svc):
        done = max(done, t_arr) + s_i
        seq_lat.append(done - t_arr)
    seq_lat = np.array(seq_lat)
    same = sum(int(got[i][0].tobytes() == (plain[i][1].cpu().numpy() + plain[i][3]).tobytes() and
                   got[i][1].tobytes() == plain[i][2].cpu().numpy().tobytes()) for i in range(len(meshes)))
    served = F_total / span / 1e6
    cap = F_total / float(svc.sum()) / 1e6
    return dict(value=round(served, 3), unit="Mfaces/s", rate_jobs_s=rate, jobs=len(jobs), faces=F_total,
                jobs_per_s=round(len(jobs) / span, 2), workload="serve: Poisson arrivals of the batch mix, "
                "seeds 5000000+i, arrival seed 5, HBM budget 200 GB",
                latency_ms=dict(p50=round(1e3 * float(np.percentile(lat, 50)), 2),
                                p99=round(1e3 * float(np.percentile(lat, 99)), 2),
                                max=round(1e3 * float(lat.max()), 2)),
                plain=dict(capacity=round(cap, 3), unit="Mfaces/s", capacity_jobs_s=round(len(jobs) / svc.sum(), 2),
                           latency_ms=dict(p50=round(1e3 * float(np.percentile(seq_lat, 50)), 2),
                                           p99=round(1e3 * float(np.percentile(seq_lat, 99)), 2)),
                           service="measured for every job (preparation, upload, pamo_run_stage2); FIFO queue over "
                                   "the same arrivals"),
                served_over_plain=round(served / cap, 2), results_equal_plain="%d of %d" % (same, len(meshes)),
                max_live_jobs=srv.stats.get("max_live"),
                device_gb_peak=round(srv.stats["device_bytes_peak"] / 1e9, 3),
                live_input_faces_peak=srv.stats["live_faces_peak"],
                device_bytes_per_live_input_face=round(srv.stats["device_bytes_peak"] /
                                                       max(srv.stats["live_faces_peak"], 1), 1),
                admission_bytes_per_face=batch.BYTES_PER_FACE)


def _cpu_full(verts, f, scale, target, round_faces, threads, out_v, out_f, placement, rounds_loop=False):
    """The oracle's whole loop on the host with the given threads, timed over
    the stage-2 window; its output against the GPU's, bit for bit.  With
    rounds_loop the loop is the reference's Python loop over one oracle
    forward per round (oracle.run_rounds, the same results), which logs its
    progress: a run of minutes reports as it goes."""
    import numpy as np
    import oracle
    threads = oracle.set_threads(threads)
    log("cpu oracle: %d threads, full run" % threads)
    t0 = time.perf_counter()
    if rounds_loop:
        last = [t0]

        def progress(r, n):
            if time.perf_counter() - last[0] > 15:
                last[0] = time.perf_counter()
                log("cpu oracle: round %d, %d faces, %.0fs" % (r, n, last[0] - t0))
        ov, of, st, orf = oracle.run_rounds(verts, f, scale, target, progress=progress)
    else:
        ov, of, st, orf = oracle.run(verts, f, scale, target, return_rounds=True,
                                     placement=1 if placement == "optimal" else 0)
    sec = time.perf_counter() - t0
    same = ov.tobytes() == np.ascontiguousarray(out_v).tobytes() and of.tobytes() == np.ascontiguousarray(out_f).tobytes()
    return dict(value=round(int(f.shape[0]) / sec / 1e6, 4), unit="Mfaces/s", cores=threads, kind="port",
                sample="the whole run (%d rounds) in %.1fs" % (st["rounds"], sec),
                rounds_match_gpu=[int(x) for x in orf] == round_faces, outputs_match_oracle=bool(same))


def _cpu_plain(verts, f, scale, target, round_faces, threads, seconds):
    """The oracle's loop on the host with the given threads, stopped after
    the first round that ends past the budget; the full-run time is the
    sample's time scaled by the fraction of face-rounds it covered."""
    import oracle
    threads = oracle.set_threads(threads)
    log("cpu oracle: %d threads, budget %.0fs" % (threads, seconds))
    F0 = int(f.shape[0])
    cst, crf, crs = oracle.run_timed(verts, f, scale, target, seconds)
    n = cst["rounds"]
    match = [int(x) for x in crf] == round_faces[:n]
    # work of a round is taken proportional to the faces entering it
    entering = [F0] + round_faces[:-1]
    frac = sum(entering[:n]) / float(sum(entering))
    est_total = crs[n - 1] / frac if n else float("nan")
    return dict(value=round(F0 / est_total / 1e6, 4), unit="Mfaces/s", cores=threads, kind="port",
                sample="oracle ran the first %d of %d rounds (%.1f%% of the face-rounds) in %.1fs; total time "
                       "scaled by face-rounds" % (n, len(round_faces), 100 * frac, crs[n - 1] if n else 0.0),
                rounds_match_gpu=bool(match))


def _class_profile(ctx, run):
    """One extra, untimed run with every kernel class timed by HIP events on
    the launch stream: achieved algorithmic GB/s per class (the traversal's
    bytes are not defined; it reports time only)."""
    from pamo_amd import _lib
    ks = [_lib.K_QUADRIC, _lib.K_EDGE_COST, _lib.K_CLAIM, _lib.K_LBVH, _lib.K_TRAVERSE]
    ctx.profile(ks)
    run()
    out = {}
    for k in ks:
        p = ctx.profile_read(k)
        n = max(p["launches"], 1)
        avg_s = p["ms"] / n / 1e3
        gbs = (p["bytes"] / n) / avg_s / 1e9 if avg_s > 0 and p["bytes"] > 0 else None
        out[_lib.KERNEL_NAMES[k]] = dict(launches=p["launches"], total_ms=round(p["ms"], 3),
                                         avg_us=round(1e6 * avg_s, 3),
                                         achieved_gbs=round(gbs, 2) if gbs else None,
                                         frac=round(gbs / HBM_PEAK_GBS, 4) if gbs else None)
    ctx.profile([])
    return out


# ------------------------------------------------------------ many meshes
def _gen(specs):
    from pamo_amd import synth
    t = time.time()
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    meshes = synth.batch_meshes(specs, workers)
    log("generated %d meshes, %.2fM faces (%.1fs, %d workers)" %
        (len(meshes), sum(m[1].shape[0] for m in meshes) / 1e6, time.time() - t, workers))
    return meshes


def _edge_cost_roofline(ctx):
    from pamo_amd import _lib
    prof = ctx.profile_read(_lib.K_EDGE_COST)
    avg_ms = prof["ms"] / max(prof["launches"], 1)
    avg_bytes = prof["bytes"] / max(prof["launches"], 1)
    achieved = (avg_bytes / (avg_ms / 1e3)) / 1e9 if avg_ms > 0 else 0.0
    return dict(kernel=_lib.KERNEL_NAMES[_lib.K_EDGE_COST], bound="hbm", achieved=round(achieved, 2),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                launches=prof["launches"], avg_launch_us=round(1e3 * avg_ms, 3), avg_algorithmic_bytes=round(avg_bytes))


def _plain_times(meshes, idx, ctx):
    """Plain path (pamo_run_stage2, one mesh at a time) on meshes[idx]:
    per-mesh seconds and outputs."""
    import torch
    from pamo_amd import batch, simplify_stage2
    dev = torch.device("cuda:0")
    out = {}
    t_log = time.perf_counter()
    for n, i in enumerate(idx):
        v, f, r = meshes[i]
        dv, dt, sc, m = batch.prepare(v, f, dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ov, of, st = simplify_stage2(dv, dt, sc, int(r * len(f)), ctx=ctx)
        torch.cuda.synchronize()
        out[i] = (time.perf_counter() - t0, ov, of)
        if time.perf_counter() - t_log > 10:
            t_log = time.perf_counter()
            log("plain path: %d of %d meshes" % (n + 1, len(idx)))
    return out


def _cpu_sample(meshes, idx, seconds):
    """The oracle on whole meshes of idx in order until the budget is used."""
    import oracle
    from pamo_amd import preprocess_numpy
    ncores = oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0)))
    faces = 0
    spent = 0.0
    n = 0
    for i in idx:
        v, f, r = meshes[i]
        verts, scale, _ = preprocess_numpy(v, f)
        t0 = time.perf_counter()
        oracle.run(verts, f, scale, int(r * len(f)))
        spent += time.perf_counter() - t0
        faces += len(f)
        n += 1
        if spent >= seconds:
            break
    return dict(value=round(faces / spent / 1e6, 4) if spent else None, unit="Mfaces/s", cores=ncores, kind="port",
                sample="oracle ran %d whole meshes (%d faces) of the plain-path sample in %.1fs" % (n, faces, spent))


def run_batch(args):
    import numpy as np
    import torch
    from pamo_amd import _lib, batch, synth
    specs = synth.batch_specs(args.meshes)
    meshes = _gen(specs)
    F_total = sum(len(m[1]) for m in meshes)
    dev = torch.device("cuda:0")
    prepared = [batch.prepare(v, f, dev) for v, f, r in meshes]
    torch.cuda.synchronize()
    ctx = _lib.Context(0)
    # warm-up: the plain path on a few small meshes loads every kernel
    sample = list(range(0, len(meshes), max(1, args.seq_sample)))
    _plain_times(meshes, sample[:2], ctx)
    steps = args.steps
    times = []
    outs = {}
    for s in range(steps):
        eng = batch.BatchEngine(ctx=ctx) if s == 0 else eng
        if s == steps - 1:
            ctx.profile([_lib.K_EDGE_COST])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        handles = {}
        for i, ((dv, dt, sc, m), (v, f, r)) in enumerate(zip(prepared, meshes)):
            handles[eng.submit(dv, dt, sc, int(r * len(f)), job=i)] = i
        pending = len(meshes)
        rounds = 0
        while pending:
            _, active = eng.step(1)
            rounds += 1
            for h, _ in eng.poll():
                gv, gf, info = eng.take(h)
                outs[handles[h]] = (gv, gf, info)
                pending -= 1
            if active == 0 and pending:
                raise RuntimeError("engine idle with %d meshes pending" % pending)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log("batch step %d: %.1f ms, %d engine rounds" % (s, 1e3 * times[-1], rounds))
    roofline = _edge_cost_roofline(ctx)
    ctx.profile([])
    sec = float(np.mean(times))
    # plain path, one mesh at a time, on every seq_sample-th mesh
    seq = _plain_times(meshes, sample, ctx)
    f_sample = sum(len(meshes[i][1]) for i in sample)
    t_sample = sum(seq[i][0] for i in sample)
    seq_rate = f_sample / t_sample / 1e6
    same = all(torch.equal(outs[i][0], seq[i][1]) and torch.equal(outs[i][1], seq[i][2]) for i in sample)
    result = dict(metric="input Mfaces/s, %d meshes batched (cfg4 mix)" % len(meshes), value=round(F_total / sec / 1e6, 3),
                  unit="Mfaces/s", n_gpus=1, steps=steps, warmup=1, ms_per_step=round(1e3 * sec, 3),
                  higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
                  data="synthetic (seeded generators, in-process)",
                  config=dict(workload="batch512", meshes=len(meshes), faces=F_total,
                              face_range=[int(min(len(m[1]) for m in meshes)), int(max(len(m[1]) for m in meshes))],
                              families="displaced icosphere / SDF union / noisy torus in thirds",
                              ratios=[0.01, 0.005, 0.001], seeds="4000000+i", tolerance=4, threshold=1e-3,
                              parallelism="single GPU, one packed engine"),
                  sequential_plain=dict(value=round(seq_rate, 3), unit="Mfaces/s", meshes=len(sample),
                                        faces=f_sample, seconds=round(t_sample, 3)),
                  batch_speedup=round((F_total / sec / 1e6) / seq_rate, 2), batch_equals_plain=bool(same),
                  roofline=roofline)
    if not args.no_cpu:
        result["cpu_baseline"] = _cpu_sample(meshes, sample, args.cpu_seconds)
    if not args.no_verify:
        from pamo_amd.verify import verify
        bad = 0
        for i in sample[:16]:
            r = verify(outs[i][0].cpu().numpy(), outs[i][1].cpu().numpy())
            bad += 0 if (r["manifold"] and r["watertight"] and r["intersecting_pairs"] == 0) else 1
        result["verify"] = dict(meshes_checked=min(16, len(sample)), failing=bad)
    print(json.dumps(result), flush=True)


def run_serve(args):
    """Poisson arrivals of the many-mesh mix into the continuous-batching
    server, one JSON line per arrival rate (--rate takes a comma list); the
    meshes and the plain-path service-time sample are shared by the rates."""
    import numpy as np
    from pamo_amd import _lib, batch, synth
    rates = [float(x) for x in str(args.rate).split(",") if x]
    specs = synth.batch_specs(args.jobs, seed_base=5_000_000)
    meshes = _gen(specs)
    ctx = _lib.Context(0)
    sample = list(range(0, len(meshes), max(1, args.seq_sample)))
    _plain_times(meshes, sample[:2], ctx)   # warm-up
    # plain path: measured service time of every seq_sample-th job, a linear
    # fit in input faces for the rest, and a FIFO single-server queue over the
    # same arrivals
    seq = _plain_times(meshes, sample, ctx)
    xs = np.array([len(meshes[i][1]) for i in sample], float)
    ys = np.array([seq[i][0] for i in sample])
    b, a = np.polyfit(xs, ys, 1)
    svc = np.array([max(a + b * len(m[1]), ys.min()) for m in meshes])
    for i in sample:
        svc[i] = seq[i][0]
    F_total = sum(len(m[1]) for m in meshes)
    # the plain path's capacity: the whole stream served one mesh at a time
    plain_mfs = F_total / float(svc.sum()) / 1e6
    plain_jps = len(meshes) / float(svc.sum())
    for rate in rates:
        rng = np.random.default_rng(5)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(meshes)))
        jobs = [dict(points=v, faces=f, ratio=r, arrival=float(t)) for (v, f, r), t in zip(meshes, arrivals)]
        eng = batch.BatchEngine(ctx=ctx)
        srv = batch.Server(engine=eng, rounds_per_step=args.serve_rounds)
        t0 = time.perf_counter()
        last = [t0]

        def progress(done, total):
            if time.perf_counter() - last[0] > 10:
                last[0] = time.perf_counter()
                log("serve %g jobs/s: %d of %d jobs done" % (rate, done, total))
        recs = srv.run(jobs, progress=progress)
        wall = time.perf_counter() - t0
        del eng
        lat = np.array([r["latency"] for r in recs])
        span = max(r["finished"] for r in recs) - float(arrivals[0])
        done = 0.0
        seq_lat = []
        for t_arr, s_i in zip(arrivals, svc):
            start = max(done, t_arr)
            done = start + s_i
            seq_lat.append(done - t_arr)
        seq_lat = np.array(seq_lat)
        log("serve %g jobs/s: p50 %.0f ms, p99 %.0f ms, %.1f jobs/s" %
            (rate, 1e3 * np.percentile(lat, 50), 1e3 * np.percentile(lat, 99), len(jobs) / span))
        result = dict(metric="served input Mfaces/s at %g jobs/s (Poisson)" % rate,
                      value=round(F_total / span / 1e6, 3), unit="Mfaces/s", n_gpus=1, steps=1, warmup=1,
                      ms_per_step=round(1e3 * wall, 3), higher_is_better=True, scaling="weak", vs_baseline=None,
                      dtype="f32", data="synthetic (seeded generators, in-process)",
                      config=dict(workload="serve", jobs=len(jobs), faces=F_total, rate_jobs_s=rate,
                                  arrival_seed=5, seeds="5000000+i", hbm_budget_gb=200,
                                  rounds_per_step=args.serve_rounds,
                                  parallelism="single GPU, continuous batching"),
                      latency_ms=dict(p50=round(1e3 * float(np.percentile(lat, 50)), 2),
                                      p99=round(1e3 * float(np.percentile(lat, 99)), 2),
                                      max=round(1e3 * float(lat.max()), 2)),
                      sequential_plain=dict(latency_ms=dict(p50=round(1e3 * float(np.percentile(seq_lat, 50)), 2),
                                                            p99=round(1e3 * float(np.percentile(seq_lat, 99)), 2)),
                                            capacity_mfaces_s=round(plain_mfs, 3),
                                            capacity_jobs_s=round(plain_jps, 2),
                                            measured_jobs=len(sample),
                                            model="measured service times on the sample, linear fit in faces "
                                                  "for the rest, FIFO queue over the same arrivals"),
                      jobs_per_s=round(len(jobs) / span, 2),
                      served_over_plain=round((F_total / span / 1e6) / plain_mfs, 2),
                      server_host_s={k: (round(v, 3) if isinstance(v, float) else v) for k, v in srv.stats.items()})
        print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
