"""Batched and served stage 2 over the engine in libpamo_hip.so.

``BatchEngine`` wraps one context's engine: meshes are submitted (appended to
the packed device state), ``step`` advances every active mesh by whole
rounds in shared launches, finished meshes are retired between rounds and
their results taken.  Each mesh follows exactly the round loop of
``PaMO.run`` (simp_cuda/pamo/__init__.py:121-158), so a mesh's result is the
same alone, in a batch, or admitted mid-stream.

``simplify_batch`` runs a list of meshes to completion.  ``Server`` is a
continuous-batching request loop: jobs arriving at given times are admitted
while the HBM budget allows, rounds run for everything on the device, and
finished jobs are replaced by waiting ones between rounds.
"""
import ctypes
import time

import numpy as np
import torch

from concurrent.futures import ThreadPoolExecutor

from . import _lib, prepare_host

# device bytes per admitted input face of the packed state (positions,
# faces and their round-start copies, fans, edges, costs, claims, LBVH
# nodes, scan scratch, growth headroom).  Serving the 1,000-job mix of 1k to
# 400k-face meshes, the context's peak device bytes (pamo_device_bytes) were
# 375 per admitted input face at the peak (12.3 GB for 32.8M faces), so the
# admission estimate stays above what the engine allocates.
BYTES_PER_FACE = 420


class BatchEngine:
    def __init__(self, device=None, shuffle=_lib.SHUFFLE_FEISTEL, ctx=None, strict=None, placement=None):
        if ctx is None:
            ctx = _lib.Context(torch.cuda.current_device() if device is None else device, shuffle)
        self.ctx = ctx
        if strict is not None:
            ctx.set_strict(strict)
        if placement is not None:
            ctx.set_placement(placement)
        L = ctx._L
        P, I, LL, Fl = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
        L.pamo_batch_submit.argtypes = [P, P, P, P, I, I, Fl, LL, Fl, I, LL, LL]
        L.pamo_batch_submit.restype = I
        L.pamo_batch_step.argtypes = [P, P, I, ctypes.POINTER(I)]
        L.pamo_batch_step.restype = I
        L.pamo_batch_poll.argtypes = [P, P, P, I]
        L.pamo_batch_poll.restype = I
        L.pamo_batch_result.argtypes = [P, I, ctypes.POINTER(I), ctypes.POINTER(I), P, P, I]
        L.pamo_batch_result.restype = I
        L.pamo_batch_take.argtypes = [P, P, I, P, P]
        L.pamo_batch_take.restype = I
        L.pamo_batch_state.argtypes = [P, P]
        L.pamo_batch_state.restype = I
        self.L = L

    def _stream(self):
        return torch.cuda.current_stream().cuda_stream

    def submit(self, verts, faces, scale, target, threshold=1e-3, tolerance=4, iters=1000000, job=-1):
        """verts: centred float32 CUDA tensor [V,3]; faces int32 [F,3]."""
        verts = verts.contiguous()
        faces = faces.contiguous()
        h = self.L.pamo_batch_submit(self.ctx.handle, self._stream(), verts.data_ptr(), faces.data_ptr(),
                                     int(verts.shape[0]), int(faces.shape[0]), float(scale), int(target),
                                     float(threshold), int(tolerance), int(iters), int(job))
        return self.ctx.check(h, "pamo_batch_submit")

    def step(self, rounds=1):
        na = ctypes.c_int()
        fin = self.L.pamo_batch_step(self.ctx.handle, self._stream(), int(rounds), ctypes.byref(na))
        self.ctx.check(fin, "pamo_batch_step")
        return fin, na.value

    def poll(self):
        n = self.L.pamo_batch_poll(self.ctx.handle, None, None, 0)
        h = np.zeros(max(n, 1), np.int32)
        j = np.zeros(max(n, 1), np.int64)
        self.L.pamo_batch_poll(self.ctx.handle, h.ctypes.data, j.ctypes.data, n)
        return list(zip(h[:n].tolist(), j[:n].tolist()))

    def result_info(self, handle):
        nv, nf = ctypes.c_int(), ctypes.c_int()
        st = np.zeros(8, np.int64)
        rf = np.zeros(4096, np.int32)
        rc = self.L.pamo_batch_result(self.ctx.handle, int(handle), ctypes.byref(nv), ctypes.byref(nf),
                                      st.ctypes.data, rf.ctypes.data, rf.size)
        self.ctx.check(rc, "pamo_batch_result")
        if rc == 1:
            return None
        rounds = int(st[0])
        return dict(V=nv.value, F=nf.value, rounds=rounds, exit_reason=int(st[1]), collapses=int(st[2]),
                    max_undo_passes=int(st[3]), job=int(st[4]), V0=int(st[5]), F0=int(st[6]),
                    seconds=float(st[7]) * 1e-6, round_faces=rf[:min(rounds, rf.size)].copy())

    def take(self, handle, device=None):
        info = self.result_info(handle)
        if info is None:
            raise RuntimeError("job %d has not finished" % handle)
        dev = device or torch.device("cuda", self.ctx.device)
        v = torch.empty((info["V"], 3), dtype=torch.float32, device=dev)
        f = torch.empty((info["F"], 3), dtype=torch.int32, device=dev)
        self.ctx.check(self.L.pamo_batch_take(self.ctx.handle, self._stream(), int(handle), v.data_ptr(),
                                              f.data_ptr()), "pamo_batch_take")
        return v, f, info

    def state(self):
        out = np.zeros(6, np.int64)
        self.L.pamo_batch_state(self.ctx.handle, out.ctypes.data)
        return dict(active=int(out[0]), V=int(out[1]), F=int(out[2]), checks=int(out[3]), query_faces=int(out[4]),
                    candidate_pairs=int(out[5]))


def _prepare_pinned(points, faces):
    """prepare_host with the vertices and faces copied into pinned host
    memory (new tensors from torch's pinned allocator), so that their upload
    is an asynchronous copy on the stream."""
    hv, ht, scale, mean = prepare_host(points, faces)
    return torch.from_numpy(hv).pin_memory(), torch.from_numpy(ht).pin_memory(), scale, mean


def prepare(points, faces, device):
    """prepare_host, with the vertices and faces moved to the device."""
    verts, tri, scale, tris_mean = prepare_host(points, faces)
    return torch.from_numpy(verts).to(device), torch.from_numpy(tri).to(device), scale, tris_mean


def simplify_batch(meshes, ratios, threshold=1e-3, tolerance=4, min_verts=0, engine=None, rounds_per_step=1):
    """Simplify many meshes together.  meshes: list of (points, faces) host
    arrays; ratios: one ratio or one per mesh.  Returns a list of
    (verts, faces, info) with host numpy arrays in the input order."""
    dev = torch.device("cuda", torch.cuda.current_device())
    eng = engine or BatchEngine()
    if np.isscalar(ratios):
        ratios = [ratios] * len(meshes)
    means = {}
    for i, ((p, f), r) in enumerate(zip(meshes, ratios)):
        v, t, s, m = prepare(p, f, dev)
        target = max(int(r * len(f)), min_verts)
        h = eng.submit(v, t, s, target, threshold, tolerance, job=i)
        means[h] = (i, m)
    out = [None] * len(meshes)
    pending = len(meshes)
    while pending:
        _, active = eng.step(rounds_per_step)
        for h, job in eng.poll():
            v, f, info = eng.take(h)
            i, m = means[h]
            out[i] = (v.cpu().numpy() + m, f.cpu().numpy(), info)
            pending -= 1
        if active == 0 and pending:
            raise RuntimeError("engine idle with %d jobs pending" % pending)
    return out


class Server:
    """Continuous batching: jobs (points, faces, ratio, arrival time) are
    admitted when they have arrived and the packed state stays under the
    HBM budget; every loop iteration runs rounds_per_step rounds for all
    admitted jobs (the engine queues each round behind the previous one on
    the device; jobs are admitted and results taken between iterations).
    Latency = completion time - arrival time (both on the host clock)."""

    def __init__(self, budget_bytes=200 << 30, threshold=1e-3, tolerance=4, min_verts=0, engine=None,
                 prep_threads=4, rounds_per_step=1):
        self.eng = engine or BatchEngine()
        self.rounds_per_step = max(1, int(rounds_per_step))
        self.budget = budget_bytes
        self.threshold = threshold
        self.tolerance = tolerance
        self.min_verts = min_verts
        self.prep_threads = prep_threads
        self.stats = {}

    def run(self, jobs, realtime=True, on_result=None, progress=None):
        """jobs: list of dicts {points, faces, ratio, arrival} (arrival in
        seconds from the start).  on_result(i, verts, faces, info) receives
        each result as host arrays; progress(done, total) is called after
        each completion.  Returns per-job records."""
        dev = torch.device("cuda", torch.cuda.current_device())
        order = sorted(range(len(jobs)), key=lambda i: jobs[i]["arrival"])
        waiting = []
        nxt = 0
        live = {}
        used = 0
        recs = [None] * len(jobs)
        # host preparation (centring and scale, then a copy into pinned
        # memory) runs on worker threads from the job's arrival on,
        # overlapping the rounds already on the device; the upload at
        # admission is then an asynchronous copy
        pool = ThreadPoolExecutor(max_workers=max(1, self.prep_threads))
        prepared = {}
        # host seconds per part of the loop, rounds run, peak admitted jobs,
        # peak device bytes of the engine's context and peak admitted input
        # faces (the measure behind BYTES_PER_FACE)
        st = dict(step_s=0.0, submit_s=0.0, prep_wait_s=0.0, take_s=0.0, idle_s=0.0, rounds=0, max_live=0,
                  device_bytes_peak=0, live_faces_peak=0)
        live_faces = 0
        self.stats = st
        t0 = time.perf_counter()
        done = 0
        try:
            while done < len(jobs):
                now = time.perf_counter() - t0
                while nxt < len(order) and (jobs[order[nxt]]["arrival"] <= now or not realtime):
                    i = order[nxt]
                    waiting.append(i)
                    prepared[i] = pool.submit(_prepare_pinned, jobs[i]["points"], jobs[i]["faces"])
                    nxt += 1
                while waiting:
                    i = waiting[0]
                    need = BYTES_PER_FACE * len(jobs[i]["faces"])
                    if live and used + need > self.budget:
                        break
                    # admit in arrival order once prepared; while the device
                    # has work, rounds go on instead of waiting for the host
                    if live and not prepared[i].done():
                        break
                    waiting.pop(0)
                    ta = time.perf_counter()
                    hv, ht, s, m = prepared.pop(i).result()
                    tb = time.perf_counter()
                    target = max(int(jobs[i]["ratio"] * len(jobs[i]["faces"])), self.min_verts)
                    h = self.eng.submit(hv.to(dev, non_blocking=True), ht.to(dev, non_blocking=True), s, target,
                                        self.threshold, self.tolerance, job=i)
                    tc = time.perf_counter()
                    st["prep_wait_s"] += tb - ta
                    st["submit_s"] += tc - tb
                    live[h] = (i, m, need, tc - t0)
                    used += need
                    live_faces += len(jobs[i]["faces"])
                    st["live_faces_peak"] = max(st["live_faces_peak"], live_faces)
                if not live:
                    if nxt < len(order):
                        ta = time.perf_counter()
                        time.sleep(max(0.0, min(0.001, jobs[order[nxt]]["arrival"] - now)))
                        st["idle_s"] += time.perf_counter() - ta
                    continue
                st["max_live"] = max(st["max_live"], len(live))
                ta = time.perf_counter()
                self.eng.step(self.rounds_per_step)
                tb = time.perf_counter()
                st["step_s"] += tb - ta
                st["rounds"] += self.rounds_per_step
                st["device_bytes_peak"] = max(st["device_bytes_peak"], self.eng.ctx.device_bytes())
                for h, _ in self.eng.poll():
                    ta = time.perf_counter()
                    v, f, info = self.eng.take(h)
                    st["take_s"] += time.perf_counter() - ta
                    i, m, need, t_adm = live.pop(h)
                    used -= need
                    live_faces -= len(jobs[i]["faces"])
                    t_end = time.perf_counter() - t0
                    recs[i] = dict(job=i, arrival=jobs[i]["arrival"], admitted=t_adm, finished=t_end,
                                   latency=t_end - jobs[i]["arrival"], faces_in=len(jobs[i]["faces"]),
                                   faces_out=info["F"], rounds=info["rounds"])
                    if on_result is not None:
                        on_result(i, v.cpu().numpy() + m, f.cpu().numpy(), info)
                    done += 1
                    if progress is not None:
                        progress(done, len(jobs))
        finally:
            pool.shutdown(wait=False, cancel_futures=True)
        return recs
