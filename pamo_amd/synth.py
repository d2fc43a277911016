"""Seeded synthetic meshes for benchmarks and tests.

Every generator draws its random parameters from ``numpy.random.default_rng``
so a (generator, seed, size) triple names one mesh exactly.  Positions are
float32 and faces int32, the layout ``PaMO.run`` takes.

* ``displaced_icosphere``: icosphere subdivided L times, radius
  r(p) = 1 + sum_i a_i exp(-|p - c_i|^2 / (2 w_i^2)) over 16 Gaussian bumps
  (c_i uniform on the sphere, a_i ~ U(-0.15, 0.15), w_i ~ U(0.1, 0.4)).
* ``noisy_torus``: torus R = 1, r ~ U(0.2, 0.5) with Gaussian bumps on the
  tube radius.
* ``sdf_union``: marching cubes of a smooth-min union of spheres (C++ in
  csrc/synth_mc.cpp, built into libpamo_synth.so).
"""
import ctypes
import os

import numpy as np

_ICO_V = None


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(level):
    """Unit icosphere with 20*4^level faces (outward orientation)."""
    v, f = _icosahedron()
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        nf = f.shape[0]
        a = base + inv[:nf]
        b = base + inv[nf:2 * nf]
        c = base + inv[2 * nf:]
        v = np.concatenate([v, mid], axis=0)
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)], axis=0)
    return v, f


def _bumps(rng, n=16, amp=0.15, wlo=0.1, whi=0.4):
    c = rng.normal(size=(n, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    a = rng.uniform(-amp, amp, n)
    w = rng.uniform(wlo, whi, n)
    return c, a, w


def displaced_icosphere(level=5, seed=1):
    """Star-shaped (hence intersection-free) bumpy sphere; level 5 gives
    10,242 vertices and 20,480 faces."""
    rng = np.random.default_rng(seed)
    c, a, w = _bumps(rng)
    v, f = icosphere(level)
    d2 = ((v[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    r = 1.0 + (a[None, :] * np.exp(-d2 / (2.0 * w[None, :] ** 2))).sum(1)
    assert r.min() > 0.05
    return (v * r[:, None]).astype(np.float32), f.astype(np.int32)


def noisy_torus(nu, nv, seed):
    """Torus R = 1, tube radius r ~ U(0.2, 0.5) displaced by 16 Gaussian
    bumps over the (u, v) angles; 2*nu*nv faces."""
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(0.2, 0.5)
    cu = rng.uniform(0, 2 * np.pi, 16)
    cv = rng.uniform(0, 2 * np.pi, 16)
    a = rng.uniform(-0.3, 0.3, 16) * r0
    w = rng.uniform(0.2, 0.6, 16)
    u = np.arange(nu) * (2 * np.pi / nu)
    vv = np.arange(nv) * (2 * np.pi / nv)
    U, Vv = np.meshgrid(u, vv, indexing="ij")
    du = np.angle(np.exp(1j * (U[..., None] - cu)))
    dv = np.angle(np.exp(1j * (Vv[..., None] - cv)))
    r = r0 + (a * np.exp(-(du ** 2 + dv ** 2) / (2 * w ** 2))).sum(-1)
    r = np.maximum(r, 0.05)
    x = (1.0 + r * np.cos(Vv)) * np.cos(U)
    y = (1.0 + r * np.cos(Vv)) * np.sin(U)
    z = r * np.sin(Vv)
    pts = np.stack([x, y, z], -1).reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    i1 = (i + 1) % nu
    j1 = (j + 1) % nv
    p00 = (i * nv + j).reshape(-1)
    p10 = (i1 * nv + j).reshape(-1)
    p01 = (i * nv + j1).reshape(-1)
    p11 = (i1 * nv + j1).reshape(-1)
    f = np.concatenate([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)], 0)
    return pts.astype(np.float32), f.astype(np.int32)


def uv_sphere(n_lon, n_lat, seed, noise=0.02):
    """Latitude-longitude sphere with radial noise: two poles of degree
    n_lon (long fans), 2 * n_lon * (n_lat - 1) faces, outward orientation."""
    rng = np.random.default_rng(seed)
    th = np.arange(1, n_lat) * (np.pi / n_lat)
    ph = np.arange(n_lon) * (2 * np.pi / n_lon)
    ring = np.stack([np.sin(th)[:, None] * np.cos(ph)[None], np.sin(th)[:, None] * np.sin(ph)[None],
                     np.cos(th)[:, None] * np.ones(n_lon)[None]], -1).reshape(-1, 3)
    v = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]])
    v = v * (1.0 + noise * rng.standard_normal((v.shape[0], 1)))
    nr = n_lat - 1
    south = 1 + nr * n_lon
    j = np.arange(n_lon)
    j1 = (j + 1) % n_lon
    faces = [np.stack([np.zeros(n_lon, np.int64), 1 + j, 1 + j1], 1)]
    for i in range(nr - 1):
        a, b = 1 + i * n_lon + j, 1 + i * n_lon + j1
        c, d = 1 + (i + 1) * n_lon + j1, 1 + (i + 1) * n_lon + j
        faces += [np.stack([a, d, c], 1), np.stack([a, c, b], 1)]
    last = 1 + (nr - 1) * n_lon
    faces.append(np.stack([np.full(n_lon, south), last + j1, last + j], 1))
    return v.astype(np.float32), np.concatenate(faces).astype(np.int32)


# ------------------------------------------------------------- SDF union
_HERE = os.path.dirname(os.path.abspath(__file__))
_SYN = os.path.join(_HERE, "libpamo_synth.so")
_syn = None


def _synlib():
    global _syn
    if _syn is None:
        if not os.path.exists(_SYN):
            raise RuntimeError("libpamo_synth.so is not built; run __graft_entry__.build()")
        _syn = ctypes.CDLL(_SYN)
        P = ctypes.c_void_p
        _syn.synth_sdf_mesh.argtypes = [P, ctypes.c_int, P, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, P, P]
        _syn.synth_sdf_mesh.restype = ctypes.c_longlong
        _syn.synth_fetch.argtypes = [P, P]
    return _syn


def sdf_mesh(spheres, slabs, k, n, threads=0):
    """Marching cubes (face-consistent, manifold) of the zero set of
    smin-union(spheres) union slabs on an n^3 grid over [0, 1]^3.

    spheres: (S, 4) float64 rows (cx, cy, cz, r); slabs: (B, 13) rows
    (centre xyz, unit normal xyz, unit tangent xyz, half-thickness,
    half-extent along tangent, half-extent along normal x tangent, unused).
    """
    lib = _synlib()
    spheres = np.ascontiguousarray(spheres, np.float64)
    slabs = np.ascontiguousarray(slabs, np.float64).reshape(-1, 13)
    counts = np.zeros(2, np.int64)
    handle = lib.synth_sdf_mesh(spheres.ctypes.data_as(ctypes.c_void_p), spheres.shape[0],
                                slabs.ctypes.data_as(ctypes.c_void_p) if slabs.size else None, slabs.shape[0],
                                float(k), int(n), int(threads), counts.ctypes.data_as(ctypes.c_void_p), None)
    if handle < 0:
        raise RuntimeError("sdf mesh generation failed (%d)" % handle)
    V, F = int(counts[0]), int(counts[1])
    pts = np.empty((V, 3), np.float32)
    faces = np.empty((F, 3), np.int32)
    lib.synth_fetch(pts.ctypes.data_as(ctypes.c_void_p), faces.ctypes.data_as(ctypes.c_void_p))
    return pts, faces


def sdf_union(n_spheres=200, grid=700, seed=2, k=0.02, n_slab_pairs=0, threads=0):
    """Smooth-min union (k) of n_spheres spheres with centres U[0.1,0.9]^3
    and radii U(0.03, 0.12), optionally with thin slab pairs, meshed on a
    grid^3 lattice."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.1, 0.9, size=(n_spheres, 3))
    r = rng.uniform(0.03, 0.12, size=n_spheres)
    spheres = np.concatenate([c, r[:, None]], 1)
    slabs = np.zeros((0, 13))
    if n_slab_pairs:
        rows = []
        h = 1.0 / grid
        for _ in range(n_slab_pairs):
            ctr = rng.uniform(0.15, 0.85, 3)
            nrm = rng.normal(size=3)
            nrm /= np.linalg.norm(nrm)
            tg = np.cross(nrm, rng.normal(size=3))
            tg /= np.linalg.norm(tg)
            th = rng.uniform(1e-3, 5e-3)
            gap = rng.uniform(0.6, 1.4) * h
            ext_t = rng.uniform(0.04, 0.12)
            ext_b = rng.uniform(0.04, 0.12)
            off = 0.5 * (th + gap)
            for s in (-1.0, 1.0):
                rows.append(np.concatenate([ctr + s * off * nrm, nrm, tg, [0.5 * th, ext_t, ext_b, 0.0]]))
        slabs = np.array(rows)
    return sdf_mesh(spheres, slabs, k, grid, threads)


def mesh_for_faces(kind, target_faces, seed, threads=0):
    """One mesh of roughly target_faces faces from a generator family
    (threads: OpenMP threads of the SDF mesher, 0 = all)."""
    if kind == "icosphere":
        level = int(np.clip(np.round(np.log(max(target_faces, 20) / 20.0) / np.log(4.0)), 1, 8))
        return displaced_icosphere(level, seed)
    if kind == "torus":
        nv = max(8, int(np.sqrt(target_faces / 2.0 / 3.0)))
        nu = max(8, int(target_faces / 2.0 / nv))
        return noisy_torus(nu, nv, seed)
    if kind == "sdf":
        rng = np.random.default_rng(seed)
        ns = int(rng.integers(5, 40))
        # faces grow roughly with the square of the grid resolution; a union
        # of few small spheres falls short of that estimate, so the lattice
        # is refined (faces ~ grid^2) until the mesh has at least 0.7 x target
        grid = int(np.clip(np.sqrt(target_faces / (2.2 * max(ns, 1) ** 0.5)) * 1.0, 24, 700))
        v, f = sdf_union(ns, grid, seed, threads=threads)
        for _ in range(4):
            if f.shape[0] >= 0.7 * target_faces or grid >= 700:
                break
            grid = int(min(700, np.ceil(grid * np.sqrt(target_faces / max(f.shape[0], 1)))))
            v, f = sdf_union(ns, grid, seed, threads=threads)
        return v, f
    raise ValueError(kind)


def batch_specs(n=512, seed_base=4_000_000, fmin=2_000, fmax=400_000, ratios=(0.01, 0.005, 0.001)):
    """The many-mesh workload: n meshes with face counts log-uniform in
    [fmin, fmax], generator families mixed in thirds (displaced icosphere at
    the matching subdivision level, SDF union, noisy torus) and ratios drawn
    uniformly from `ratios`; mesh i is seeded seed_base + i.  Returns a list
    of (kind, target_faces, ratio, seed)."""
    kinds = ("icosphere", "sdf", "torus")
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed_base + i)
        faces = int(np.exp(rng.uniform(np.log(fmin), np.log(fmax))))
        ratio = float(ratios[int(rng.integers(len(ratios)))])
        out.append((kinds[i % 3], faces, ratio, seed_base + i))
    return out


def _spec_mesh(spec, threads=0):
    kind, faces, ratio, seed = spec
    v, f = mesh_for_faces(kind, faces, seed, threads)
    return v, f, ratio


def _spec_mesh_1t(spec):
    return _spec_mesh(spec, 1)   # pool workers: one thread each


def batch_meshes(specs, workers=1):
    """Generate the meshes of batch_specs: a list of (points, faces, ratio)
    in spec order; `workers` > 1 generates them in a process pool."""
    if workers <= 1 or len(specs) < 2:
        return [_spec_mesh(s) for s in specs]
    import concurrent.futures as cf
    import multiprocessing as mp
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_spec_mesh_1t, specs, chunksize=4))
