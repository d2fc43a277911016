"""Pins the CPU oracle (oracle/pamo_oracle.cpp) and the exact verifier
(pamo_amd/csrc/verify.cpp) with checks that do not restate their formulas:
golden vectors from rocThrust, closed forms, exact rational arithmetic,
brute force on tiny inputs and invariants of the collapse."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from pamo_amd import synth
from pamo_amd import verify as pv

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------ shuffle (D7)
def test_shuffle_matches_rocthrust_golden():
    """thrust::shuffle(default_random_engine) permutations produced by
    rocThrust itself (tests/golden/make_shuffle_golden.py)."""
    gold = json.load(open(os.path.join(HERE, "golden", "shuffle_thrust.json")))
    assert len(gold) >= 10
    for m, perm in gold.items():
        p = oracle.shuffle_perm(int(m))
        assert list(p) == perm, m


@pytest.mark.parametrize("m", [1, 7, 16, 17, 1000, 65537])
def test_shuffle_is_bijection(m):
    p = oracle.shuffle_perm(m)
    assert sorted(p.tolist()) == list(range(m))


# ------------------------------------------------------ exact predicates
def _orient_frac(a, b, c, d):
    a, b, c, d = [[Fraction(float(x)) for x in p] for p in (a, b, c, d)]
    m = [[a[i] - d[i] for i in range(3)], [b[i] - d[i] for i in range(3)], [c[i] - d[i] for i in range(3)]]
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    return (det > 0) - (det < 0)


def test_orient3d_exact_vs_rationals():
    rng = np.random.default_rng(11)
    n = 0
    for _ in range(400):
        a, b, c = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        # d nearly on the plane of abc: a float32 rounding of an exact combination
        s, t = rng.uniform(-1, 1, 2)
        d = (a + s * (b - a) + t * (c - a)).astype(np.float32)
        d = d + rng.integers(-2, 3, 3).astype(np.float32) * np.spacing(np.abs(d).astype(np.float32))
        for pts in ((a, b, c, d), (a, b, c, rng.uniform(-1, 1, 3).astype(np.float32))):
            want = _orient_frac(*pts)
            assert pv.orient3d(*[p.astype(np.float64) for p in pts]) == want
            assert pv.orient3d(*[p.astype(np.float64) for p in pts], exact_only=True) == want
            n += want == 0
    # exact zeros: integer-grid points on a common plane
    for _ in range(100):
        a, b, c = rng.integers(-50, 50, (3, 3))
        d = a + 3 * (b - a) - 2 * (c - a)
        pts = [np.asarray(p, np.float64) * 2.0 ** -20 for p in (a, b, c, d)]
        assert pv.orient3d(*pts) == 0 == _orient_frac(*pts)


def _tri(rng):
    return rng.uniform(0, 1, (3, 3)).astype(np.float32)


def test_reference_tritri_agrees_with_exact_on_generic_pairs():
    """Away from the epsilon filters (no shared vertex, not near-coplanar,
    no contact within 1e-3), the reference interval test and the exact test
    agree."""
    rng = np.random.default_rng(5)
    agree = checked = 0
    for _ in range(3000):
        a, b = _tri(rng), _tri(rng) * 0.8 + 0.1
        # skip near-degenerate configurations (orientation margins)
        na = np.cross(a[1] - a[0], a[2] - a[0])
        nb = np.cross(b[1] - b[0], b[2] - b[0])
        da = np.abs((b - a[0]) @ na) / np.linalg.norm(na)
        db = np.abs((a - b[0]) @ nb) / np.linalg.norm(nb)
        if da.min() < 1e-3 or db.min() < 1e-3 or np.linalg.norm(np.cross(na, nb)) < 1e-3 * np.linalg.norm(na) * np.linalg.norm(nb):
            continue
        hit, r1, r2 = oracle.isect_raw(a, b)
        exact = pv.tri_pair_intersects(a, b, [0, 1, 2], [3, 4, 5])
        if hit and np.abs(r2 - r1).max() < 1e-3:
            continue
        checked += 1
        agree += hit == exact
    assert checked > 1000
    assert agree == checked


def test_narrow_phase_filters():
    """Outcome codes of self_intersect.cuh:284-375 on constructed pairs."""
    t = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    # coplanar overlapping pair -> skipped as coplanar (code 2)
    u = np.array([[0.2, 0.2, 0], [1.2, 0.2, 0], [0.2, 1.2, 0]], np.float32)
    assert oracle.narrow(t, u, [0, 1, 2], [3, 4, 5]) == 2
    # pair sharing an edge by index, folded through -> code 3
    w = np.array([[1, 0, 0], [0, 1, 0], [0.3, 0.3, 0.5]], np.float32)
    assert oracle.narrow(t, w, [0, 1, 2], [1, 2, 6]) == 3
    # the same geometry with distinct indices -> shared-edge coordinates (code 4)
    assert oracle.narrow(t, w, [0, 1, 2], [7, 8, 6]) == 4
    # a crossing pair -> hit
    x = np.array([[0.2, 0.2, -0.5], [0.3, 0.3, 0.5], [0.9, -0.5, 0.2]], np.float32)
    assert oracle.narrow(t, x, [0, 1, 2], [3, 4, 5]) == 1
    assert pv.tri_pair_intersects(t, x, [0, 1, 2], [3, 4, 5])
    # far apart -> miss
    assert oracle.narrow(t, x + 5, [0, 1, 2], [3, 4, 5]) == 0
    # shared vertex with a short contact -> filtered (code 5)
    y = np.array([[0, 0, 0], [0.0004, 0.0004, 0.5], [0.0004, 0.0004, -0.5]], np.float32)
    assert oracle.narrow(t, y, [0, 1, 2], [0, 4, 5]) == 5
    assert oracle.is_coplanar(t, u) and not oracle.is_coplanar(t, x)


# ----------------------------------------------------------- quadrics
def _octahedron():
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]], np.int32)
    return v, f


def test_quadrics_closed_form():
    """Each face plane (n, -n.v0) of the octahedron is (s/sqrt3, -1/sqrt3)
    with s in {+-1}^3; Q_v = sum over the 4 faces at v of p p^T."""
    v, f = _octahedron()
    o = oracle.Oracle(shuffle=oracle.SHUFFLE_IDENTITY)
    o.forward(v, f, np.zeros(0, np.int32), 2.0, 1e-3, False, True)
    Q = o.get("Q").reshape(-1, 4, 4).astype(np.float64)
    for vi in range(6):
        want = np.zeros((4, 4))
        for face in f:
            if vi in face:
                a, b, c = v[face].astype(np.float64)
                n = np.cross(b - a, c - a)
                n /= np.linalg.norm(n)
                p = np.append(n, -n @ a)
                want += np.outer(p, p)
        assert np.allclose(Q[vi], want, atol=1e-6)
        assert np.allclose(Q[vi], Q[vi].T, atol=0)            # exactly symmetric
        assert abs(np.append(v[vi], 1) @ Q[vi] @ np.append(v[vi], 1)) < 1e-6


def test_edge_cost_closed_form_octahedron():
    """Edge (x+, y+) of the unit octahedron (identity shuffle): midpoint
    m = (1/2, 1/2, 0).  QEM: the 8 fan planes give (n.m + d)^2 = 0 twice for
    the two shared faces in each fan and 1/3 for the other two, so
    v^T(Q_u+Q_v)v = 4/3 and the cost term is (4/3)/s^2 with s = 2.  Length
    terms: every edge is sqrt2, so L = (sqrt2/s + sqrt2/s) tau.  Skinny
    terms: the moved faces are (m, y-, z+-) and (m, x-, z+-); with the
    reference precedence 4*sqrt3*A/e01^2 + e12^2 + e20^2 > 1 for all four,
    so the penalty is 0."""
    v, f = _octahedron()
    tau = 1e-3
    o = oracle.Oracle(shuffle=oracle.SHUFFLE_IDENTITY)
    o.forward(v, f, np.zeros(0, np.int32), 2.0, tau, False, True)
    edges = o.get("edges")
    cost = o.get("edge_cost")
    idx = [i for i, (a, b) in enumerate(edges) if (a, b) == (0, 2)]
    assert len(idx) == 1
    c = (4.0 / 3.0) / 4.0 + (np.sqrt(2) / 2 + np.sqrt(2) / 2) * tau
    want = c / 10.0 * 2.0 ** 32
    assert abs(int(cost[idx[0]]) - want) < 2e-6 * want
    # by symmetry all 12 edges cost the same
    assert len(set(cost.tolist())) == 1


def test_unguarded_edge_cost_closed_form_octahedron():
    """The unguarded cost (cusimp.cu:283-482) of the same octahedron edge: the
    fan length term is counted twice (edge_len += edge_len + ..., :378) and the
    skinny term carries no x5 weight (it is 0 here anyway), so
    c = (4/3)/s^2 + (2 sqrt2/s + sqrt2/s) tau with s = 2."""
    v, f = _octahedron()
    tau = 1e-3
    o = oracle.Oracle(shuffle=oracle.SHUFFLE_IDENTITY)
    o.forward_sp(v, f, 2.0, tau, True)
    edges = o.get("edges")
    cost = o.get("edge_cost")
    idx = [i for i, (a, b) in enumerate(edges) if (a, b) == (0, 2)]
    assert len(idx) == 1
    c = (4.0 / 3.0) / 4.0 + (2 * np.sqrt(2) / 2 + np.sqrt(2) / 2) * tau
    want = c / 10.0 * 2.0 ** 32
    assert abs(int(cost[idx[0]]) - want) < 2e-6 * want
    assert len(set(cost.tolist())) == 1


def test_unguarded_rounds_keep_topology():
    """Rounds of the unguarded variant (no guard, no skinny restore, no
    invalid-face removal): winners collapse independently, each removes its
    two faces, no face gets a repeated corner, and the compacted mesh stays
    closed and manifold with the same Euler characteristic; the rounds end
    when one removes no face (PaSP.run, simp_cuda/pamo/__init__.py:196-222)."""
    v, f = synth.displaced_icosphere(3, 9)
    verts = np.asarray(v, np.float32)
    scale = float(np.max(verts.max(0) - verts.min(0)))
    chi0 = pv.verify(verts, f)["euler"]
    o = oracle.Oracle()
    rounds = 0
    for r in range(200):
        n0 = f.shape[0]
        P, T, occ, vmap = o.forward_sp(verts, f, scale, 1e-3, r == 0)
        rounds += 1
        alive = T[:, 0] >= 0
        assert T.shape[0] - alive.sum() == 2 * int((occ == 0).sum())
        At = T[alive]
        assert not np.any((At[:, 0] == At[:, 1]) | (At[:, 1] == At[:, 2]) | (At[:, 2] == At[:, 0]))
        verts = P[occ.ravel() == 1]
        f = vmap.ravel()[At].astype(np.int32)
        res = pv.verify(verts, f)
        assert res["manifold"] and res["watertight"] and res["euler"] == chi0, res
        if f.shape[0] == n0:
            break
    assert 1 < rounds < 200


def test_cost_quantization_edges():
    """collapse_t for tau=1e-3 is 429496 (float32 arithmetic of
    cusimp_free.cu:923)."""
    o = oracle.Oracle()
    v, f = _octahedron()
    o.forward(v, f, np.zeros(0, np.int32), 2.0, 1e-3, False, True)
    assert o.sizes()["collapse_t"] == 429496
    o.forward(v, f, np.zeros(0, np.int32), 2.0, 10.0, False, True)
    assert o.sizes()["collapse_t"] == 0xFFFFFFFF
    o.forward(v, f, np.zeros(0, np.int32), 2.0, 0.0, False, True)
    assert o.sizes()["collapse_t"] == 0


def test_flat_grid_has_zero_qem():
    """On a planar patch every quadric annihilates every in-plane point, so
    interior edge costs come only from the length and skinny terms."""
    n = 8
    xs, ys = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v = np.stack([xs.ravel(), ys.ravel(), np.zeros(n * n)], 1).astype(np.float32)
    f = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1
            f += [[a, b, c], [a, c, d]]
    f = np.array(f, np.int32)
    o = oracle.Oracle(shuffle=oracle.SHUFFLE_IDENTITY)
    o.forward(v, f, np.zeros(0, np.int32), float(n - 1), 1e-3, False, True)
    Q = o.get("Q").reshape(-1, 4, 4)
    for vi in range(v.shape[0]):
        if Q[vi].any():
            p = np.append(v[vi] + np.array([0.3, -0.7, 0], np.float32), 1).astype(np.float64)
            assert abs(p @ Q[vi] @ p) < 1e-9


# ------------------------------------------------------ brute force / sets
def test_candidate_pairs_brute_force():
    v, f = synth.displaced_icosphere(2, 4)
    cand, hits = oracle.full_check(v, f)
    lo = v[f].min(1)
    hi = v[f].max(1)
    F = f.shape[0]
    ov = np.ones((F, F), bool)
    for k in range(3):
        ov &= ~((hi[:, None, k] < lo[None, :, k]) | (hi[None, :, k] < lo[:, None, k]))
    np.fill_diagonal(ov, False)
    want = np.argwhere(ov)
    assert np.array_equal(cand, want)
    assert hits.shape[0] == 0   # a star-shaped surface has no intersections


def _fans(tris, V):
    fans = [[] for _ in range(V)]
    for fi, t in enumerate(tris):
        if t[0] == -1 and t[1] == -1 and t[2] == -1:
            continue
        for x in t:
            fans[x].append(fi)
    return fans


def test_claims_and_independent_set_brute_force():
    v, f = synth.displaced_icosphere(3, 2)
    verts, scale, _ = oracle.preprocess(v, f)
    o = oracle.Oracle()
    o.forward(verts, f, np.zeros(0, np.int32), scale, 1e-3, False, True)
    T0 = o.get("orig_tris")
    edges = o.get("edges")
    cost = o.get("edge_cost").astype(np.uint64)
    fans = _fans(T0, verts.shape[0])
    # fan order is ascending face index (D0)
    off = o.get("fan_off")
    fan = o.get("fan")
    for x in range(verts.shape[0]):
        assert fan[off[x]:off[x + 1]].tolist() == fans[x]
    keys = (cost << np.uint64(32)) | np.arange(len(edges), dtype=np.uint64)
    want = np.full(T0.shape[0], np.iinfo(np.uint64).max, np.uint64)
    for e, (a, b) in enumerate(edges):
        for fi in fans[a] + fans[b]:
            want[fi] = min(want[fi], keys[e])
    assert np.array_equal(o.get("tri_min"), want)
    coll = o.get("collapsed")
    owned = {}
    for e in coll:
        a, b = edges[e]
        for fi in set(fans[a] + fans[b]):
            assert fi not in owned          # winners own disjoint fans
            owned[fi] = e
        assert cost[e] <= o.sizes()["collapse_t"]


# ---------------------------------------------------- invariants per round
@pytest.mark.parametrize("mesh", ["ico", "torus"])
def test_round_invariants(mesh):
    """Each round keeps the surface closed and manifold, keeps the Euler
    characteristic, removes exactly two faces per net collapse, and restored
    fans equal the round-start faces bit for bit."""
    if mesh == "ico":
        v, f = synth.displaced_icosphere(4, 3)
    else:
        v, f = synth.noisy_torus(40, 16, 2)
    verts, scale, _ = oracle.preprocess(v, f)
    chi0 = pv.verify(verts, f)["euler"]
    o = oracle.Oracle()
    undo = np.zeros(0, np.int32)
    for r in range(12):
        P, T, occ, vmap, undo = o.forward(verts, f, undo, scale, 1e-3, False, r == 0)
        T0 = o.get("orig_tris")
        alive = T[:, 0] >= 0
        removed_v = int((occ == 0).sum())
        assert T.shape[0] - alive.sum() == 2 * removed_v
        changed = np.any(T != T0, axis=1)
        assert np.all(changed <= (~alive | np.isin(T, np.nonzero(occ.ravel() == 1)[0]).all(1)))
        verts = P[occ.ravel() == 1]
        f = vmap.ravel()[T[alive]].astype(np.int32)
        res = pv.verify(verts, f)
        assert res["manifold"] and res["watertight"], res
        assert res["euler"] == chi0


def _zero_length_edge(level=2):
    """Unit icosphere whose first face's first edge has length zero (its
    larger vertex moved onto the smaller)."""
    v, f = synth.icosphere(level)
    v = np.asarray(v, np.float32).copy()
    f = np.asarray(f, np.int32)
    a, b = sorted(int(x) for x in f[0, :2])
    v[b] = v[a]
    return v, f, a, b


def test_line_edge_collapse():
    """cusimp_free.cu:234-274 on a closed sphere with one zero-length edge
    (a, b), a < b, and a threshold that admits no regular collapse: the
    edge is the only collapsed one, b dies, the two faces on (a, b) are
    removed, every other face of b now uses a, and the compacted result is
    closed, manifold, has the same Euler characteristic and no alive face
    with two coincident corners."""
    v, f, a, b = _zero_length_edge()
    chi0 = pv.verify(v, f)["euler"]
    verts, scale, _ = oracle.preprocess(v, f)
    assert np.array_equal(verts[a], verts[b])
    o = oracle.Oracle(shuffle=oracle.SHUFFLE_IDENTITY)
    P, T, occ, vmap, undo = o.forward(verts, f, np.zeros(0, np.int32), scale, 0.0, False, True)
    sz = o.sizes()
    assert sz["n_winners"] == 0 and sz["n_collapsed"] == 1
    E = o.get("edges")
    assert sorted(E[o.get("collapsed")[0]].tolist()) == [a, b]
    assert occ[b, 0] == 0 and np.all(P[b] == 0) and int(occ.sum()) == v.shape[0] - 1
    both = np.isin(f, [a]).any(1) & np.isin(f, [b]).any(1)
    assert both.sum() == 2 and np.all(T[both] == -1)
    alive = T[:, 0] >= 0
    assert alive.sum() == f.shape[0] - 2 and not np.isin(T[alive], [b]).any()
    had_b = np.isin(f, [b]).any(1) & ~both
    assert np.all(np.where(f[had_b] == b, a, f[had_b]) == T[had_b])
    assert np.all(T[~had_b & ~both] == f[~had_b & ~both])
    nv = P[occ.ravel() == 1]
    nf = vmap.ravel()[T[alive]].astype(np.int32)
    res = pv.verify(nv, nf)
    assert res["manifold"] and res["watertight"] and res["euler"] == chi0, res
    corners = nv[nf]
    assert not np.any(np.all(corners[:, 0] == corners[:, 1], 1) | np.all(corners[:, 1] == corners[:, 2], 1) |
                      np.all(corners[:, 2] == corners[:, 0], 1))


def test_run_reaches_target_or_tolerance():
    v, f = synth.displaced_icosphere(4, 8)
    verts, scale, _ = oracle.preprocess(v, f)
    target = int(0.05 * f.shape[0])
    ov, of, st, rf = oracle.run(verts, f, scale, target, return_rounds=True)
    assert st["exit_reason"] in (1, 3)
    if st["exit_reason"] == 1:
        assert of.shape[0] <= target
    else:
        assert rf[-1] == rf[-2] == rf[-3] == rf[-4] == rf[-5]
    assert np.all(np.diff(rf) <= 0)
    assert pv.verify(ov, of)["ok"]


@pytest.mark.parametrize("target_ratio", [0.05, 0.0])
def test_python_round_loop_equals_c_loop(target_ratio):
    """run_rounds (the reference's loop in Python over one forward per round,
    reporting progress) gives the C loop's output bit for bit, round by
    round, to the target and to the tolerance stop."""
    v, f = synth.displaced_icosphere(4, 8)
    verts, scale, _ = oracle.preprocess(v, f)
    target = int(target_ratio * f.shape[0])
    ov, of, st, rf = oracle.run(verts, f, scale, target, return_rounds=True)
    seen = []
    pv_, pf_, pst, prf = oracle.run_rounds(verts, f, scale, target, progress=lambda r, n: seen.append((r, n)))
    assert pv_.tobytes() == ov.tobytes() and pf_.tobytes() == of.tobytes()
    assert list(prf) == list(rf) and [n for _, n in seen] == list(rf)
    assert pst["rounds"] == st["rounds"] and pst["exit_reason"] == st["exit_reason"]


@pytest.mark.parametrize("tol", [1, 4, 7])
def test_run_without_collapses_stops_after_tolerance_rounds(tol):
    """Threshold 0 admits no collapse on a sphere (every quantized cost is
    above 0): the face count never changes, so PaMO.run's counter reaches
    `tolerance` after exactly `tolerance` rounds (simp_cuda/pamo/__init__.py:
    147-158) and the mesh comes back unchanged (identity shuffle)."""
    v, f = synth.displaced_icosphere(3, 4)
    verts, scale, _ = oracle.preprocess(v, f)
    ov, of, st, rf = oracle.run(verts, f, scale, 0, threshold=0.0, tolerance=tol, shuffle=oracle.SHUFFLE_IDENTITY,
                                return_rounds=True)
    assert st["rounds"] == tol and st["exit_reason"] == 3
    assert list(rf) == [f.shape[0]] * tol
    assert np.array_equal(of, f) and np.array_equal(ov, verts)


def test_undo_loop_on_intersecting_input():
    """Two interpenetrating spheres: collapses touching the intersection are
    undone, up to 5 passes (cusimp_free.cu:1024-1040)."""
    v1, f1 = synth.displaced_icosphere(3, 3)
    v2, f2 = synth.displaced_icosphere(3, 4)
    v = np.concatenate([v1, v2 + np.array([0.9, 0.05, 0], np.float32)]).astype(np.float32)
    f = np.concatenate([f1, f2 + v1.shape[0]]).astype(np.int32)
    verts, scale, _ = oracle.preprocess(v, f)
    o = oracle.Oracle()
    o.forward(verts, f, np.zeros(0, np.int32), scale, 1e-3, False, True)
    s = o.sizes()
    assert s["undo_passes"] >= 1 and s["n_vertices_undo"] > 0
    und = o.get("undo")
    assert und.shape[0] == s["n_vertices_undo"]
    assert np.all(und == 0)          # first round: the stale map is zero-filled (D1)


def test_undo_leaves_no_flagged_fan():
    """A round whose undo loop ends before its cap (SDF union, round 76):
    the final geometry has no kept collapse of the scanned prefix whose
    round-start fans hold a face of an intersecting pair (brute force over
    all ordered pairs of the final mesh), every undone edge has its far
    vertex back at its round-start position, and with one pass the undo
    list holds exactly the undone edges' two vertices."""
    v, f = synth.sdf_union(60, 100, 5)
    verts, scale, _ = oracle.preprocess(v, f)
    o = oracle.Oracle()
    undo = np.zeros(0, np.int32)
    for r in range(76):
        P0 = verts
        P, T, occ, vmap, undo = o.forward(verts, f, undo, scale, 1e-3, False, r == 0)
        if r < 75:
            alive = T[:, 0] >= 0
            verts = P[occ.ravel() == 1]
            f = vmap.ravel()[T[alive]].astype(np.int32)
    s = o.sizes()
    assert s["undo_passes"] == 1, s
    E, coll = o.get("edges"), o.get("collapsed")
    off, fan = o.get("fan_off"), o.get("fan")
    occ = occ.ravel()
    scanned = coll[:s["undo_cut"]]
    kept = [e for e in scanned if occ[E[e][1]] == 0]
    undone = [e for e in scanned if occ[E[e][1]] == 1]
    assert len(undone) > 0 and 2 * len(undone) == s["n_vertices_undo"]
    for e in undone:
        assert np.array_equal(P[E[e][1]], P0[E[e][1]])
    _, hits = oracle.full_check(P, T)
    flagged = np.zeros(T.shape[0], bool)
    flagged[hits.ravel()] = True
    for e in kept:
        a, b = E[e]
        fs = np.concatenate([fan[off[a]:off[a + 1]], fan[off[b]:off[b + 1]]])
        assert not flagged[fs].any(), e


# ------------------------------------------------------------- verifier
def test_verifier_detects_defects():
    v, f = synth.displaced_icosphere(2, 1)
    assert pv.verify(v, f)["ok"]
    r = pv.verify(v, f[1:])
    assert r["boundary_edges"] == 3 and not r["watertight"]
    g = f.copy()
    g[0] = g[0][[0, 2, 1]]
    assert pv.verify(v, g)["misoriented_edges"] > 0
    w = v.copy()
    w[0] = -w[0] * 1.5           # pull a vertex out through the opposite side
    assert pv.verify(w, f)["intersecting_pairs"] > 0


def test_sdf_generator_is_closed_manifold():
    v, f = synth.sdf_union(40, 80, seed=9)
    r = pv.verify(v, f)
    assert r["ok"], r
    v, f = synth.sdf_union(30, 100, seed=5, n_slab_pairs=4)
    r = pv.verify(v, f)
    assert r["manifold"] and r["watertight"] and r["degenerate_faces"] == 0


# ------------------------------------------------------- FP64 precision
def test_cost_twin_float_equals_oracle_and_fp64_bounds():
    """The float instance of the cost twin quantizes to the oracle's own
    first-round costs, and its winners (the claim rule in numpy) are the
    oracle's; the double instance gives the FP64 reference values.  The
    collapse positions (midpoints) agree to one float rounding."""
    v, f = synth.displaced_icosphere(4, 7)
    verts, scale, _ = oracle.preprocess(v, f)
    tw = oracle.cost_twin(verts, f, scale)
    o = oracle.Oracle(oracle.SHUFFLE_IDENTITY)
    o.forward(verts, f.astype(np.int32), np.zeros(0, np.int32), float(scale), 1e-3, False, True)
    ec = o.get("edge_cost")[:len(tw["q32"])]
    sel = tw["st32"] != 2
    assert sel.any() and np.array_equal(tw["q32"][sel], ec[sel])
    r = oracle.precision_report(verts, f, scale)
    assert r["winners_fp32"] == o.sizes()["n_winners"]
    assert r["status_differs"] == 0
    assert r["position_rel_err_max"] <= 2.0 ** -23
    # FP32 evaluation of the reference's expressions stays near FP64 for
    # most edges (the QEM term cancels near the surface, hence the tail)
    assert r["cost_rel_err_median"] < 0.05


# ------------------------------------------------------- strict guard
def test_strict_guard_removes_what_the_reference_keeps():
    """A thin-slab union (32,436 faces): the reference's guard skips
    coplanar pairs, shared-edge coordinates and short shared-vertex
    contacts, and its output keeps exact intersections.  The strict guard
    (exact pair test, every collapse scanned, undo until clean) leaves none,
    and on a mesh where nothing is skipped it changes nothing."""
    v, f = synth.sdf_union(20, 80, 3, n_slab_pairs=6)
    assert pv.verify(v, f)["ok"]
    verts, scale, _ = oracle.preprocess(v, f)
    target = int(0.001 * len(f))
    ov, of, st = oracle.run(verts, f, scale, target)
    r = pv.verify(ov, of)
    assert r["manifold"] and r["watertight"] and r["intersecting_pairs"] > 0
    sv, sf, sst = oracle.run(verts, f, scale, target, strict=True)
    rs = pv.verify(sv, sf)
    assert rs["ok"], rs
    assert sst["max_undo_passes"] < 64
    # a star-shaped surface: the strict and the reference guard agree
    v, f = synth.displaced_icosphere(4, 5)
    verts, scale, _ = oracle.preprocess(v, f)
    a = oracle.run(verts, f, scale, int(0.01 * len(f)))
    b = oracle.run(verts, f, scale, int(0.01 * len(f)), strict=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ------------------------------------------- cost and collapse branches
# Hand-built meshes (tests/cost_eval.py) that each decide one branch, so a
# misreading shared by the oracle and the device would fail here.
import warnings  # noqa: E402

import cost_eval as ce  # noqa: E402


def _identity_round(P, F, s, tau, undo=None, sp=False):
    o = oracle.Oracle(shuffle=oracle.SHUFFLE_IDENTITY)
    if sp:
        o.forward_sp(P, F, s, tau, True)
    elif undo is None:
        o.forward(P, F, np.zeros(0, np.int32), s, tau, False, True)
    else:
        # a first call on the same vertices with every face removed sizes the
        # invalid-vertex table, so the second call's undo list marks it (the
        # table is written on the storage as it stands, cusimp_free.cu:895-916)
        o.forward(P, np.full_like(F, -1), np.zeros(0, np.int32), s, tau, False, True)
        o.forward(P, F, undo, s, tau, False, False)
    return o


def _cost_of(o, u, v):
    E = o.get("edges")
    idx = [i for i, (a, b) in enumerate(E) if (a, b) == (u, v)]
    assert len(idx) == 1
    return int(o.get("edge_cost")[idx[0]])


def test_cost_restatement_matches_oracle_every_edge():
    """A second restatement of the edge cost (numpy float32 scalars, written
    separately from the C++ oracle) gives the oracle's quantized cost on every
    edge whose cost the reference writes, guarded and unguarded."""
    meshes = [ce.flat_patch(), ce.flat_patch(lift=ce.FLIP_LIFT), synth.displaced_icosphere(2, 5),
              synth.noisy_torus(12, 8, 3)]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")   # 0/0 of degenerate normals gives NaN, as in the reference
        for P, F in meshes:
            verts, scale, _ = oracle.preprocess(P, F)
            for sp in (False, True):
                o = _identity_round(verts, F, scale, 1e-3, sp=sp)
                E, cost = o.get("edges"), o.get("edge_cost")
                n = 0
                for e, (a, b) in enumerate(E):
                    want = ce.edge_cost(verts, F, int(a), int(b), scale, 1e-3, unguarded=sp)
                    if want is None:
                        continue
                    n += 1
                    assert int(cost[e]) == want, (F.shape, sp, a, b)
                assert n >= 20


def test_skinny_penalty_closed_form():
    """Edge (u, v) of the flat patch, s = 1, tau = 1.  Both fans are flat, so
    the QEM term is exactly 0.  Of the 8 moved faces (the fan faces without
    the other endpoint: num_tri = 8) only (m, p, q) has a skinny term below 1:
    with k = 1/8, m = (k/2, 0), p = (-k/2, k/10), q = (-k, 0), its area is
    0.075 k^2, |m-p|^2 = 1.01 k^2, |p-q|^2 = 0.26 k^2, |q-m|^2 = 2.25 k^2, so
    t = 4 sqrt3 * 0.075 / 1.01 + 2.51 k^2 + 1e-7 = 0.55369 (the reference's
    precedence: only the area is divided by e01^2).  The other seven terms
    are 1.76 to 3.51 and clamp to 1.  The cost is the length term plus
    5 (1 - t) / 8 * tau (cusimp_free.cu:573-584); the penalty is about half
    of it, so a lost x5 weight or /num_tri, or a wrong num_tri, moves the
    cost far outside float rounding."""
    k = 0.125
    P, F = ce.flat_patch()
    o = _identity_round(P, F, 1.0, 1.0)
    got = _cost_of(o, 0, 1)
    P64 = P.astype(np.float64)
    el, ne = 0.0, 0
    for f in [0, 1, 2, 3, 4, 5, 0, 5, 6, 7, 8, 9]:   # fan of u, then fan of v (ascending face ids)
        i, j, kk = F[f]
        for a, b in ((i, j), (j, kk), (kk, i)):
            if a > b:
                el += np.linalg.norm(P64[a] - P64[b])
                ne += 1
    L = el / ne + np.linalg.norm(P64[0] - P64[1])
    t = 4 * np.sqrt(3) * 0.075 / 1.01 + (0.26 + 2.25) * k * k + 1e-7
    penalty = 5.0 * (1.0 - t) / 8.0
    want = (L + penalty) / 10.0 * 2.0 ** 32
    assert abs(t - 0.55369) < 1e-5
    assert penalty > 0.4 * (L + penalty)
    assert abs(got - want) <= 1e-5 * want, (got, want)


def test_normal_flip_sets_max_cost():
    """A closed manifold mesh without intersections (verified) in which the
    collapse of (u, v) to its midpoint turns the moved face (u, p, q) over:
    the cost is UINT32_MAX (cusimp_free.cu:526-531,566-571).  The same
    restatement without the flip test gives a finite cost, so the flip is
    what decides it."""
    P, F = ce.flat_patch(lift=ce.FLIP_LIFT)
    assert pv.verify(P, F)["ok"]
    o = _identity_round(P, F, 1.0, 1e-3)
    assert _cost_of(o, 0, 1) == 0xFFFFFFFF
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert ce.edge_cost(P, F, 0, 1, 1.0, 1e-3) == 0xFFFFFFFF
        assert ce.edge_cost(P, F, 0, 1, 1.0, 1e-3, flip_check=False) < 0xFFFFFFFF


@pytest.mark.parametrize("case", range(len(ce.RESTORE_CASES)))
def test_skinny_restore_threshold_formed_in_double(case):
    """The winning collapse of (u, v) leaves the face (m, p, q) with a float
    area exactly at, or one ulp beside, the restore threshold
    1e-6 * max_len^2 (cusimp_free.cu:617-637).  1e-6 is a double literal, so
    the threshold is formed in double and rounded once; at these scales a
    float-formed threshold is one ulp away, and the face at the lower of the
    two would be judged the other way.  The collapse is restored (v occupied
    again, u back at its position) exactly when area < the double-formed
    threshold."""
    k, y, restored = ce.RESTORE_CASES[case]
    P, F = ce.restore_mesh(k, y)
    assert pv.verify(P, F)["ok"]
    m = (P[0] + P[1]) / np.float32(2)
    area, thr_d, thr_f = ce.skinny_thresholds(m, P[3], P[4])
    assert thr_d != thr_f
    assert (area < thr_d) == restored
    if area in (thr_d, thr_f):
        assert (area < thr_f) != restored     # a float-formed threshold decides the other way
    o = _identity_round(P, F, 1.0, 1e-3, undo=ce.RESTORE_UNDO)
    E = o.get("edges")
    assert [tuple(int(c) for c in E[e]) for e in o.get("collapsed")] == [(0, 1)]
    occ, pts = o.get("occ"), o.get("points")
    assert int(occ[1]) == (1 if restored else 0)
    assert pts[0].tobytes() == (P[0] if restored else m).tobytes()


def test_unguarded_quotient_in_double():
    """The unguarded variant writes sqrt(3) with an integer argument
    (cusimp.cu:423,463), the double overload, so its skinny quotient is
    formed in double and rounded to float once (D10).  At this geometry the
    double and a float evaluation give quantized costs 16 apart; the oracle
    gives the double one."""
    P, F = ce.flat_patch(p=ce.D10_P)
    o = _identity_round(P, F, 1.0, 1.0, sp=True)
    got = _cost_of(o, 0, 1)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        want = ce.edge_cost(P, F, 0, 1, 1.0, 1.0, unguarded=True, d10=True)
        alt = ce.edge_cost(P, F, 0, 1, 1.0, 1.0, unguarded=True, d10=False)
    assert want != alt
    assert got == want


# ------------------------------------------- optimal placement (opt-in)
def _quadric(planes):
    """10 unique entries of sum p p^T over planes p = (n, d) (n unit)."""
    Q = np.zeros((4, 4))
    for p in planes:
        p = np.asarray(p, np.float64)
        Q += np.outer(p, p)
    return np.array([Q[0, 0], Q[0, 1], Q[0, 2], Q[0, 3], Q[1, 1], Q[1, 2], Q[1, 3], Q[2, 2], Q[2, 3], Q[3, 3]],
                    np.float32)


def test_placement_closed_forms():
    """Each branch of the placement on a quadric with a known minimiser.
    Three orthogonal planes through p0: A = I, so the FP32 solve returns p0
    exactly.  The same with the third plane weighted 2^-7 (entries exact in
    float; det / tr^3 = 7.6e-6, below the FP32 bound): the FP64 solve,
    again p0.  A flat patch
    (every plane z = 0): A is singular, the rule takes the midpoint (all
    three errors are 0).  A crease (planes x = 0 and y = 0): singular again,
    the rule takes the endpoint on the crease.  An optimum farther from the
    midpoint than the edge is long: rejected, the rule."""
    p0 = np.array([0.25, -0.5, 0.75], np.float32)
    a, b = p0 + np.float32(0.125), p0 - np.float32(0.0625)
    q = _quadric([(1, 0, 0, -0.25), (0, 1, 0, 0.5), (0, 0, 1, -0.75)])
    x, how = oracle.place(q, a, b)
    assert how == 1 and x.tobytes() == p0.tobytes()
    q = _quadric([(1, 0, 0, -0.25), (0, 1, 0, 0.5), (0, 0, 2.0 ** -7, -0.75 * 2.0 ** -7)])
    x, how = oracle.place(q, a, b)
    assert how == 2 and x.tobytes() == p0.tobytes()
    q = _quadric([(0, 0, 1, 0)] * 6)
    a2, b2 = np.array([0, 0, 0], np.float32), np.array([1, 0.5, 0], np.float32)
    x, how = oracle.place(q, a2, b2)
    assert how == 3 and x.tobytes() == ((a2 + b2) / np.float32(2)).tobytes()
    q = _quadric([(1, 0, 0, 0), (0, 1, 0, 0)])
    on, off = np.array([0, 0, 0.3], np.float32), np.array([0.2, 0.1, 0.3], np.float32)
    x, how = oracle.place(q, off, on)
    assert how == 5 and x.tobytes() == on.tobytes()
    x, how = oracle.place(q, on, off)
    assert how == 4 and x.tobytes() == on.tobytes()
    far = np.array([10, 10, 10], np.float32)
    q = _quadric([(1, 0, 0, -10), (0, 1, 0, -10), (0, 0, 1, -10)])
    x, how = oracle.place(q, np.zeros(3, np.float32), np.array([0.1, 0, 0], np.float32))
    assert how in (3, 4, 5) and np.linalg.norm(x - far) > 1


def _surface_error(v0, f0, v1, f1):
    """Mean and max distance from the input vertices to the output surface
    (brute force over the output faces; small meshes)."""
    tri = v1[f1].astype(np.float64)
    p = v0.astype(np.float64)
    best = np.full(len(p), np.inf)
    for t in tri:
        a, b, c = t
        ab, ac, ap = b - a, c - a, p - a
        d1, d2 = ap @ ab, ap @ ac
        # closest point by barycentric projection clamped to the triangle (Ericson 2005, 5.1.5)
        n = np.cross(ab, ac)
        nn = n @ n
        if nn == 0:
            continue
        w = (p - a) - np.outer((p - a) @ n / nn, n)
        M = np.array([[ab @ ab, ab @ ac], [ab @ ac, ac @ ac]])
        st = np.linalg.solve(M, np.stack([w @ ab, w @ ac]))
        s, t2 = st
        inside = (s >= 0) & (t2 >= 0) & (s + t2 <= 1)
        dist = np.where(inside, np.abs((p - a) @ n) / np.sqrt(nn), np.inf)
        for u, v in ((a, b), (b, c), (c, a)):
            e = v - u
            h = np.clip(((p - u) @ e) / (e @ e), 0, 1)
            dist = np.minimum(dist, np.linalg.norm(p - (u + h[:, None] * e), axis=1))
        best = np.minimum(best, dist)
    return best.mean(), best.max()


def test_optimal_placement_runs_stay_valid_and_lower_error():
    """Whole runs with the optimal placement: outputs stay closed, manifold
    and free of intersections, and sit closer to the input surface than the
    midpoint runs at the same target."""
    for v, f in (synth.displaced_icosphere(4, 5), synth.sdf_union(10, 40, 7)):
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(0.02 * f.shape[0])
        res = {}
        for pl in (0, 1):
            ov, of, st = oracle.run(verts, f, scale, target, placement=pl)
            r = pv.verify(ov, of)
            assert r["manifold"] and r["watertight"] and r["intersecting_pairs"] == 0, (pl, r)
            res[pl] = (_surface_error(verts, f, ov, of), of.shape[0])
        (m0, _), n0 = res[0]
        (m1, _), n1 = res[1]
        print("midpoint", res[0], "optimal", res[1])
        assert m1 < m0


def test_fp64_shadow_follows_the_fp32_run():
    """The whole-run FP64 shadow runs the same FP32 loop as oracle.run
    (rounds, final size), scores every edge of every round, and its double
    positions follow the FP32 collapses: on a mesh with dyadic coordinates
    every midpoint is exact in float, so the position difference is exactly
    0; on a generic mesh it stays at float rounding."""
    v, f = synth.displaced_icosphere(4, 5)
    verts, scale, _ = oracle.preprocess(v, f)
    target = int(0.01 * f.shape[0])
    _, of, st = oracle.run(verts, f, scale, target)
    r = oracle.fp64_shadow(verts, f, scale, target)
    assert r["rounds"] == st["rounds"] and r["final_faces"] == of.shape[0]
    assert r["winners_fp32"] == st["total_collapsed"]
    assert r["edges_scored"] > 0 and 0 < r["edges_valued"] <= r["edges_scored"]
    assert 0 < r["position_err_max"] < 1e-6
    P, F = ce.flat_patch(p=(-0.5, 0.125))
    assert np.all(P * 1024 == np.round(P * 1024))   # dyadic coordinates
    r = oracle.fp64_shadow(P, F, 1.0, 4, threshold=1e-3)
    assert r["winners_fp32"] > 0 and r["position_err_max"] == 0.0
