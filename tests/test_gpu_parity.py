"""GPU parity: libpamo_hip.so against the CPU oracle, bit for bit.

Every comparison is exact (integer, index and float bit patterns).  The
oracle pins D0-D9 (see oracle/pamo_oracle.cpp) and runs the same inputs."""
import numpy as np
import pytest
import torch

import oracle
from pamo_amd import _C, debug, synth, simplify_stage2, PaMO

pytestmark = pytest.mark.gpu

NAMES = ["points", "tris", "occ", "map", "undo"]


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


def _two_spheres(level=4, seed=3, offset=0.9):
    """Two bumpy spheres interpenetrating: a mesh with built-in
    self-intersections, so the undo loop runs."""
    v1, f1 = synth.displaced_icosphere(level, seed)
    v2, f2 = synth.displaced_icosphere(level, seed + 1)
    v2 = v2 + np.array([offset, 0.05, 0.0], np.float32)
    return np.concatenate([v1, v2]).astype(np.float32), np.concatenate([f1, f2 + v1.shape[0]]).astype(np.int32)


def _zero_length_edge(level=4, seed=5):
    """Bumpy sphere with one zero-length edge: exercises the line-edge
    collapse (cusimp_free.cu:234-274) after the regular winners."""
    v, f = synth.displaced_icosphere(level, seed)
    v = v.copy()
    a, b = sorted(int(x) for x in f[0, :2])
    v[b] = v[a]
    return v, f


MESHES = {
    "line_edge": _zero_length_edge,
    "ico3": lambda: synth.displaced_icosphere(3, 7),
    "ico5": lambda: synth.displaced_icosphere(5, 1),
    "torus": lambda: synth.noisy_torus(60, 24, 5),
    "sdf": lambda: synth.sdf_union(12, 60, 4),
    "two_spheres": lambda: _two_spheres(),
}


@pytest.mark.parametrize("name", ["ico3", "torus", "sdf", "two_spheres", "line_edge"])
def test_forward_round_exact(cuda, name):
    v, f = MESHES[name]()
    verts, scale, _ = oracle.preprocess(v, f)
    print("case", name, f.shape, flush=True)
    g = _C.CUDSP_Free()
    o = oracle.Oracle()
    undo_t = torch.empty(0, dtype=torch.int32, device=cuda)
    gout = g.forward(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), undo_t, 0, float(scale), 1e-3,
                     False, True)
    oout = o.forward(verts, f, np.zeros(0, np.int32), float(scale), 1e-3, False, True)
    for n, a, b in zip(NAMES, gout, oout):
        assert _same(a.cpu().numpy(), b), n
    for n in ["fan_off", "fan", "edges", "Q", "edge_cost", "tri_min", "orig_tris", "table"]:
        a, b = g.debug(n), o.get(n)
        assert _same(a, b), n
    a, b = g.debug("collapsed"), o.get("collapsed")
    assert _same(a, b), "collapsed"
    if o.sizes()["undo_cut"] > 0:
        assert _same(g.debug("S"), o.get("S")), "S"


@pytest.mark.parametrize("name", ["ico5", "two_spheres"])
def test_rounds_exact(cuda, name):
    """The reference round loop (engine='rounds') against the oracle, every
    round's five outputs bit for bit."""
    v, f = MESHES[name]()
    verts, scale, _ = oracle.preprocess(v, f)
    target = int(0.01 * f.shape[0])
    g = _C.CUDSP_Free()
    o = oracle.Oracle()
    gv, gf = torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda)
    ov, of = verts.copy(), f.copy()
    gu = torch.empty(0, dtype=torch.int32, device=cuda)
    ou = np.zeros(0, np.int32)
    count = is_stuck = 0
    init = True
    for it in range(400):
        nprev = of.shape[0]
        go = g.forward(gv, gf, gu, gu.shape[0], float(scale), 1e-3, is_stuck, init)
        oo = o.forward(ov, of, ou, float(scale), 1e-3, is_stuck, init)
        init = False
        for n, a, b in zip(NAMES, go, oo):
            assert _same(a.cpu().numpy(), b), "round %d %s" % (it, n)
        gverts, gtris, gocc, gmap, gu = go
        gv = gverts[gocc.view(-1).bool()]
        gf = gtris[gtris[:, 0] >= 0]
        gf = gmap.view(-1)[gf.long()].int().contiguous()
        overts, otris, oocc, omap, ou = oo
        ov = overts[oocc.reshape(-1).astype(bool)]
        of = otris[otris[:, 0] >= 0]
        of = omap.reshape(-1)[of].astype(np.int32)
        n = of.shape[0]
        if n <= target or n <= 10:
            break
        if n == nprev:
            count += 1
        else:
            count = 0
            is_stuck = 0
        if count >= 2:
            is_stuck = 1
        if count == 4:
            break
    assert it > 5


@pytest.mark.parametrize("name", ["ico5", "two_spheres", "sdf", "line_edge"])
def test_device_loop_exact(cuda, name):
    """pamo_run_stage2 (device-resident loop) == the oracle's PaMO.run loop."""
    v, f = MESHES[name]()
    verts, scale, _ = oracle.preprocess(v, f)
    target = int(0.01 * f.shape[0])
    ov, of, ost = oracle.run(verts, f, scale, target)
    gv, gf, gst = simplify_stage2(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, target)
    print(name, ost, gst, flush=True)
    assert gst["rounds"] == ost["rounds"]
    assert _same(gv.cpu().numpy(), ov)
    assert _same(gf.cpu().numpy(), of)


def test_pamo_api_engines_agree(cuda):
    v, f = synth.displaced_icosphere(4, 9)
    outs = []
    for engine in ("device", "rounds"):
        m = PaMO((v, f), use_stage1=False, use_stage3=False, verbose=False, engine=engine)
        outs.append(m.run(torch.from_numpy(v).to(cuda), torch.from_numpy(f).to(cuda), 0.01, min_verts=0))
    assert _same(outs[0][0], outs[1][0]) and _same(outs[0][1], outs[1][1])


@pytest.mark.parametrize("name", ["sdf", "two_spheres"])
def test_candidate_pairs_exact(cuda, name):
    """LBVH candidate pairs and narrow-phase verdicts over every alive face
    == the oracle's brute-force BVH enumeration."""
    v, f = MESHES[name]()
    cand, hits = oracle.full_check(v, f)
    rows = debug.candidate_pairs(torch.from_numpy(v).to(cuda), torch.from_numpy(f).to(cuda))
    assert rows.shape[0] == cand.shape[0]
    assert np.array_equal(rows[:, :2], cand)
    gh = rows[rows[:, 2] == 1][:, :2]
    assert np.array_equal(gh, hits)


@pytest.mark.parametrize("name", ["ico3", "torus", "two_spheres", "line_edge"])
def test_sp_forward_exact(cuda, name):
    """CUDSP.forward (the unguarded round) against the oracle: the four
    outputs bit for bit, and the edge costs."""
    v, f = MESHES[name]()
    verts = np.asarray(v, np.float32)
    scale = float(np.max(verts.max(0) - verts.min(0)))
    print("case", name, f.shape, flush=True)
    g = _C.CUDSP()
    o = oracle.Oracle()
    gout = g.forward(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), scale, 1e-3, True)
    oout = o.forward_sp(verts, f, scale, 1e-3, True)
    for n, a, b in zip(NAMES[:4], gout, oout):
        assert _same(a.cpu().numpy(), b), n


def test_pasp_run_matches_oracle_loop(cuda):
    """PaSP.run against the same loop over the oracle's unguarded round."""
    from pamo_amd import PaSP
    v, f = MESHES["ico5"]()
    verts = np.asarray(v, np.float32)
    gv, gf = PaSP().run(torch.from_numpy(verts).to(cuda), torch.from_numpy(f).to(cuda), threshold=1e-3)
    scale = float(np.max(verts.max(0) - verts.min(0)))
    o = oracle.Oracle()
    ov, of = verts, f
    for it in range(1000):
        n0 = of.shape[0]
        P, T, occ, vmap = o.forward_sp(ov, of, scale, 1e-3, it == 0)
        ov = P[occ.ravel() == 1]
        of = vmap.ravel()[T[T[:, 0] >= 0]].astype(np.int32)
        if of.shape[0] == n0:
            break
    print("PaSP", f.shape, "->", of.shape, it + 1, flush=True)
    assert _same(gv.cpu().numpy(), ov) and _same(gf.cpu().numpy(), of)


def _crafted_cases():
    """Hand-built meshes of tests/cost_eval.py that each decide one branch:
    (name, points, faces, scale, threshold, undo list or None, unguarded)."""
    import cost_eval as ce
    cases = [("skinny_penalty", *ce.flat_patch(), 1.0, 1.0, None, False),
             ("normal_flip", *ce.flat_patch(lift=ce.FLIP_LIFT), 1.0, 1e-3, None, False),
             ("d10_quotient", *ce.flat_patch(p=ce.D10_P), 1.0, 1.0, None, True)]
    for i, (k, y, _) in enumerate(ce.RESTORE_CASES):
        cases.append(("restore_threshold_%d" % i, *ce.restore_mesh(k, y), 1.0, 1e-3, ce.RESTORE_UNDO, False))
    return cases


@pytest.mark.parametrize("case", range(7))
def test_crafted_branches_exact(cuda, case):
    """The device's skinny weight and divide, normal flip, unguarded quotient
    and skinny-restore threshold on the meshes that pin them in the oracle's
    tests: every output and internal buffer bit for bit (identity shuffle;
    the restore cases first size the tables with a call whose faces are all
    removed, then pass the undo list)."""
    name, P, F, s, tau, undo, sp = _crafted_cases()[case]
    print("case", name, flush=True)
    dv, df = torch.from_numpy(P).to(cuda), torch.from_numpy(F).to(cuda)
    if sp:
        g = _C.CUDSP(shuffle=oracle.SHUFFLE_IDENTITY)
        o = oracle.Oracle(shuffle=oracle.SHUFFLE_IDENTITY)
        gout = g.forward(dv, df, s, tau, True)
        oout = o.forward_sp(P, F, s, tau, True)
        for n, a, b in zip(NAMES, gout, oout):
            assert _same(a.cpu().numpy(), b), n
        assert _same(g.debug("edge_cost"), o.get("edge_cost"))
        return
    g = _C.CUDSP_Free(shuffle=oracle.SHUFFLE_IDENTITY)
    o = oracle.Oracle(shuffle=oracle.SHUFFLE_IDENTITY)
    none_g = torch.empty(0, dtype=torch.int32, device=cuda)
    if undo is not None:
        dead = np.full_like(F, -1)
        g.forward(dv, torch.from_numpy(dead).to(cuda), none_g, 0, s, tau, False, True)
        o.forward(P, dead, np.zeros(0, np.int32), s, tau, False, True)
        ug = torch.from_numpy(undo).to(cuda)
        gout = g.forward(dv, df, ug, ug.shape[0], s, tau, False, False)
        oout = o.forward(P, F, undo, s, tau, False, False)
    else:
        gout = g.forward(dv, df, none_g, 0, s, tau, False, True)
        oout = o.forward(P, F, np.zeros(0, np.int32), s, tau, False, True)
    for n, a, b in zip(NAMES, gout, oout):
        assert _same(a.cpu().numpy(), b), n
    for n in ["edges", "Q", "edge_cost", "tri_min", "collapsed", "table"]:
        assert _same(g.debug(n), o.get(n)), n
