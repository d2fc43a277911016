"""Stage 3 (safe projection) on the GPU against the float64 oracle
(oracle/stage3_oracle.cpp, pinned by tests/test_stage3_oracle.py).

Exact where the device and the oracle evaluate the same IEEE operations in
the same order (topology, Voronoi areas, nearest target points, the contact
list and its distances); within stated tolerances where the device sums in
another fixed order (energies, forces, H dx, CG) or where a step goes
through atan2 / log (hinge angles, barrier), whose last bits may differ
between the device's math library and the host's."""
import numpy as np
import pytest
import torch

from oracle import stage3 as o3
from s3_cases import close_spheres, stage2_pair, tetra

pytestmark = pytest.mark.gpu


def _pair(level=4, seed=3, ratio=0.02):
    gv, gf, v, f = stage2_pair(level, seed, ratio)
    sc, t = o3.normalization(gv)
    return gv * sc + t, gf, v * sc + t, f


def _systems(gv, gf, v, f, radius=None):
    from pamo_amd import safe_project as s3
    cfg_o, cfg_g = o3.Config(), s3.Stage3Config()
    if radius is not None:
        cfg_o.contact_detection_radius = cfg_g.contact_detection_radius = radius
    o = o3.System(cfg_o)
    o.register_mesh(gv, gf, v, f)
    g = s3.Stage3System(cfg_g)
    g.register_mesh(gv, gf, v, f)
    return o, g


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _ulps(a, b):
    a = np.ascontiguousarray(a, np.float64).view(np.int64)
    b = np.ascontiguousarray(b, np.float64).view(np.int64)
    return int(np.abs(a - b).max()) if a.size else 0


def test_s3_register_topology(cuda):
    gv, gf, v, f = _pair()
    o, g = _systems(gv, gf, v, f)
    so, sg = o.sizes(), g.sizes()
    assert {k: so[k] for k in ("nV", "nF", "nE", "hinges")} == {k: sg[k] for k in ("nV", "nF", "nE", "hinges")}
    e, h, r, a = g.topology()
    assert np.array_equal(e, o.edges())
    hi, hr = o.hinges()
    assert np.array_equal(h, hi)
    assert np.array_equal(r[:, 1], hr[:, 1])            # rest lengths: same operations
    assert np.abs(r[:, 0] - hr[:, 0]).max() <= 1e-14    # rest angles: atan2 of both libraries
    assert _ulps(a, o.voronoi_areas()) == 0
    print("register: %d vertices, %d faces, %d edges" % (sg["nV"], sg["nF"], sg["nE"]))


def test_s3_register_rejects_open_mesh(cuda):
    from pamo_amd import safe_project as s3
    v, f = tetra()
    bad = f.copy()
    bad[3] = bad[3][::-1]   # one face turned over: its edges are not shared oppositely
    g = s3.Stage3System()
    with pytest.raises(ValueError):
        g.register_mesh(v, f, v, bad)


@pytest.mark.parametrize("level", [4, 6])
def test_s3_targets_exact(cuda, level):
    gv, gf, v, f = _pair(level, 5, 0.02 if level == 4 else 0.005)
    o, g = _systems(gv, gf, v, f)
    o.update_target()
    g.update_target()
    tg = g.get("target").cpu().numpy()
    assert np.array_equal(tg, o.get("target")), _ulps(tg, o.get("target"))
    print("targets: %d vertices against %d target faces, bit-exact" % (len(v), len(gf)))


@pytest.mark.parametrize("level,ratio,scale,radius", [(5, 0.1, 1.0, None), (3, 0.05, 0.5, 0.05)])
def test_s3_contacts_exact(cuda, level, ratio, scale, radius):
    gv, gf, v, f = _pair(level, 2, ratio)
    o, g = _systems(gv, gf, v * scale, f, radius)
    o.part("detect")
    g.part("detect")
    ro, do = o.contacts()
    rg, dg = g.contacts()
    assert rg.shape == ro.shape and np.array_equal(rg, ro)
    assert _ulps(dg, do) == 0
    print("contacts: %d pairs, list and distances bit-exact" % len(ro))


def test_s3_contacts_active_barrier_exact(cuda):
    o, g = _systems(*close_spheres())
    o.part("detect")
    g.part("detect")
    ro, do = o.contacts()
    rg, dg = g.contacts()
    assert np.array_equal(rg, ro) and _ulps(dg, do) == 0
    assert (ro[:, 0] == o3.PT_CONTACT).any() and (ro[:, 0] == o3.EE_CONTACT).any() and (do < 1e-3).sum() > 10
    print("contacts: %d pairs (%d within d_hat), bit-exact" % (len(ro), int((do < 1e-3).sum())))


@pytest.mark.parametrize("mode", [1, 2])
def test_s3_contact_search_modes(cuda, mode):
    """The LBVH walk (mode 1) and the exhaustive tiled search (mode 2) list
    the oracle's pairs in its order, on the close spheres and on a denser
    stage-2 output with a wide radius."""
    for gv, gf, v, f, radius in (close_spheres() + (None,), _pair(5, 2, 0.1) + (0.03,)):
        o, g = _systems(gv, gf, v, f, radius)
        g.ctx.debug_guard(mode)
        o.part("detect")
        g.part("detect")
        ro, do = o.contacts()
        rg, dg = g.contacts()
        assert np.array_equal(rg, ro) and _ulps(dg, do) == 0
        print("mode %d: %d pairs bit-exact" % (mode, len(ro)))


def test_s3_cg_paths_identical(cuda):
    """The one-workgroup CG step and the multi-kernel step give the same
    bits (the launch shape never changes a result)."""
    gv, gf, v, f = close_spheres()
    out = []
    for one in (1 << 30, 0):
        o, g = _systems(gv, gf, v, f)
        g.ctx.set_tuning("s3_cg_one_block", one)
        g.update_target()
        for p in ("detect", "energy", "diff", "cg"):
            g.part(p)
        out.append(g.get("p").cpu().numpy())
    assert np.array_equal(out[0], out[1])


def test_s3_graph_replays_equal_eager(cuda):
    """The recorded Newton iteration, replayed after each eager contact
    detection, gives the eager run's positions bit for bit, iteration by
    iteration."""
    gv, gf, v, f = close_spheres()
    runs = []
    for graph in (1, 0):
        o, g = _systems(gv, gf, v, f)
        g.ctx.set_tuning("s3_graph", graph)
        g.update_target()
        qs = []
        for _ in range(6):
            g.part("newton")
            qs.append(g.get("q").cpu().numpy())
        runs.append(qs)
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_s3_radius_halving(cuda):
    """More pairs than max_blocks: the radius is halved until they fit
    (energy.py:937-947), on both sides."""
    from pamo_amd import safe_project as s3
    gv, gf, v, f = close_spheres()
    cfg_o, cfg_g = o3.Config(), s3.Stage3Config()
    cfg_o.max_blocks = cfg_g.max_blocks = 300
    o = o3.System(cfg_o)
    o.register_mesh(gv, gf, v, f)
    g = s3.Stage3System(cfg_g)
    g.register_mesh(gv, gf, v, f)
    o.part("detect")
    g.part("detect")
    assert g.scalars()["radius"] == o.scalars()["radius"] < 0.01
    assert np.array_equal(g.contacts()[0], o.contacts()[0])


@pytest.mark.parametrize("mode", [1, 2])
def test_s3_halving_without_end_fails(cuda, mode):
    """More pairs within d_hat alone than max_blocks: halving the radius
    stops changing the threshold, and the detection fails with an error
    (the reference recurses without end there) under both searches; a
    max_blocks beyond the int offsets is refused."""
    from pamo_amd import safe_project as s3
    from pamo_amd._lib import PamoError
    gv, gf, v, f = close_spheres()
    cfg = s3.Stage3Config()
    cfg.max_blocks = 10
    g = s3.Stage3System(cfg)
    g.register_mesh(gv, gf, v, f)
    g.ctx.debug_guard(mode)
    with pytest.raises(PamoError, match="max_blocks"):
        g.part("detect")
    cfg.max_blocks = 1 << 31
    with pytest.raises(PamoError, match="invalid configuration"):
        s3.Stage3System(cfg, ctx=g.ctx)


CASES = {"icosphere": lambda: _pair(4, 3), "close_spheres": close_spheres}


def _state_pair(case):
    o, g = _systems(*CASES[case]())
    o.update_target()
    g.update_target()
    return o, g


@pytest.mark.parametrize("case", sorted(CASES))
def test_s3_energy_forces_hdx(cuda, case):
    o, g = _state_pair(case)
    for s in (o, g):
        s.part("detect")
        s.part("energy")
        s.part("diff")
    q = o.get("q")
    eo, eg = o.energy_parts(q), g.energy_parts(q)
    for k in eo:
        assert abs(eg[k] - eo[k]) <= 1e-13 * max(abs(eo[k]), 1e-300) + 1e-300, (k, eg[k], eo[k])
    assert _rel(g.get("grad").cpu().numpy(), o.get("grad")) <= 1e-12
    assert _rel(g.get("diag").cpu().numpy(), o.get("diag")) <= 1e-12
    dx = np.random.default_rng(1).standard_normal(q.shape)
    assert _rel(g.hess_dx(dx), o.hess_dx(dx)) <= 1e-12
    print("%s: energy parts %s" % (case, {k: "%.6e" % v for k, v in eg.items()}))


@pytest.mark.parametrize("case", sorted(CASES))
def test_s3_cg_solution(cuda, case):
    o, g = _state_pair(case)
    for s in (o, g):
        for p in ("detect", "energy", "diff", "cg"):
            s.part(p)
    assert _rel(g.get("p").cpu().numpy(), o.get("p")) <= 1e-9


def test_s3_tetra_system_solved(cuda):
    """The tetrahedron's CG step against numpy's solve of the assembled H
    (the same system test_stage3_oracle.py solves on the oracle)."""
    from pamo_amd import safe_project as s3
    v, f = tetra()
    g = s3.Stage3System()
    g.register_mesh(v, f, v, f)
    g.set("target", v + 0.01)
    g.set("q", v + np.array([[0.0, 0, 0], [0.05, 0, 0], [0, -0.02, 0], [0, 0, 0.03]]))
    for p in ("detect", "energy", "diff", "cg"):
        g.part(p)
    H = np.stack([g.hess_dx(np.eye(12)[k].reshape(4, 3)).reshape(-1) for k in range(12)], axis=1)
    gr = g.get("grad").cpu().numpy().reshape(-1)
    p = np.linalg.solve(0.5 * (H + H.T), gr)
    assert np.abs(g.get("p").cpu().numpy().reshape(-1) - p).max() <= 1e-8 * np.abs(p).max()


@pytest.mark.parametrize("case", sorted(CASES))
def test_s3_newton_iterations(cuda, case):
    o, g = _state_pair(case)
    for it in range(5):
        o.part("newton")
        g.part("newton")
        so, sg = o.scalars(), g.scalars()
        qo, qg = o.get("q"), g.get("q").cpu().numpy()
        err = np.abs(qg - qo).max()
        print("%s newton %d: energy %.12e / %.12e, ccd %.6f / %.6f, halvings %d / %d, |dq| %.2e" %
              (case, it, sg["energy"], so["energy"], sg["ccd_step"], so["ccd_step"], sg["ls_halvings"],
               so["ls_halvings"], err))
        assert sg["ls_halvings"] == so["ls_halvings"]
        assert abs(sg["ccd_step"] - so["ccd_step"]) <= 1e-9
        assert abs(sg["energy"] - so["energy"]) <= 1e-10 * abs(so["energy"])
        assert err <= 1e-9


def test_s3_large_mesh_lbvh_contacts_and_newton(cuda):
    """A 19,928-face mesh (above the exhaustive contact search's 16,384-face
    limit, so the LBVH walks run by default) with a 81,920-face target:
    targets and the ~490k-pair contact list bit-exact, the exhaustive search
    forced on the same positions equal too, and one Newton iteration within
    tolerance."""
    gv, gf, v, f = _pair(6, 8, 0.25)
    assert len(f) > 16384
    o, g = _systems(gv, gf, v, f)
    o.update_target()
    g.update_target()
    assert np.array_equal(g.get("target").cpu().numpy(), o.get("target"))
    o.part("detect")
    g.part("detect")
    ro, do = o.contacts()
    rg, dg = g.contacts()
    assert len(ro) > 100000 and np.array_equal(rg, ro) and _ulps(dg, do) == 0
    g.ctx.debug_guard(2)
    g.part("detect")
    rg2, dg2 = g.contacts()
    assert np.array_equal(rg2, ro) and _ulps(dg2, do) == 0
    g.ctx.debug_guard(0)
    g.set("q", o.get("q"))
    for p in ("energy", "diff", "cg", "clamp", "ccd", "line_search"):
        o.part(p)
        g.part(p)
    so, sg = o.scalars(), g.scalars()
    err = np.abs(g.get("q").cpu().numpy() - o.get("q")).max()
    print("large: %d faces, %d contacts, energy %.12e / %.12e, |dq| %.2e" % (len(f), len(ro), sg["energy"],
                                                                          so["energy"], err))
    assert sg["ls_halvings"] == so["ls_halvings"] and abs(sg["ccd_step"] - so["ccd_step"]) <= 1e-9
    assert abs(sg["energy"] - so["energy"]) <= 1e-10 * abs(so["energy"]) and err <= 1e-9


def test_s3_pamo_all_stages(cuda):
    """PaMO(use_stage1=True, use_stage3=True): remeshing, stage 2, then stage
    3 toward the input; the stage-3 part against the oracle on the same
    stage-2 output."""
    from pamo_amd import PaMO, synth
    v, f = synth.displaced_icosphere(4, 9)
    m = PaMO((v, f), use_stage1=True, use_stage3=True, verbose=False)
    out_v, out_f = m.run(torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda(), 0.05, min_verts=0)
    s2v, s2f = m.last_stage2_output
    vo, _ = o3.process(v, f, s2v, s2f, 5)
    ext = float(np.max(v.max(0) - v.min(0)))
    err = float(np.abs(out_v - vo).max()) / ext
    print("PaMO stages 1+2+3: %d -> %d faces, max |GPU - oracle| / extent = %.3e" % (len(f), len(out_f), err))
    assert np.array_equal(out_f, s2f) and np.isfinite(out_v).all() and err <= 1e-6


@pytest.mark.parametrize("level,ratio", [(4, 0.02), (6, 0.05)])
def test_s3_process_matches_oracle(cuda, level, ratio):
    from pamo_amd import safe_project as s3
    gv, gf, v, f = stage2_pair(level, 4, ratio)
    vo, fo = o3.process(gv, gf, v, f, 2)
    vg, fg = s3.process(gv, gf, v, f, 2)
    ext = float(np.max(gv.max(0) - gv.min(0)))
    err = float(np.abs(vg - vo).max()) / ext
    print("process (2 steps, %d vertices, %d target faces): max |GPU - oracle| / extent = %.3e" %
          (len(v), len(gf), err))
    assert np.array_equal(np.asarray(fg), np.asarray(fo))
    assert err <= 1e-7


def test_s3_pamo_run_with_stage3(cuda):
    """PaMO(use_stage3=True): stage 2 then stage 3 with 5 steps (the
    reference's pamo/__init__.py:160-176), against the oracle on the same
    stage-2 output."""
    from pamo_amd import PaMO, synth
    v, f = synth.displaced_icosphere(4, 6)
    m = PaMO((v, f), use_stage1=False, use_stage3=True, verbose=False)
    out_v, out_f = m.run(torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda(), 0.02, min_verts=0)
    s2v, s2f = m.last_stage2_output
    vo, _ = o3.process(v, f, s2v, s2f, 5)
    ext = float(np.max(v.max(0) - v.min(0)))
    err = float(np.abs(out_v - vo).max()) / ext
    print("PaMO stage 2 + 3: %d -> %d faces, max |GPU - oracle| / extent = %.3e" % (len(f), len(out_f), err))
    assert np.array_equal(out_f, s2f)
    assert err <= 1e-6
