"""Pins the float64 stage-3 oracle (oracle/stage3_oracle.cpp) with checks that
do not restate its formulas: independent closest-point and segment-distance
routines, central differences, closed forms on hand-built elements, numpy's
symmetric eigensolver and linear solver, and invariants of a Newton step.

The one part no check reaches is the curvature term of the reference's
hinge Hessian on the two edge vertices (hinge_energy.py:190-236): it is not
the derivative of the reference's own force (test_hinge_hessian_edge_blocks
records how far it is), so the oracle keeps the reference's expressions and
that term stays parity unpinned."""
import numpy as np
import pytest

from oracle import stage3 as s3

from s3_cases import stage2_pair, tetra


def _fd(f, x, h=1e-6):
    x = np.array(x, np.float64)
    g = np.zeros_like(x)
    for i in np.ndindex(x.shape):
        a, b = x.copy(), x.copy()
        a[i] += h
        b[i] -= h
        g[i] = (f(a) - f(b)) / (2 * h)
    return g


def _seg_closest(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return a + t * ab


def _tri_dist(p, a, b, c):
    """Point-triangle distance: the plane foot when it is inside (barycentric
    test), else the nearest of the three edges."""
    n = np.cross(b - a, c - a)
    foot = p - np.dot(p - a, n) / np.dot(n, n) * n
    m = np.array([b - a, c - a]).T
    uv = np.linalg.lstsq(m, foot - a, rcond=None)[0]
    if uv[0] >= 0 and uv[1] >= 0 and uv.sum() <= 1:
        return np.linalg.norm(p - foot)
    return min(np.linalg.norm(p - _seg_closest(p, x, y)) for x, y in ((a, b), (b, c), (c, a)))


def _seg_seg_dist(x0, x1, x2, x3):
    """Segment-segment distance: the interior solution of the two lines when
    both parameters are inside [0, 1], else the best endpoint-to-segment."""
    c = [np.linalg.norm(x0 - _seg_closest(x0, x2, x3)), np.linalg.norm(x1 - _seg_closest(x1, x2, x3)),
         np.linalg.norm(x2 - _seg_closest(x2, x0, x1)), np.linalg.norm(x3 - _seg_closest(x3, x0, x1))]
    u, v, w = x1 - x0, x3 - x2, x0 - x2
    A, B, C, D, E = u @ u, u @ v, v @ v, u @ w, v @ w
    den = A * C - B * B
    if den > 1e-12 * A * C:
        s, t = (B * E - C * D) / den, (A * E - B * D) / den
        if 0 <= s <= 1 and 0 <= t <= 1:
            c.append(np.linalg.norm(x0 + s * u - x2 - t * v))
    return min(c)


# ------------------------------------------------------------ distances
def test_point_triangle_distance_matches_closest_point():
    rng = np.random.default_rng(0)
    seen = set()
    for _ in range(3000):
        x = rng.normal(size=(4, 3))
        t, d, _ = s3.pt(x)
        seen.add(t)
        ref = _tri_dist(*x)
        assert abs(d - ref) <= 1e-9 * max(1.0, ref), (t, d, ref)
        cp = s3.closest(x[0], x[1:])
        assert abs(np.linalg.norm(cp - x[0]) - ref) <= 1e-9 * max(1.0, ref)
    assert seen == set(range(7))   # every region of the classification reached


def test_point_triangle_types_by_hand():
    tri = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    cases = {(0.2, 0.2, 0.5): 0, (-1, -1, 0.3): 1, (2, -0.5, 0.1): 2, (0.5, -1, 0.2): 3, (-0.5, 2, 0.1): 4,
             (-1, 0.5, 0.2): 5, (1, 1, 0.2): 6}
    for p, t in cases.items():
        x = np.vstack([np.array(p, float), tri])
        assert s3.pt(x)[0] == t, p


def test_degenerate_triangle_takes_nearest_vertex():
    x = np.array([[0.2, 0.3, 0.0], [0, 0, 0], [1, 0, 0], [2, 0, 0]])   # collinear triangle
    t, d, _ = s3.pt(x)
    assert t == 1 and d == pytest.approx(np.hypot(0.2, 0.3))


def test_edge_edge_distance_matches_segment_distance():
    rng = np.random.default_rng(1)
    seen = set()
    for _ in range(3000):
        x = rng.normal(size=(4, 3))
        t, d, _ = s3.ee(x, 1e-3)
        seen.add(t)
        ref = _seg_seg_dist(*x)
        assert abs(d - ref) <= 1e-9 * max(1.0, ref), (t, d, ref)
    assert seen == set(range(9))


def test_parallel_edges_never_take_the_line_distance():
    x = np.array([[0.0, 0, 0], [1, 0, 0], [0.2, 0.1, 0], [0.9, 0.1, 1e-5]])   # sin ~ 1.4e-5 < 1e-3
    t, d, _ = s3.ee(x, 1e-3)
    assert t != 8 and d == pytest.approx(_seg_seg_dist(*x), rel=1e-12)


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_distance_gradients_by_differences(kind):
    rng = np.random.default_rng(2)
    want = set(range(7 if kind == "pt" else 9))
    seen = set()
    f = s3.pt if kind == "pt" else (lambda x: s3.ee(x, 1e-3))
    while seen != want:
        x = rng.normal(size=(4, 3))
        t, d, g = f(x)
        if d < 1e-2:
            continue
        gf = _fd(lambda y: f(y)[1] if f(y)[0] == t else np.nan, x)
        if np.isnan(gf).any():
            continue   # a region boundary within the step
        assert np.abs(g - gf).max() <= 1e-7, (t, g, gf)
        seen.add(t)


# --------------------------------------------------------------- barrier
def test_barrier_closed_form_and_derivatives():
    dh = 1e-3
    assert np.allclose(s3.barrier(dh, dh), 0.0) and np.allclose(s3.barrier(2 * dh, dh), 0.0)
    for d in (1e-5, 1e-4, 5e-4, 9e-4):
        b, b1, b2 = s3.barrier(d, dh)
        assert b == pytest.approx(-(d - dh) ** 2 * np.log(d / dh), rel=1e-14)
        h = d * 1e-5
        assert b1 == pytest.approx((s3.barrier(d + h, dh)[0] - s3.barrier(d - h, dh)[0]) / (2 * h), rel=1e-6)
        assert b2 == pytest.approx((s3.barrier(d + h, dh)[1] - s3.barrier(d - h, dh)[1]) / (2 * h), rel=1e-6)
        assert b > 0 and b1 < 0 and b2 > 0


# --------------------------------------------------------------- elastic
def test_elastic_pure_stretch_closed_form():
    rest = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    mu, la = 0.05, 0.03
    for a, b in ((1.0, 1.0), (1.2, 0.9), (0.5, 2.0)):
        x = rest * np.array([a, b, 1.0])
        E, g, _ = s3.elastic(rest, x, np.zeros((3, 3)), mu, la)
        ex, ey = (a * a - 1) / 2, (b * b - 1) / 2
        assert E == pytest.approx(0.5 * (mu * (ex * ex + ey * ey) + 0.5 * la * (ex + ey) ** 2), rel=1e-12, abs=1e-15)


def test_elastic_force_and_hessian_by_differences():
    rng = np.random.default_rng(3)
    for _ in range(20):
        rest = rng.normal(size=(3, 3))
        x = rest @ np.linalg.qr(rng.normal(size=(3, 3)))[0] + 0.2 * rng.normal(size=(3, 3))
        E, g, _ = s3.elastic(rest, x, np.zeros((3, 3)), 0.05, 0.02)
        assert np.abs(g - _fd(lambda y: s3.elastic(rest, y, np.zeros((3, 3)), 0.05, 0.02)[0], x)).max() \
            <= 1e-7 * max(1.0, np.abs(g).max())
        # H dx for a displacement in the triangle's plane: the derivative of the force
        n = np.cross(x[1] - x[0], x[2] - x[0])
        n /= np.linalg.norm(n)
        dx = rng.normal(size=(3, 3))
        dx -= np.outer(dx @ n, n)
        _, _, hdx = s3.elastic(rest, x, dx, 0.05, 0.02)
        h = 1e-6
        fd = (s3.elastic(rest, x + h * dx, dx, 0.05, 0.02)[1] - s3.elastic(rest, x - h * dx, dx, 0.05, 0.02)[1]) / (2 * h)
        assert np.abs(hdx - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max())


# ----------------------------------------------------------------- hinge
def _flap(beta):
    """Flaps (x0; x1, x2) in the plane z = 0 and (x3; x1, x2) raised by beta
    about the edge x1x2 on the x axis: the signed angle is -beta."""
    return np.array([[0.5, 1.0, 0.0], [0, 0, 0], [1.0, 0, 0], [0.5, -np.cos(beta), np.sin(beta)]])


def test_hinge_angle_closed_form():
    for beta in (0.0, 0.3, -0.7, 1.5, 2.9):
        th, E, _, _, _ = s3.hinge(_flap(0.0), _flap(beta), 1e-2)
        assert th == pytest.approx(-beta, abs=1e-12)
        assert E == pytest.approx(0.5 * 1e-2 * beta * beta * 1.0, rel=1e-12, abs=1e-16)


def test_hinge_force_by_differences():
    rng = np.random.default_rng(4)
    for _ in range(20):
        rest = rng.normal(size=(4, 3))
        x = rest + 0.3 * rng.normal(size=(4, 3))
        _, _, g, _, _ = s3.hinge(rest, x)
        gf = _fd(lambda y: s3.hinge(rest, y)[1], x)
        assert np.abs(g - gf).max() <= 1e-7 * max(1e-3, np.abs(gf).max())


def _hinge_fd_hessian(rest, x):
    H = np.zeros((12, 12))
    for i in range(12):
        a, b = x.reshape(-1).copy(), x.reshape(-1).copy()
        a[i] += 1e-6
        b[i] -= 1e-6
        H[:, i] = (s3.hinge(rest, a)[2].reshape(-1) - s3.hinge(rest, b)[2].reshape(-1)) / 2e-6
    return H


def test_hinge_hessian_at_rest_and_wing_blocks():
    """At the rest angle the Hessian is k l0 grad(theta) grad(theta)^T; away from
    it the blocks of the two wing vertices x0, x3 still equal the differences."""
    rng = np.random.default_rng(5)
    for _ in range(10):
        x = rng.normal(size=(4, 3))
        _, _, _, H, _ = s3.hinge(x, x)
        Hf = _hinge_fd_hessian(x, x)
        assert np.abs(H - Hf).max() <= 1e-6 * np.abs(Hf).max()
        rest = x + 0.3 * rng.normal(size=(4, 3))
        _, _, _, H, _ = s3.hinge(rest, x)
        Hf = _hinge_fd_hessian(rest, x)
        for i, j in ((0, 0), (3, 3), (0, 3), (3, 0)):
            blk = (slice(3 * i, 3 * i + 3), slice(3 * j, 3 * j + 3))
            assert np.abs(H[blk] - Hf[blk]).max() <= 1e-6 * np.abs(Hf).max()


def test_hinge_hessian_edge_blocks():
    """Records the reference's curvature term on the edge vertices: it differs
    from the derivative of the force (so it is restated, not pinned), and it
    is symmetric."""
    rng = np.random.default_rng(6)
    x = rng.normal(size=(4, 3))
    rest = x + 0.3 * rng.normal(size=(4, 3))
    _, _, _, H, _ = s3.hinge(rest, x)
    Hf = _hinge_fd_hessian(rest, x)
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
    assert np.abs(H - Hf).max() > 1e-2 * np.abs(Hf).max()


def test_spd_projection_matches_eigendecomposition():
    rng = np.random.default_rng(7)
    for _ in range(20):
        A = rng.normal(size=(12, 12))
        A = A + A.T
        w, Q = np.linalg.eigh(A[:9, :9])
        ref9 = (Q * np.maximum(w, 0)) @ Q.T
        # converged sweeps give the nearest PSD matrix; the three the
        # reference's configuration asks for are close to it and, being a
        # rotation of clamped values, exactly PSD
        P = s3.spd_project(A, 30)
        assert np.abs(P[:9, :9] - ref9).max() <= 1e-10 * np.abs(ref9).max()
        P = s3.spd_project(A, 3)
        assert np.abs(P[:9, :9] - ref9).max() <= 2e-2 * np.abs(ref9).max()
        assert np.linalg.eigvalsh(P).min() >= -1e-9 * np.abs(P).max()
        assert np.abs(P.reshape(4, 3, 12).sum(0)).max() <= 1e-12 * np.abs(P).max()   # zero row sums
        assert np.abs(P - P.T).max() <= 1e-12 * np.abs(P).max()


# ---------------------------------------------------------- mesh helpers
def test_voronoi_areas_by_hand():
    eq = np.array([[0.0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]])
    a = s3.voronoi(eq, np.array([[0, 1, 2]]))
    assert np.allclose(a, np.sqrt(3) / 12)
    right = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    assert np.allclose(s3.voronoi(right, np.array([[0, 1, 2]])), [0.25, 0.125, 0.125])
    obtuse = np.array([[0.0, 0, 0], [1, 0, 0], [-1, 0.5, 0]])   # obtuse at vertex 0
    A = 0.5 * 0.5
    assert np.allclose(s3.voronoi(obtuse, np.array([[0, 1, 2]])), [A / 2, A / 4, A / 4])


def test_voronoi_areas_partition_a_closed_mesh():
    gv, gf, v, f = stage2_pair(3, 1, 0.05)
    a = s3.voronoi(v, f)
    area = 0.5 * np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]), axis=1).sum()
    assert a.sum() == pytest.approx(area, rel=1e-12) and (a > 0).all()


# ------------------------------------------------------------------ ACCD
def test_accd_stops_short_of_a_head_on_contact():
    x = np.array([[-1.0, -1.0, 0.1], [-5, -5, 0], [5, -5, 0], [-5, 5, 0]])   # above the interior
    v = np.zeros((4, 3))
    v[0] = [0, 0, -0.2]   # contact at t = 0.5
    t = s3.accd(x, v, s3.PT_CONTACT)
    assert 0.0 < t < 0.5
    # the gap left at t is at least (1 - slackness) of the start; later
    # steps of 0.8 x the gap cannot have passed it
    gap = 0.1 - 0.2 * t
    assert gap >= (1 - 0.7) * 0.1 * (1 - 1e-9)
    assert s3.accd(x, -v, s3.PT_CONTACT) == 1.0      # moving away: no limit
    assert s3.accd(x, np.zeros((4, 3)), s3.PT_CONTACT) == 1.0


def test_accd_edge_edge_crossing():
    x = np.array([[-1, 0, 0.05], [1, 0, 0.05], [0, -1, 0], [0, 1, 0]])
    v = np.zeros((4, 3))
    v[0] = v[1] = [0, 0, -0.1]   # edges meet at t = 0.5
    t = s3.accd(x, v, s3.EE_CONTACT)
    assert 0.0 < t < 0.5


# ---------------------------------------------------------------- system
def _system(v, f, gv=None, gf=None, cfg=None):
    s = s3.System(cfg)
    s.register_mesh(v if gv is None else gv, f if gf is None else gf, v, f)
    return s


def test_register_edges_hinges_and_open_mesh():
    v, f = tetra()
    s = _system(v, f)
    sz = s.sizes()
    assert (sz["nE"], sz["hinges"]) == (6, 6)
    idx, rest = s.hinges()
    for h, (th, l) in zip(idx, rest):   # rest angle of each tetrahedron edge, recomputed
        assert l == pytest.approx(np.linalg.norm(v[h[1]] - v[h[2]]), rel=1e-12)
    with pytest.raises(ValueError):
        _system(v, f[:3])


def test_contacts_equal_brute_force():
    gv, gf, v, f = stage2_pair(3, 2, 0.05)
    cfg = s3.Config()
    cfg.contact_detection_radius = 0.05   # wide enough to catch many pairs
    s = _system(v * 0.5, f, cfg=cfg)
    s.part("detect")
    rows, d = s.contacts()
    q = s.get("q")
    thr = 2 * 0.05 + cfg.d_hat
    want = set()
    for i in range(len(q)):
        for t in f:
            if i in t:
                continue
            if _tri_dist(q[i], *q[t]) <= thr:
                want.add((3, i, *t))
    E = s.edges()
    for a in range(len(E)):
        for b in range(a + 1, len(E)):
            if set(E[a]) & set(E[b]):
                continue
            if _seg_seg_dist(q[E[a][0]], q[E[a][1]], q[E[b][0]], q[E[b][1]]) <= thr:
                want.add((4, *E[a], *E[b]))
    got = set(map(tuple, rows[:, :5].tolist()))
    assert got == want and len(got) == len(rows) and len(rows) > 100


def test_cg_solves_the_tetrahedron_system():
    v, f = tetra()
    cfg = s3.Config()
    s = _system(v, f, cfg=cfg)
    s.set("target", v + 0.01)
    s.set("q", v + np.array([[0.0, 0, 0], [0.05, 0, 0], [0, -0.02, 0], [0, 0, 0.03]]))
    s.part("detect")
    s.part("energy")
    s.part("diff")
    s.part("cg")
    H = np.stack([s.hess_dx(np.eye(12)[k].reshape(4, 3)).reshape(-1) for k in range(12)], axis=1)
    g = s.get("grad").reshape(-1)
    p = np.linalg.solve(0.5 * (H + H.T), g)
    assert np.abs(s.get("p").reshape(-1) - p).max() <= 1e-8 * np.abs(p).max()
    # the gradient is minus the derivative of the energy
    q0 = s.get("q")
    gf = _fd(lambda y: sum(s.energy_parts(y).values()), q0)
    assert np.abs(g.reshape(4, 3) + gf).max() <= 1e-6 * np.abs(gf).max()


def test_newton_step_lowers_the_energy():
    gv, gf, v, f = stage2_pair(4, 3, 0.02)
    sc, t = s3.normalization(gv)
    s = _system(v * sc + t, f, gv * sc + t, gf)
    s.update_target()
    for _ in range(5):
        s.part("newton")
        sc_ = s.scalars()
        assert sc_["energy"] <= sc_["energy_prev"] or sc_["ls_halvings"] == 9
        assert 0.0 < sc_["ccd_step"] <= 1.0


def test_process_moves_the_mesh_toward_the_target():
    gv, gf, v, f = stage2_pair(4, 4, 0.02)
    out, of = s3.process(gv, gf, v, f, 2)
    sc, t = s3.normalization(gv)
    s = _system(v * sc + t, f, gv * sc + t, gf)
    s.update_target()
    before = np.linalg.norm(s.get("q") - s.get("target"), axis=1).mean()
    s.set("q", out * sc + t)
    s.update_target()
    after = np.linalg.norm(s.get("q") - s.get("target"), axis=1).mean()
    assert after < 0.5 * before
    assert np.array_equal(of, f)
