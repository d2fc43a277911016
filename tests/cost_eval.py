"""Test helpers: hand-built meshes that drive single branches of the edge
cost and of the collapse, and a second, independent restatement of the
reference's edge cost in numpy float32 scalars.

The restatement follows simp_cuda/src/cusimp_free.cu:382-585 (guarded) and
simp_cuda/src/cusimp.cu:283-482 (unguarded) expression by expression, with
the operator order of the reference's vector types
(simp_cuda/src/cusimp_free.h:25-152: dot, cross, norm, vTMv), Heron's area
(cusimp_free.cu:368-375) and clamp (cusimp_free.cu:29-34).  Every numpy
float32 scalar operation rounds once, as the reference built without FMA
contraction does (simp_cuda/setup.py:34).  Fans are listed in ascending face
order (definition D0 of the oracle).
"""
import numpy as np

f32 = np.float32
U32MAX = 0xFFFFFFFF


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _norm(a):
    return np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def _div(a, s):
    return (a[0] / s, a[1] / s, a[2] / s)


def _clamp(x, lo, hi):
    m = lo if lo > x else x
    return m if m < hi else hi


def _area(p0, p1, p2):
    a, b, c = _norm(_sub(p0, p1)), _norm(_sub(p1, p2)), _norm(_sub(p2, p0))
    s = (a + b + c) / f32(2)
    return np.sqrt(s * (s - a) * (s - b) * (s - c))


def _u32(x):
    """uint32_t(x / COST_RANGE * UINT32_MAX) of a value clamped to [0, 10]."""
    x = f32(x) / f32(10) * f32(4294967295.0)
    if not (x > 0):
        return 0
    if x >= f32(4294967295.0):
        return U32MAX
    return int(x)


def edge_cost(P, T, u, v, s, tau, unguarded=False, d10=True, flip_check=True, skinny_weight=5.0):
    """Quantized cost of edge (u, v) on points P [V,3] and faces T [F,3].
    Returns None where the reference leaves the cost unwritten (link count
    other than 2).  unguarded: the variant of cusimp.cu; d10: its skinny
    quotient evaluated in double (sqrt(3) of an int is the double overload)
    and rounded to float once.  flip_check and skinny_weight exist so a test
    can show which branch decides a result."""
    P = [tuple(f32(c) for c in p) for p in np.asarray(P, np.float32)]
    T = [tuple(int(c) for c in t) for t in np.asarray(T)]
    s, tau = f32(s), f32(tau)
    fan = {x: [f for f in range(len(T)) if x in T[f]] for x in (u, v)}

    def succ(t, x):
        return t[(t.index(x) + 1) % 3]

    dup = 0                                   # cusimp_free.cu:405-441
    for fu in fan[u]:
        iu = succ(T[fu], u)
        for fv in fan[v]:
            if iu == succ(T[fv], v):
                dup += 1
            if dup > 2:
                return U32MAX
    if dup != 2:
        return None
    el, ne = f32(0), 0                        # cusimp_free.cu:444-486
    for x in (u, v):
        for f in fan[x]:
            i, j, k = T[f]
            if i > j:
                el = el + _norm(_sub(P[i], P[j]))
                ne += 1
            if j > k:
                el = el + _norm(_sub(P[j], P[k]))
                ne += 1
            if k > i:
                el = el + _norm(_sub(P[k], P[i]))
                ne += 1
    el = el / f32(ne) / s * tau
    if unguarded:                             # cusimp.cu:378: edge_len += edge_len + ...
        el = el + (el + _norm(_sub(P[u], P[v])) / s * tau)
    else:                                     # cusimp_free.cu:489
        el = el + _norm(_sub(P[u], P[v])) / s * tau
    m = _div(_add(P[u], P[v]), f32(2))        # cusimp_free.cu:491
    qa, nt = f32(0), 0                        # cusimp_free.cu:497-575
    for mv, ot in ((u, v), (v, u)):
        for f in fan[mv]:
            t = T[f]
            if ot in t:
                continue
            o = [P[c] for c in t]
            n = [m if c == mv else P[c] for c in t]
            on = _cross(_sub(o[1], o[0]), _sub(o[2], o[0]))
            on = _div(on, _norm(on))
            nn = _cross(_sub(n[1], n[0]), _sub(n[2], n[0]))
            nn = _div(nn, _norm(nn))
            if flip_check and _dot(on, nn) < 0:
                return U32MAX
            e01, e12, e20 = _norm(_sub(n[0], n[1])), _norm(_sub(n[1], n[2])), _norm(_sub(n[2], n[0]))
            if unguarded and d10:
                term = f32(4.0 * np.sqrt(3.0) * np.float64(_area(*n)) / np.float64(e01 * e01) +
                           np.float64(e12 * e12) + np.float64(e20 * e20) + np.float64(f32(1e-7)))
            else:
                term = f32(4) * np.sqrt(f32(3)) * _area(*n) / (e01 * e01) + e12 * e12 + e20 * e20 + f32(1e-7)
            qa = qa + (f32(1) - _clamp(term, f32(0), f32(1)))
            nt += 1
    if not unguarded:
        qa = f32(skinny_weight) * qa          # cusimp_free.cu:578-579

    def plane(t):                             # cusimp_free.cu:337-346
        v0, v1, v2 = P[t[0]], P[t[1]], P[t[2]]
        nrm = _cross(_sub(v1, v0), _sub(v2, v0))
        nrm = _div(nrm, _norm(nrm))
        return (nrm[0], nrm[1], nrm[2], -_dot(nrm, v0))

    Q = {}
    for x in (u, v):                          # cusimp_free.cu:348-366
        K = [f32(0)] * 16
        for f in fan[x]:
            p = plane(T[f])
            K = [K[4 * r + c] + p[r] * p[c] for r in range(4) for c in range(4)]
        Q[x] = K
    M = [Q[u][i] + Q[v][i] for i in range(16)]
    w = (m[0], m[1], m[2], f32(1))
    tv = [M[c] * w[0] + M[4 + c] * w[1] + M[8 + c] * w[2] + M[12 + c] * w[3] for c in range(4)]
    qe = tv[0] * w[0] + tv[1] * w[1] + tv[2] * w[2] + tv[3] * w[3]
    cost = qe / (s * s)                       # cusimp_free.cu:581-584
    return _u32(_clamp(cost + el + qa / f32(nt) * tau, f32(0), f32(10)))


def flat_patch(k=0.125, p=(-0.5, 0.1), h=1.0, lift=None):
    """A closed 18-face mesh: a flat patch in z = 0 whose one interior edge
    (u, v) = (0, 1) has hexagonal one-rings, closed by a cone to an apex
    below.  u = (0,0), v = (1,0), ring of u: v, a, p, q, r, b; ring of v: u,
    b, s, t, w, a (times k).  Every face of both fans lies in z = 0, so the
    QEM term of (u, v) at its midpoint is exactly 0.  lift: {vertex: (dx, dz)}
    moves ring vertices off the plane."""
    pts = [(0, 0), (1, 0), (0.5, 1), p, (-1, 0), (-0.5, -1), (0.5, -1), (2, 0), (1.5, 1), (1.5, -1)]
    P = np.array([[x * k, y * k, 0.0] for x, y in pts] + [[0.5 * k, 0.0, -h * k]], np.float32)
    if lift:
        for i, (dx, dz) in lift.items():
            P[i, 0] += np.float32(dx)
            P[i, 2] = np.float32(dz)
    u, v, a, pp, q, r, b, t, w, s, ap = range(11)
    F = [(u, v, a), (u, a, pp), (u, pp, q), (u, q, r), (u, r, b), (u, b, v), (v, t, w), (v, w, a), (v, b, s),
         (v, s, t)]
    ring = [a, pp, q, r, b, s, t, w]
    for i in range(8):
        F.append((ring[(i + 1) % 8], ring[i], ap))
    return P, np.array(F, np.int32)


# A valid closed mesh (verified) whose collapse of (u, v) to the midpoint
# turns the moved face (u, p, q) over: the oracle's cost must be UINT32_MAX.
FLIP_LIFT = {3: (0.0661, -0.0227), 4: (0.0158, -0.1778), 5: (-0.0643, -0.0719)}

# Restore threshold of the collapse (cusimp_free.cu:617-637): 1e-6 is a
# double literal, so the threshold 1e-6 * max_len * max_len is formed in
# double and rounded to float once.  At these scales k the float-evaluated
# product differs from it by one ulp, and the apex height y of p puts the
# moved face's float area exactly on the lower of the two:
#   (k, y, restored by the reference)
RESTORE_CASES = [
    (0.1002499982714653, 3.0074997425799666e-07, True),    # area = thr_f < thr_d: skinny
    (0.1002499982714653, 3.007500026797061e-07, False),    # one ulp up: area > thr_d
    (0.10140000283718109, 3.041999718789157e-07, False),   # area = thr_d < thr_f: not skinny
    (0.10140000283718109, 3.0419994345720625e-07, True),   # one ulp down: area < thr_d
]


def restore_mesh(k, y):
    P, F = flat_patch(k=k, p=(-0.5, 0.0))
    P[3, 1] = np.float32(y)
    return P, F


# every vertex but v = 1: the undo list that makes every edge not touching v
# cost UINT32_MAX (both endpoints in the invalid table, cusimp_free.cu:395),
# so (u, v) is the round's winner
RESTORE_UNDO = np.array([x for x in range(11) if x != 1], np.int32)


def skinny_thresholds(p1, q1, r1):
    """(float area, threshold formed in double, threshold formed in float)
    of is_skinny_triangle (cusimp_free.cu:617-637) on one face."""
    p1, q1, r1 = [tuple(f32(c) for c in x) for x in (p1, q1, r1)]
    pq, pr, qr = _sub(q1, p1), _sub(r1, p1), _sub(r1, q1)
    ml = max(_norm(pq), max(_norm(pr), _norm(qr)))
    c = _cross(pq, pr)
    area = f32(0.5) * np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2])
    return area, f32(1e-6 * np.float64(ml) * np.float64(ml)), f32(1e-6) * ml * ml


# the unguarded variant's quotient (D10): at this p the double evaluation and
# a float evaluation give quantized costs 16 apart
D10_P = (-0.4672, 0.0964)
