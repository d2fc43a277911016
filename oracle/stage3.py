"""ctypes bindings of the float64 stage-3 (safe projection) oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's CPU leg, always
as the checker.  The product package ``pamo_amd`` never imports it.

The oracle (stage3_oracle.cpp) restates the reference's
``pamo_safe_project.process`` (simp_cuda/safe_project/src/pamo_safe_project/
processing.py:23-108) and its ``Stage3System`` (system.py:17-470).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libpamo_stage3_oracle.so")
_lib = None

# BlockTypes (defs.py:1-9)
PT_CONTACT = 3
EE_CONTACT = 4


class Config:
    """The default configuration of the reference (config.py:6-73)."""

    def __init__(self):
        self.n_newton_iters = 10
        self.n_cg_iters = 40
        self.n_line_search_iters = 10
        self.spd_max_iters = 3
        self.ccd_max_iters = 100
        self.max_blocks = 1 << 25
        self.d_hat = 1e-3
        self.coll_stiffness = 1e2
        self.contact_detection_radius = 1e-2
        self.ee_classify_thres = 1e-3
        self.ccd_slackness = 0.7
        self.ccd_thickness = 1e-6
        self.elas_young_modulus = 1e-1
        self.elas_poisson_ratio = 0.0
        self.hinge_stiffness = 1e-2
        self.mesh2gt_dist_stiffness = 1e3
        self.min_angle_thres = 0.0

    def arrays(self):
        d = np.array([self.d_hat, self.coll_stiffness, self.contact_detection_radius, self.ee_classify_thres,
                      self.ccd_slackness, self.ccd_thickness, self.elas_young_modulus, self.elas_poisson_ratio,
                      self.hinge_stiffness, self.mesh2gt_dist_stiffness, self.min_angle_thres], np.float64)
        i = np.array([self.n_newton_iters, self.n_cg_iters, self.n_line_search_iters, self.spd_max_iters,
                      self.ccd_max_iters, self.max_blocks], np.int64)
        return d, i


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            build()
        L = ctypes.CDLL(_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.s3o_pt.argtypes = [P, P, P, P]
        L.s3o_ee.argtypes = [P, D, P, P, P]
        L.s3o_closest.argtypes = [P, P, P]
        L.s3o_barrier.argtypes = [D, D, P]
        L.s3o_elastic.argtypes = [P, P, P, D, D, P, P, P]
        L.s3o_hinge.argtypes = [P, P, D, P, P, P, P, P, I]
        L.s3o_spd_project.argtypes = [P, I]
        L.s3o_accd.argtypes = [P, P, I, D, D, I, D]
        L.s3o_accd.restype = D
        L.s3o_voronoi.argtypes = [P, I, P, I, P]
        L.s3o_create.argtypes = [P, P]
        L.s3o_create.restype = P
        L.s3o_destroy.argtypes = [P]
        L.s3o_register.argtypes = [P, P, I, P, I, P, I, P, I]
        L.s3o_sizes.argtypes = [P, P]
        L.s3o_step.argtypes = [P, I]
        L.s3o_update_target.argtypes = [P]
        L.s3o_newton_part.argtypes = [P, I]
        L.s3o_get.argtypes = [P, I, P]
        L.s3o_set.argtypes = [P, I, P]
        L.s3o_scalars.argtypes = [P, P]
        L.s3o_energy_parts.argtypes = [P, P, P]
        L.s3o_contacts.argtypes = [P, P, P]
        L.s3o_hess_dx.argtypes = [P, P, P]
        L.s3o_edges.argtypes = [P, P]
        L.s3o_hinges.argtypes = [P, P, P]
        L.s3o_voronoi_of.argtypes = [P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _d(a, shape=None):
    a = np.ascontiguousarray(a, np.float64)
    return a if shape is None else a.reshape(shape)


# ------------------------------------------------------------ functions
def pt(x):
    """Point-triangle pair (x0; x1, x2, x3): (type, distance, gradient [4,3])."""
    x = _d(x, (4, 3))
    t, d, g = ctypes.c_int(), ctypes.c_double(), np.zeros((4, 3))
    lib().s3o_pt(_p(x), ctypes.byref(t), ctypes.byref(d), _p(g))
    return t.value, d.value, g


def ee(x, thres=1e-3):
    """Edge-edge pair (x0, x1; x2, x3): (type, distance, gradient [4,3])."""
    x = _d(x, (4, 3))
    t, d, g = ctypes.c_int(), ctypes.c_double(), np.zeros((4, 3))
    lib().s3o_ee(_p(x), float(thres), ctypes.byref(t), ctypes.byref(d), _p(g))
    return t.value, d.value, g


def closest(p, tri):
    out = np.zeros(3)
    lib().s3o_closest(_p(_d(p, (3,))), _p(_d(tri, (3, 3))), _p(out))
    return out


def barrier(d, d_hat):
    out = np.zeros(3)
    lib().s3o_barrier(float(d), float(d_hat), _p(out))
    return out


def elastic(rest, x, dx, mu=0.05, la=0.0):
    """(energy, +dE/dx [3,3], H dx [3,3]) of one membrane triangle."""
    E, g, h = ctypes.c_double(), np.zeros((3, 3)), np.zeros((3, 3))
    lib().s3o_elastic(_p(_d(rest, (3, 3))), _p(_d(x, (3, 3))), _p(_d(dx, (3, 3))), float(mu), float(la),
                      ctypes.byref(E), _p(g), _p(h))
    return E.value, g, h


def hinge(rest, x, k=1e-2, spd_iters=3):
    """(angle, energy, +dE/dx [4,3], H [12,12], projected H [12,12])."""
    th, E = ctypes.c_double(), ctypes.c_double()
    g, H, Hp = np.zeros((4, 3)), np.zeros((12, 12)), np.zeros((12, 12))
    lib().s3o_hinge(_p(_d(rest, (4, 3))), _p(_d(x, (4, 3))), float(k), ctypes.byref(th), ctypes.byref(E), _p(g),
                    _p(H), _p(Hp), int(spd_iters))
    return th.value, E.value, g, H, Hp


def spd_project(H, it=3):
    H = np.array(H, np.float64).reshape(12, 12).copy()
    lib().s3o_spd_project(_p(H), int(it))
    return H


def accd(x, v, type1, slack=0.7, thick=1e-6, iters=100, ee_thres=1e-3):
    return lib().s3o_accd(_p(_d(x, (4, 3))), _p(_d(v, (4, 3))), int(type1), float(slack), float(thick), int(iters),
                          float(ee_thres))


def voronoi(V, F):
    V = _d(V)
    F = np.ascontiguousarray(F, np.int32)
    out = np.zeros(V.shape[0])
    lib().s3o_voronoi(_p(V), V.shape[0], _p(F), F.shape[0], _p(out))
    return out


# --------------------------------------------------------------- system
def normalization(gt_V):
    """processing.py:15-20: scale by the largest extent, centre at the mean."""
    gt_V = np.asarray(gt_V, np.float64)
    center = np.mean(gt_V, axis=0)
    extents = np.max(gt_V, axis=0) - np.min(gt_V, axis=0)
    scale = 1.0 / np.max(extents)
    return scale, -center * scale


class System:
    """Stage3System over the oracle (register, step, parts of a Newton step)."""

    def __init__(self, config=None):
        self.config = config or Config()
        d, i = self.config.arrays()
        self._h = lib().s3o_create(_p(d), _p(i))
        self.nV = 0

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.s3o_destroy(self._h)
            self._h = None

    def register_mesh(self, gt_V, gt_F, V, F):
        gv, v = _d(gt_V), _d(V)
        gf, f = np.ascontiguousarray(gt_F, np.int32), np.ascontiguousarray(F, np.int32)
        rc = lib().s3o_register(self._h, _p(gv), gv.shape[0], _p(gf), gf.shape[0], _p(v), v.shape[0], _p(f),
                                f.shape[0])
        if rc != 0:
            raise ValueError("number of hinges differs from the number of edges (mesh not closed and manifold)")
        self.nV = v.shape[0]

    def sizes(self):
        out = np.zeros(6, np.int64)
        lib().s3o_sizes(self._h, _p(out))
        return dict(zip(["nV", "nF", "nE", "hinges", "contacts", "do_nothing"], out.tolist()))

    def step(self, n=1):
        lib().s3o_step(self._h, int(n))

    def update_target(self):
        lib().s3o_update_target(self._h)

    PARTS = {"detect": 0, "energy": 1, "diff": 2, "cg": 3, "clamp": 4, "ccd": 5, "line_search": 6, "newton": 7}
    VECS = {"q": 0, "target": 1, "grad": 2, "diag": 3, "p": 4, "q_rest": 5, "q_newton": 6, "q_det": 7}

    def part(self, name):
        lib().s3o_newton_part(self._h, self.PARTS[name])

    def get(self, name):
        out = np.zeros((self.nV, 3))
        lib().s3o_get(self._h, self.VECS[name], _p(out))
        return out

    def set(self, name, a):
        lib().s3o_set(self._h, self.VECS[name], _p(_d(a, (self.nV, 3))))

    def scalars(self):
        out = np.zeros(5)
        lib().s3o_scalars(self._h, _p(out))
        return dict(zip(["energy", "energy_prev", "ccd_step", "radius", "ls_halvings"], out.tolist()))

    def energy_parts(self, x):
        out = np.zeros(4)
        lib().s3o_energy_parts(self._h, _p(_d(x, (self.nV, 3))), _p(out))
        return dict(zip(["m2gt", "elastic", "hinge", "collision"], out.tolist()))

    def contacts(self):
        n = self.sizes()["contacts"]
        rows = np.zeros((n, 6), np.int32)
        d = np.zeros(n)
        lib().s3o_contacts(self._h, _p(rows), _p(d))
        return rows, d

    def hess_dx(self, dx):
        out = np.zeros((self.nV, 3))
        lib().s3o_hess_dx(self._h, _p(_d(dx, (self.nV, 3))), _p(out))
        return out

    def edges(self):
        out = np.zeros((self.sizes()["nE"], 2), np.int32)
        lib().s3o_edges(self._h, _p(out))
        return out

    def hinges(self):
        n = self.sizes()["hinges"]
        idx, rest = np.zeros((n, 4), np.int32), np.zeros((n, 2))
        lib().s3o_hinges(self._h, _p(idx), _p(rest))
        return idx, rest

    def voronoi_areas(self):
        out = np.zeros(self.nV)
        lib().s3o_voronoi_of(self._h, _p(out))
        return out

    def get_vertices(self):
        return self.get("q")


def process(gt_V, gt_F, stage2_V, stage2_F, n_iters, config=None):
    """pamo_safe_project.process (processing.py:23-108) on the oracle."""
    s, t = normalization(gt_V)
    gv = np.asarray(gt_V, np.float64) * s + t
    v = np.asarray(stage2_V, np.float64) * s + t
    sysm = System(config)
    sysm.register_mesh(gv, gt_F, v, stage2_F)
    sysm.step(n_iters)
    return (sysm.get_vertices() - t) / s, np.asarray(stage2_F)
