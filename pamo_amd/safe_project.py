"""Stage 3 (safe projection) on the GPU: the reference's ``pamo_safe_project``
package (simp_cuda/safe_project/src/pamo_safe_project/), in float64.

``process(gt_V, gt_F, stage2_V, stage2_F, n_iters, system=None, config=None)``
is ``pamo_safe_project.process`` (processing.py:23-108): the target (the input
mesh) and the stage-2 mesh are normalised by the target's extent and mean,
the stage-2 vertices are moved by ``n_iters`` steps of a Newton solver, and
the result is mapped back.  Each step (system.py:407-436) finds every vertex's
nearest target point, then runs Newton iterations on the sum of four
energies: distance to the target points (weighted by Voronoi area), a
membrane (St. Venant-Kirchhoff on the rest triangles), hinge bending (rest
dihedral angles) and an IPC log barrier on point-triangle and edge-edge
contacts within d_hat.  Each Newton iteration detects contacts, solves the
Jacobi-preconditioned CG system, limits the step to the detection radius,
shortens it by additive CCD so no contact closes, and halves it until the
energy drops.

Everything runs in libpamo_hip.so (the ``pamo_s3_*`` calls of
include/pamo_hip.h); positions never leave the device between steps.
"""
import ctypes

import numpy as np
import torch

from . import _lib

# vectors of pamo_s3_get / pamo_s3_set
VECS = {"q": 0, "target": 1, "grad": 2, "diag": 3, "p": 4, "q_rest": 5, "q_newton": 6, "q_det": 7}
# parts of a Newton iteration (pamo_s3_newton_part)
PARTS = {"detect": 0, "energy": 1, "diff": 2, "cg": 3, "clamp": 4, "ccd": 5, "line_search": 6, "newton": 7}
# contact types (defs.py:1-9)
PT_CONTACT = 3
EE_CONTACT = 4


def _bind(L):
    if getattr(L, "_pamo_s3_bound", False):
        return L
    P, I, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
    L.pamo_s3_configure.argtypes = [P, P, P]
    L.pamo_s3_register.argtypes = [P, P, P, I, P, I, P, I, P, I]
    L.pamo_s3_step.argtypes = [P, P, I]
    L.pamo_s3_sizes.argtypes = [P, P]
    L.pamo_s3_get.argtypes = [P, P, I, P]
    L.pamo_s3_set.argtypes = [P, P, I, P]
    L.pamo_s3_scalars.argtypes = [P, P, P]
    L.pamo_s3_update_target.argtypes = [P, P]
    L.pamo_s3_newton_part.argtypes = [P, P, I]
    L.pamo_s3_energy_parts.argtypes = [P, P, P, P]
    L.pamo_s3_contacts.argtypes = [P, P, P, P]
    L.pamo_s3_contacts.restype = LL
    L.pamo_s3_hess_dx.argtypes = [P, P, P, P]
    L.pamo_s3_topology.argtypes = [P, P, P, P, P, P]
    for n in ["configure", "register", "step", "sizes", "get", "set", "scalars", "update_target", "newton_part",
              "energy_parts", "hess_dx", "topology"]:
        getattr(L, "pamo_s3_" + n).restype = I
    L._pamo_s3_bound = True
    return L


class Stage3Config:
    """config.py:6-73, the parameters the default energies read."""

    def __init__(self):
        self.max_blocks = 1 << 25
        self.n_newton_iters = 10
        self.n_cg_iters = 40
        self.n_line_search_iters = 10
        self.d_hat = 1e-3
        self.coll_stiffness = 1e2
        self.contact_detection_radius = 1e-2
        self.ee_classify_thres = 1e-3
        self.ccd_slackness = 0.7
        self.ccd_thickness = 1e-6
        self.ccd_max_iters = 100
        self.spd_max_iters = 3
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


def _dev(device):
    d = torch.device(device)
    if d.type != "cuda":
        raise RuntimeError("the stage-3 system runs on a CUDA (HIP) device")
    return d.index if d.index is not None else torch.cuda.current_device()


def _stream():
    return torch.cuda.current_stream().cuda_stream


class Stage3System:
    """system.py:17-470 over one library context: ``register_mesh``,
    ``step``, ``get_vertices`` and ``clear``, plus the parts of a step for
    parity tests."""

    def __init__(self, config=None, device="cuda:0", ctx=None):
        self.config = config or Stage3Config()
        self.device = torch.device("cuda", _dev(device))
        self.ctx = ctx if ctx is not None else _lib.Context(self.device.index)
        self._L = _bind(self.ctx._L)
        self._configure()
        self.n_particles = 0

    def _configure(self):
        d, i = self.config.arrays()
        self.ctx.check(self._L.pamo_s3_configure(self.ctx.handle, d.ctypes.data, i.ctypes.data), "pamo_s3_configure")

    def clear(self):
        """system.py:99-107: forget the mesh; the configuration is re-read."""
        self._configure()
        self.n_particles = 0

    def _d(self, a, dtype):
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=dtype).contiguous()
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64 if dtype == torch.float64 else np.int32)).to(
            self.device)

    def register_mesh(self, gt_V, gt_F, V, F):
        gv, gf = self._d(gt_V, torch.float64), self._d(gt_F, torch.int32)
        v, f = self._d(V, torch.float64), self._d(F, torch.int32)
        rc = self._L.pamo_s3_register(self.ctx.handle, _stream(), gv.data_ptr(), gv.shape[0], gf.data_ptr(),
                                      gf.shape[0], v.data_ptr(), v.shape[0], f.data_ptr(), f.shape[0])
        if rc == -2:
            raise ValueError("number of hinges differs from the number of edges: " + self.ctx.error())
        self.ctx.check(rc, "pamo_s3_register")
        self.n_particles = v.shape[0]

    def sizes(self):
        out = np.zeros(6, np.int64)
        self.ctx.check(self._L.pamo_s3_sizes(self.ctx.handle, out.ctypes.data), "pamo_s3_sizes")
        return dict(zip(["nV", "nF", "nE", "hinges", "contacts", "do_nothing"], out.tolist()))

    @property
    def do_nothing(self):
        return bool(self.sizes()["do_nothing"])

    def step(self, n=1):
        self.ctx.check(self._L.pamo_s3_step(self.ctx.handle, _stream(), int(n)), "pamo_s3_step")

    def get(self, name):
        """A [nV, 3] float64 device tensor of the named vector (VECS)."""
        out = torch.empty((self.n_particles, 3), dtype=torch.float64, device=self.device)
        self.ctx.check(self._L.pamo_s3_get(self.ctx.handle, _stream(), VECS[name], out.data_ptr()), "pamo_s3_get")
        return out

    def set(self, name, a):
        a = self._d(a, torch.float64).reshape(self.n_particles, 3)
        self.ctx.check(self._L.pamo_s3_set(self.ctx.handle, _stream(), VECS[name], a.data_ptr()), "pamo_s3_set")

    def get_vertices(self):
        """system.py:438-440: the positions as a float64 numpy array."""
        return self.get("q").cpu().numpy()

    # -------------------------------------------------- parts (parity)
    def update_target(self):
        self.ctx.check(self._L.pamo_s3_update_target(self.ctx.handle, _stream()), "pamo_s3_update_target")

    def part(self, name):
        self.ctx.check(self._L.pamo_s3_newton_part(self.ctx.handle, _stream(), PARTS[name]), "pamo_s3_newton_part")

    def scalars(self):
        out = np.zeros(5)
        self.ctx.check(self._L.pamo_s3_scalars(self.ctx.handle, _stream(), out.ctypes.data), "pamo_s3_scalars")
        return dict(zip(["energy", "energy_prev", "ccd_step", "radius", "ls_halvings"], out.tolist()))

    def energy_parts(self, x):
        x = self._d(x, torch.float64).reshape(self.n_particles, 3)
        out = np.zeros(4)
        self.ctx.check(self._L.pamo_s3_energy_parts(self.ctx.handle, _stream(), x.data_ptr(), out.ctypes.data),
                       "pamo_s3_energy_parts")
        return dict(zip(["m2gt", "elastic", "hinge", "collision"], out.tolist()))

    def contacts(self):
        """(rows [n, 6] int32: type, i0, i1, i2, i3, distance type; distances [n])."""
        n = self.ctx.check(self._L.pamo_s3_contacts(self.ctx.handle, _stream(), None, None), "pamo_s3_contacts")
        rows = torch.empty((max(n, 1), 6), dtype=torch.int32, device=self.device)
        d = torch.empty(max(n, 1), dtype=torch.float64, device=self.device)
        self.ctx.check(self._L.pamo_s3_contacts(self.ctx.handle, _stream(), rows.data_ptr(), d.data_ptr()),
                       "pamo_s3_contacts")
        return rows[:n].cpu().numpy(), d[:n].cpu().numpy()

    def hess_dx(self, dx):
        dx = self._d(dx, torch.float64).reshape(self.n_particles, 3)
        out = torch.empty_like(dx)
        self.ctx.check(self._L.pamo_s3_hess_dx(self.ctx.handle, _stream(), dx.data_ptr(), out.data_ptr()),
                       "pamo_s3_hess_dx")
        return out.cpu().numpy()

    def topology(self):
        """(edges [nE, 2], hinges [nE, 4], rest (angle, length) [nE, 2], Voronoi areas [nV])."""
        s = self.sizes()
        e = torch.empty((max(s["nE"], 1), 2), dtype=torch.int32, device=self.device)
        h = torch.empty((max(s["hinges"], 1), 4), dtype=torch.int32, device=self.device)
        r = torch.empty((max(s["hinges"], 1), 2), dtype=torch.float64, device=self.device)
        a = torch.empty(max(s["nV"], 1), dtype=torch.float64, device=self.device)
        self.ctx.check(self._L.pamo_s3_topology(self.ctx.handle, _stream(), e.data_ptr(), h.data_ptr(), r.data_ptr(),
                                                a.data_ptr()), "pamo_s3_topology")
        return (e[:s["nE"]].cpu().numpy(), h[:s["hinges"]].cpu().numpy(), r[:s["hinges"]].cpu().numpy(),
                a[:s["nV"]].cpu().numpy())


def get_normalization_transform(V):
    """processing.py:15-20: scale by the largest extent, centre at the mean."""
    V = np.asarray(V, np.float64)
    center = np.mean(V, axis=0)
    extents = np.max(V, axis=0) - np.min(V, axis=0)
    scale = 1.0 / np.max(extents)
    return scale, -center * scale


def process(gt_V, gt_F, stage2_V, stage2_F, n_iters, system=None, config=None):
    """processing.py:23-108: (stage3_V float64 [V, 3], stage3_F)."""
    if system is None:
        system = Stage3System(config)
    else:
        system.clear()
    s, t = get_normalization_transform(gt_V)
    gv = np.asarray(gt_V, np.float64) * s + t
    v = np.asarray(stage2_V, np.float64) * s + t
    system.register_mesh(gv, gt_F, v, stage2_F)
    system.step(n_iters)
    return (system.get_vertices() - t) / s, stage2_F
