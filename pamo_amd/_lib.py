"""ctypes binding of libpamo_hip.so (the C ABI declared in include/pamo_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``python -m pamo_amd.build``).  Loading fails loudly when it is missing:
there is no CPU fallback anywhere in the package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpamo_hip.so")

SHUFFLE_IDENTITY = 0
SHUFFLE_FEISTEL = 1

# kernel ids of pamo_profile (include/pamo_hip.h)
K_QUADRIC, K_EDGE_COST, K_CLAIM, K_LBVH, K_TRAVERSE = range(5)
KERNEL_NAMES = {K_QUADRIC: "k_quadric", K_EDGE_COST: "k_edge_cost", K_CLAIM: "k_claim", K_LBVH: "lbvh_build",
                K_TRAVERSE: "k_traverse"}

_lib = None


class PamoError(RuntimeError):
    pass


def lib():
    """The loaded library (raises PamoError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PamoError("libpamo_hip.so not found at %s: build it with __graft_entry__.build() "
                        "(hipcc --offload-arch=gfx950)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    LL = ctypes.c_longlong
    F = ctypes.c_float
    L.pamo_ctx_create.argtypes = [ctypes.POINTER(P), I, I]
    L.pamo_ctx_create.restype = I
    L.pamo_ctx_destroy.argtypes = [P]
    L.pamo_ctx_destroy.restype = None
    L.pamo_last_error.argtypes = [P]
    L.pamo_last_error.restype = ctypes.c_char_p
    L.pamo_device_bytes.argtypes = [P]
    L.pamo_device_bytes.restype = ctypes.c_longlong
    L.pamo_free_forward.argtypes = [P, P, P, P, P, I, I, I, F, F, I, I]
    L.pamo_free_forward.restype = I
    L.pamo_sp_forward.argtypes = [P, P, P, P, I, I, F, F, I]
    L.pamo_sp_forward.restype = I
    L.pamo_free_outputs.argtypes = [P] + [P] * 5 + [ctypes.POINTER(I)] * 3
    L.pamo_free_outputs.restype = I
    L.pamo_free_copy_outputs.argtypes = [P, P, P, P, P, P, P]
    L.pamo_free_copy_outputs.restype = I
    L.pamo_run_stage2.argtypes = [P, P, P, P, I, I, F, LL, F, I, LL, ctypes.POINTER(I), ctypes.POINTER(I), P]
    L.pamo_run_stage2.restype = I
    L.pamo_run_stage2_noalloc.argtypes = [P, P, P, P, I, I, F, LL, F, I, LL, P, P, ctypes.POINTER(I),
                                          ctypes.POINTER(I), P]
    L.pamo_run_stage2_noalloc.restype = I
    L.pamo_run_rounds.argtypes = [P, P, I]
    L.pamo_run_rounds.restype = I
    L.pamo_run_fetch.argtypes = [P, P, P, P]
    L.pamo_run_fetch.restype = I
    L.pamo_debug_copy.argtypes = [P, P, I, P, LL]
    L.pamo_debug_copy.restype = LL
    L.pamo_debug_pairs.argtypes = [P, P, P, P, I, I, P, LL]
    L.pamo_debug_pairs.restype = LL
    L.pamo_profile.argtypes = [P, ctypes.c_uint]
    L.pamo_profile.restype = I
    L.pamo_profile_read.argtypes = [P, I, P]
    L.pamo_profile_read.restype = I
    L.pamo_debug_pair_cap.argtypes = [P, LL]
    L.pamo_debug_pair_cap.restype = I
    L.pamo_debug_guard.argtypes = [P, I]
    L.pamo_debug_walk_split.argtypes = [P, I]
    L.pamo_debug_walk_split.restype = I
    L.pamo_set_strict.argtypes = [P, I]
    L.pamo_set_tuning.argtypes = [P, ctypes.c_char_p, LL]
    L.pamo_set_tuning.restype = I
    L.pamo_get_tuning.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(LL)]
    L.pamo_get_tuning.restype = I
    L.pamo_set_placement.argtypes = [P, I]
    L.pamo_set_placement.restype = I
    L.pamo_debug_guard.restype = I
    L.pamo_simplify.argtypes = [P, P, P, P, P, F]
    L.pamo_simplify.restype = None
    L.pamo_prepare_host.argtypes = [P, P, I, I, P, P, P]
    L.pamo_prepare_host.restype = I
    _lib = L
    return L


# launch shapes of the engine (pamo_set_tuning); python -m pamo_amd.tune
# writes the measured choice to TUNING_FILE, which every Context applies
TUNING_NAMES = ("ec_lanes", "morton_bits", "bf_max_faces", "bf_split_max_faces", "walk8_max", "walk4_max",
                "emit_groups", "pipeline", "graph", "graph_max_faces", "s3_bf_max_faces", "s3_cg_one_block",
                "s3_graph")
TUNING_FILE = os.path.join(_HERE, "tuning.json")


def tuning_cache():
    """The tuner's cache (name -> value), or {} when there is none.  The
    environment variable PAMO_TUNING_FILE, read by the library itself, takes
    precedence over this file."""
    import json
    if os.environ.get("PAMO_TUNING_FILE") or not os.path.exists(TUNING_FILE):
        return {}
    with open(TUNING_FILE) as fp:
        d = json.load(fp)
    return {k: int(v) for k, v in d.items() if k in TUNING_NAMES}


class Context:
    """Owns one ``pamo_ctx`` (one simplifier's device state)."""

    def __init__(self, device=0, shuffle=SHUFFLE_FEISTEL, tuning=True):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.pamo_ctx_create(ctypes.byref(h), int(device), int(shuffle))
        if rc != 0:
            raise PamoError("pamo_ctx_create failed (%d): no usable GPU %d" % (rc, device))
        self.handle = h
        self.device = device
        if tuning:
            for k, v in tuning_cache().items():
                self.set_tuning(k, v)

    def set_tuning(self, name, value):
        """One launch-shape parameter of the engine (results do not change)."""
        self.check(self._L.pamo_set_tuning(self.handle, name.encode(), int(value)), "pamo_set_tuning(%s)" % name)

    def tuning(self):
        """Every launch-shape parameter in effect."""
        out = {}
        for k in TUNING_NAMES:
            v = ctypes.c_longlong()
            self.check(self._L.pamo_get_tuning(self.handle, k.encode(), ctypes.byref(v)), "pamo_get_tuning")
            out[k] = v.value
        return out

    def close(self):
        if getattr(self, "handle", None):
            self._L.pamo_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self):
        """Device bytes this context holds (its buffers, the batched
        engine's and the stage-3 system's)."""
        return int(self._L.pamo_device_bytes(self.handle))

    def error(self):
        msg = self._L.pamo_last_error(self.handle)
        return msg.decode() if msg else ""

    def profile(self, kernels):
        mask = 0
        for k in kernels:
            mask |= 1 << k
        self.check(self._L.pamo_profile(self.handle, mask), "pamo_profile")

    def profile_read(self, kernel):
        import numpy as np
        out = np.zeros(3, np.float64)
        self.check(self._L.pamo_profile_read(self.handle, int(kernel), out.ctypes.data), "pamo_profile_read")
        return dict(launches=int(out[0]), ms=float(out[1]), bytes=float(out[2]))

    def debug_pair_cap(self, cap):
        """Limit the engine's candidate-pair buffer (-1: default); test hook."""
        self.check(self._L.pamo_debug_pair_cap(self.handle, int(cap)), "pamo_debug_pair_cap")

    def debug_guard(self, mode):
        """Candidate search of the engine's check: 0 by size, 1 LBVH, 2 exhaustive; test hook."""
        self.check(self._L.pamo_debug_guard(self.handle, int(mode)), "pamo_debug_guard")

    def debug_walk_split(self, kshift):
        """Lanes per query of the engine's LBVH walk, log2 (-1: by size); test hook."""
        self.check(self._L.pamo_debug_walk_split(self.handle, int(kshift)), "pamo_debug_walk_split")

    def set_strict(self, strict):
        """Strict guard for the engine (off by default; not in the reference):
        exact pair test added, every collapse scanned, undo until clean."""
        self.check(self._L.pamo_set_strict(self.handle, 1 if strict else 0), "pamo_set_strict")

    PLACEMENTS = {"midpoint": 0, "optimal": 1}

    def set_placement(self, placement):
        """Where the engine places a collapse: "midpoint" (the reference's,
        the default) or "optimal" (the optimal vertex of Q_u + Q_v with FP64
        and midpoint/endpoint fallbacks; not in the reference)."""
        mode = self.PLACEMENTS[placement] if isinstance(placement, str) else int(placement)
        self.check(self._L.pamo_set_placement(self.handle, mode), "pamo_set_placement")

    def check(self, rc, what):
        if rc < 0:
            raise PamoError("%s failed (%d): %s" % (what, rc, self.error()))
        return rc
