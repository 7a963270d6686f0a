"""HipSolver: the ClothSolver backend (src/backend/solver.ts:5-13) on one
MI355X, over the C ABI of include/cloth_hip.h.

Mirrors CpuSolver (src/backend/cpu-solver.ts:7-49): `initialize` keeps the
caller's scene by reference, `step` advances it and writes the new state back
into `scene.positions` / `scene.velocities` in place, so `getPositions()`
returns the same live array every tick.  Errors raise like the reference's
thrown Errors.  There is no CPU fallback: without the HIP library or a device
every call fails.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from .scene import ClothMaterial, ClothScene, to_desc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libclothhip.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)


class ClothError(RuntimeError):
    """A non-zero status from the C ABI (message from cs_last_error)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def load_library(path=LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"HIP cloth library not built: {path} (run __graft_entry__.build())")
    L = C.CDLL(path)
    sig = {
        "cs_options_default": (None, [C.POINTER(_abi.CsOptions)]),
        "cs_create": (C.c_int, [C.POINTER(_abi.CsSceneDesc), C.POINTER(_abi.CsOptions), C.POINTER(C.c_void_p)]),
        "cs_destroy": (None, [C.c_void_p]),
        "cs_step": (C.c_int, [C.c_void_p, C.c_double]),
        "cs_set_material": (C.c_int, [C.c_void_p, C.POINTER(_abi.CsMaterial)]),
        "cs_pin_vertex": (C.c_int, [C.c_void_p, C.c_uint32, _dp]),
        "cs_set_static_mesh": (C.c_int, [C.c_void_p, _fp, C.c_uint32, _up, C.c_uint32]),
        "cs_unpin_vertex": (C.c_int, [C.c_void_p, C.c_uint32]),
        "cs_get_positions": (C.c_int, [C.c_void_p, _dp, C.c_size_t]),
        "cs_get_velocities": (C.c_int, [C.c_void_p, _dp, C.c_size_t]),
        "cs_set_state": (C.c_int, [C.c_void_p, _dp, _dp, C.c_size_t]),
        "cs_last_stats": (C.c_int, [C.c_void_p, C.POINTER(_abi.CsStepStats)]),
        "cs_last_error": (C.c_char_p, []),
        "cs_device_info": (C.c_int, [C.c_int, C.c_char_p, C.c_size_t]),
        "cs_abi_version": (C.c_int, []),
        "cs_kernel_timing": (C.c_int, [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_uint64), _dp]),
        "cs_reset_timing": (C.c_int, [C.c_void_p]),
        "cs_time_hvp": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
        "cs_schwarz_partition": (C.c_int, [C.POINTER(_abi.CsSceneDesc), C.c_int, C.c_int, C.c_int,
                                           C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "cs_hvp_bytes": (C.c_int, [C.c_void_p, _dp, _dp]),
        "cs_tuned_launch": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
        "cs_batch_step": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_double, C.c_int, C.POINTER(C.c_int32)]),
        "cs_batch_step_packed": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_double, C.POINTER(C.c_int32)]),
        "cs_step_adjoint": (C.c_int, [C.c_void_p, C.c_double, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int, _dp,
                                      _dp, _dp, C.POINTER(C.c_int), _dp]),
        "cs_probe_hvp": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, _dp]),
        "cs_probe_hvp_contact": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_double, _dp, C.POINTER(C.c_uint32)]),
        "cs_coarse_pattern": (C.c_int, [C.c_void_p, _up, _up, C.c_uint64, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint32)]),
        "cs_probe_two_level": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "cs_hvp_tiles": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), _up, _up, _up, C.POINTER(C.c_uint16),
                                   C.POINTER(C.c_uint16)]),
        "cs_hvp_layout": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
        "cs_probe_schwarz": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_uint8),
                                       C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)]),
        "cs_probe_active_set": (C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(C.c_int32), _dp, C.c_uint64,
                                          C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "cs_probe_contact_adjoint": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
        "cs_probe_check_trial": (C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(C.c_int32), _dp, _dp]),
        "cs_probe_candidates": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int32, _up, C.c_uint64,
                                          C.POINTER(C.c_uint64), _up, C.c_uint64, C.POINTER(C.c_uint64)]),
        "cs_probe_ccd": (C.c_int, [C.c_int, _dp, _dp, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.c_uint64, C.c_double, _dp]),
        "cs_probe_closest": (C.c_int, [C.c_int, _dp, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                       C.c_uint64, _dp]),
        "cs_device_alloc": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
        "cs_device_free": (C.c_int, [C.c_int, C.c_void_p]),
        "cs_device_copy": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
        "cs_probe_hinges": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_uint64, _dp, _dp]),
        "cs_ccd_queries": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64,
                                     C.c_double, C.c_void_p, _dp]),
        "cs_broadphase_self_pairs": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                               C.c_uint32, C.c_double, C.c_void_p, C.c_uint64,
                                               C.POINTER(C.c_uint64), _dp, _dp]),
        "cs_broadphase_self_pairs_noalloc": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                                       C.c_uint32, C.c_double, C.c_void_p, C.c_uint64,
                                                       C.POINTER(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != _abi.CS_OK:
        raise ClothError(rc, load_library().cs_last_error().decode(errors="replace"))


def device_info(device=0):
    buf = C.create_string_buffer(256)
    _check(load_library().cs_device_info(device, buf, 256))
    return buf.value.decode()


def schwarz_partition(scene, target=16, min_size=8, max_size=32):
    """buildSchwarzDomains of the library (host only): (domain_of, local_of, count)."""
    desc, keep = to_desc(scene)
    n = scene.mesh.count
    dom = np.zeros(n, np.uint32)
    loc = np.zeros(n, np.uint32)
    cnt = C.c_uint32()
    _check(load_library().cs_schwarz_partition(C.byref(desc), target, min_size, max_size,
                                               dom.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               loc.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(cnt)))
    del keep
    return dom, loc, cnt.value


def broadphase_self_pairs(x0, x1, idx, pad, device=0, cap=None, timing=False):
    """BroadPhase.build (broadphase.ts:61-64) of the cloth's self pairs on the
    device: sorted (tA, tB) pairs as an (k, 2) uint32 array; with timing, also
    (build_ms, query_ms).  The buffer grows to the reported count on overflow."""
    from .devmem import DeviceArray
    L = load_library()
    x0 = np.ascontiguousarray(x0, np.float64)
    x1 = np.ascontiguousarray(x1, np.float64)
    idx = np.ascontiguousarray(idx, np.uint32)
    n, m = len(x0) // 3, len(idx) // 3
    d0, d1, di = DeviceArray(x0, device=device), DeviceArray(x1, device=device), DeviceArray(idx, device=device)
    cap = max(1, 8 * m if cap is None else int(cap))
    cnt = C.c_uint64()
    bms, qms = C.c_double(), C.c_double()
    while True:
        out = DeviceArray(dtype=np.uint32, shape=(cap, 2), device=device)
        rc = L.cs_broadphase_self_pairs(device, d0.ptr, d1.ptr, n, di.ptr, m, float(pad), out.ptr, cap,
                                        C.byref(cnt), C.byref(bms) if timing else None,
                                        C.byref(qms) if timing else None)
        if rc == _abi.CS_E_OVERFLOW and cnt.value > cap:
            cap = int(cnt.value)
            out.free()
            continue
        _check(rc)
        pairs = out.numpy(cnt.value)
        out.free()
        return (pairs, bms.value, qms.value) if timing else pairs


def probe_hinges(X, P, rest, length, asum, k, device=0):
    """Analytic gradient and Hessian-vector product of hingeEnergy on the
    device for independent hinges: X, P (count, 12) -> (grad, hvp) (count, 12)."""
    L = load_library()
    X = np.ascontiguousarray(X, np.float64).reshape(-1, 12)
    P = np.ascontiguousarray(P, np.float64).reshape(-1, 12)
    r, ln, a = (np.ascontiguousarray(np.broadcast_to(v, (len(X),)), np.float64) for v in (rest, length, asum))
    g = np.zeros_like(X)
    hv = np.zeros_like(X)
    _check(L.cs_probe_hinges(device, X.ctypes.data_as(_dp), P.ctypes.data_as(_dp), r.ctypes.data_as(_dp),
                             ln.ctypes.data_as(_dp), a.ctypes.data_as(_dp), float(k), len(X), g.ctypes.data_as(_dp),
                             hv.ctypes.data_as(_dp)))
    return g, hv


def expand_pairs(pairs, idx):
    """expandTriPair (self-collision.ts:23-41) of sorted triangle pairs as
    CCD primitives: per pair the 3 + 3 vertex-triangle prims (each vertex of
    one triangle against the other) and the 3 x 3 edge-edge prims; returns
    (kinds int32, ids4 int32 (k, 4))."""
    T = np.asarray(idx, np.int64).reshape(-1, 3)
    A, B = T[pairs[:, 0]], T[pairs[:, 1]]
    vt = [np.stack([P[:, c], Q[:, 0], Q[:, 1], Q[:, 2]], 1) for P, Q in ((A, B), (B, A)) for c in range(3)]
    ea = [(A[:, 0], A[:, 1]), (A[:, 1], A[:, 2]), (A[:, 2], A[:, 0])]
    eb = [(B[:, 0], B[:, 1]), (B[:, 1], B[:, 2]), (B[:, 2], B[:, 0])]
    ee = [np.stack([a, b, c, d], 1) for a, b in ea for c, d in eb]
    ids = np.concatenate(vt + ee).astype(np.int32)
    kinds = np.concatenate([np.zeros(6 * len(pairs), np.int32), np.ones(9 * len(pairs), np.int32)])
    return kinds, np.ascontiguousarray(ids)


def hvp_tiles(scene):
    """The vertex-owner tile layout of the mixed HVP for a scene (cs_hvp_tiles,
    host only): dict of hdr (tiles x 8), loc, lidx (tile triangles x 2), off,
    ent; None when the mesh is not tiled."""
    L = load_library()
    desc, keep = to_desc(scene)
    sizes = (C.c_uint64 * 5)()
    _check(L.cs_hvp_tiles(C.byref(desc), sizes, None, None, None, None, None))
    if sizes[0] == 0:
        return None
    hdr = np.zeros((sizes[0], 8), np.uint32)
    loc = np.zeros(sizes[1], np.uint32)
    lidx = np.zeros((max(sizes[2], 1), 2), np.uint32)
    off = np.zeros(sizes[3], np.uint16)
    ent = np.zeros(max(sizes[4], 1), np.uint16)
    _check(L.cs_hvp_tiles(C.byref(desc), sizes, hdr.ctypes.data_as(_up), loc.ctypes.data_as(_up),
                          lidx.ctypes.data_as(_up), off.ctypes.data_as(C.POINTER(C.c_uint16)),
                          ent.ctypes.data_as(C.POINTER(C.c_uint16))))
    del keep
    return dict(hdr=hdr, loc=loc, lidx=lidx[:sizes[2]], off=off, ent=ent[:sizes[4]])


def ccd_queries(x0, x1, kinds, ids4, thickness, device=0, timing=False):
    """vtCCD / eeCCD over device copies of the inputs; with timing, also the
    kernel time in ms."""
    from .devmem import DeviceArray
    L = load_library()
    x0 = np.ascontiguousarray(x0, np.float64)
    d = [DeviceArray(a, device=device) for a in (x0, np.ascontiguousarray(x1, np.float64),
                                                  np.ascontiguousarray(kinds, np.int32),
                                                  np.ascontiguousarray(ids4, np.int32))]
    out = DeviceArray(dtype=np.float64, shape=(len(kinds),), device=device)
    ms = C.c_double()
    _check(L.cs_ccd_queries(device, d[0].ptr, d[1].ptr, len(x0) // 3, d[2].ptr, d[3].ptr, len(kinds),
                            float(thickness), out.ptr, C.byref(ms) if timing else None))
    toi = out.numpy()
    return (toi, ms.value) if timing else toi


def probe_ccd(x0, x1, kinds, ids4, thickness, device=0):
    """vtCCD / eeCCD on the device for a batch of primitives."""
    L = load_library()
    x0 = np.ascontiguousarray(x0, np.float64)
    x1 = np.ascontiguousarray(x1, np.float64)
    kinds = np.ascontiguousarray(kinds, np.int32)
    ids4 = np.ascontiguousarray(ids4, np.int32)
    out = np.zeros(len(kinds))
    _check(L.cs_probe_ccd(device, x0.ctypes.data_as(_dp), x1.ctypes.data_as(_dp), len(x0) // 3,
                          kinds.ctypes.data_as(C.POINTER(C.c_int32)), ids4.ctypes.data_as(C.POINTER(C.c_int32)),
                          len(kinds), thickness, out.ctypes.data_as(_dp)))
    return out


def probe_closest(x0, kinds, ids4, device=0):
    """Closest point (s, t, dist) on the device for a batch of VT / EE primitives."""
    L = load_library()
    x0 = np.ascontiguousarray(x0, np.float64)
    kinds = np.ascontiguousarray(kinds, np.int32)
    ids4 = np.ascontiguousarray(ids4, np.int32)
    out = np.zeros((len(kinds), 3))
    _check(L.cs_probe_closest(device, x0.ctypes.data_as(_dp), len(x0) // 3, kinds.ctypes.data_as(C.POINTER(C.c_int32)),
                              ids4.ctypes.data_as(C.POINTER(C.c_int32)), len(kinds), out.ctypes.data_as(_dp)))
    return out


class HipSolver:
    """ClothSolver on the GPU.  `precision` is "fp64" (the reference operator,
    bit-faithful to CpuSolver) or "mixed" (FP32 inner solve, FP64 Newton)."""

    def __init__(self, precision="fp64", options=None, **kw):
        self._lib = load_library()
        if options is None:
            prec = _abi.CS_PREC_MIXED if precision == "mixed" else _abi.CS_PREC_FP64
            options = _abi.options(precision=prec, **kw)
        self._opts = options
        self._h = None
        self.scene = None

    # ---- ClothSolver ---------------------------------------------------------
    def initialize(self, scene: ClothScene):
        self.close()
        desc, keep = to_desc(scene)
        h = C.c_void_p()
        _check(self._lib.cs_create(C.byref(desc), C.byref(self._opts), C.byref(h)))
        self._h = h
        self.scene = scene
        self._static_seen = self._static_state()
        del keep

    def _static_state(self):
        c = self.scene.contact
        if c is None or c.staticPos is None:
            return None
        return c.staticPos, np.array(c.staticPos, np.float32, copy=True)

    def _sync_static(self):
        """The device copy of scene.contact's static mesh follows the caller's
        setStaticMesh calls (a moving collider), checked before every step."""
        c = self.scene.contact
        if c is None or c.staticPos is None:
            return
        seen = self._static_seen
        if seen is not None and c.staticPos is seen[0] and np.array_equal(c.staticPos, seen[1]):
            return
        pos = np.ascontiguousarray(c.staticPos, np.float32)
        idx = np.ascontiguousarray(c.staticIdx, np.uint32)
        _check(self._lib.cs_set_static_mesh(self._h, pos.ctypes.data_as(_fp), len(pos) // 3,
                                            idx.ctypes.data_as(_up), len(idx) // 3))
        self._static_seen = self._static_state()

    def _need(self):
        if self._h is None:
            raise ClothError(_abi.CS_E_NOT_INIT, "HipSolver not initialized")

    def step(self, dt):
        self._need()
        self._sync_static()
        _check(self._lib.cs_step(self._h, float(dt)))
        self._copy_out()

    def setMaterial(self, material: ClothMaterial):
        self._need()
        c = material.to_c()
        _check(self._lib.cs_set_material(self._h, C.byref(c)))
        from dataclasses import replace
        self.scene.material = replace(material)

    def pinVertex(self, vertexId, position=None):
        self._need()
        if position is None:
            _check(self._lib.cs_pin_vertex(self._h, vertexId, None))
        else:
            p = np.ascontiguousarray(position, np.float64)
            _check(self._lib.cs_pin_vertex(self._h, vertexId, p.ctypes.data_as(_dp)))
        self._copy_out(velocities=False)
        i = int(vertexId)
        self.scene.pinned[i] = tuple(float(v) for v in self.scene.positions[i * 3:i * 3 + 3])

    def unpinVertex(self, vertexId):
        if self._h is None:
            return
        _check(self._lib.cs_unpin_vertex(self._h, vertexId))
        self.scene.pinned.pop(int(vertexId), None)

    def getPositions(self):
        self._need()
        return self.scene.positions

    def getVelocities(self):
        self._need()
        return self.scene.velocities

    @property
    def lastStats(self):
        if self._h is None:
            return None
        st = _abi.CsStepStats()
        _check(self._lib.cs_last_stats(self._h, C.byref(st)))
        return _abi.stats_to_dict(st)

    # ---- extras --------------------------------------------------------------
    def set_state(self, x, v):
        self._need()
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        _check(self._lib.cs_set_state(self._h, x.ctypes.data_as(_dp), v.ctypes.data_as(_dp), len(x)))
        self.scene.positions[:] = x
        self.scene.velocities[:] = v

    def step_adjoint(self, dt, x_prev, v_prev, x_next, dL_dx_next, dL_dv_next, tol=1e-10, max_iters=4000):
        """Adjoint of one step (cs_step_adjoint): returns dL/dx_prev, dL/dv_prev,
        dL/d(stretchWarp, stretchWeft, stretchCoupling, shear, kb), and the
        adjoint solve's iteration count and relative residual."""
        self._need()
        a = [np.ascontiguousarray(q, np.float64) for q in (x_prev, v_prev, x_next, dL_dx_next, dL_dv_next)]
        dx = np.zeros_like(a[0])
        dv = np.zeros_like(a[0])
        dth = np.zeros(5)
        it = C.c_int()
        rr = C.c_double()
        _check(self._lib.cs_step_adjoint(self._h, dt, *[q.ctypes.data_as(_dp) for q in a], tol, int(max_iters),
                                         dx.ctypes.data_as(_dp), dv.ctypes.data_as(_dp), dth.ctypes.data_as(_dp),
                                         C.byref(it), C.byref(rr)))
        return dx, dv, dth, it.value, rr.value

    def kernel_timing(self):
        self._need()
        out = {}
        for k, name in enumerate(_abi.KERNEL_NAMES):
            ms = C.c_double()
            cnt = C.c_uint64()
            b = C.c_double()
            _check(self._lib.cs_kernel_timing(self._h, k, C.byref(ms), C.byref(cnt), C.byref(b)))
            out[name] = dict(total_ms=ms.value, count=cnt.value, bytes=b.value)
        return out

    def reset_timing(self):
        self._need()
        _check(self._lib.cs_reset_timing(self._h))

    def time_hvp(self, reps=200):
        """Mean HVP launch time (ms, back-to-back launches bracketed by HIP
        events on the solver stream) and algorithmic bytes per launch."""
        self._need()
        ms = C.c_double()
        b = C.c_double()
        _check(self._lib.cs_time_hvp(self._h, int(reps), C.byref(ms), C.byref(b)))
        return ms.value, b.value

    def hvp_layout(self):
        """Tile count of this solver's FP32 HVP (0: the per-vertex gather
        kernel runs), cs_hvp_layout."""
        self._need()
        t = C.c_uint32()
        _check(self._lib.cs_hvp_layout(self._h, C.byref(t)))
        return t.value

    def tuned_launch(self):
        """Grid caps the launch autotuner chose for the contact kernels."""
        self._need()
        buf = C.create_string_buffer(256)
        _check(self._lib.cs_tuned_launch(self._h, buf, 256))
        return buf.value.decode()

    def hvp_bytes(self):
        """(compulsory bytes, bytes this build's kernel moves) of one HVP
        launch on the current operands (cs_hvp_bytes)."""
        self._need()
        a = C.c_double()
        r = C.c_double()
        _check(self._lib.cs_hvp_bytes(self._h, C.byref(a), C.byref(r)))
        return a.value, r.value

    # ---- component probes (parity tests) ------------------------------------
    def probe_hvp(self, x, p, inv_h2=0.0):
        """One launch of the mixed solve's FP32 HVP kernel at (x, p), no
        contact, pinned rows zeroed (cs_probe_hvp)."""
        self._need()
        x = np.ascontiguousarray(x, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        out = np.zeros_like(x)
        _check(self._lib.cs_probe_hvp(self._h, x.ctypes.data_as(_dp), p.ctypes.data_as(_dp), float(inv_h2),
                                      out.ctypes.data_as(_dp)))
        return out

    def probe_hvp_contact(self, x0, x, p, inv_h2=0.0):
        """probe_hvp plus the frozen barrier term of the active set built at x
        for a step begun at x0 (cs_probe_hvp_contact): (out, active count)."""
        self._need()
        x0 = np.ascontiguousarray(x0, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        out = np.zeros_like(x)
        k = C.c_uint32()
        _check(self._lib.cs_probe_hvp_contact(self._h, x0.ctypes.data_as(_dp), x.ctypes.data_as(_dp),
                                              p.ctypes.data_as(_dp), float(inv_h2), out.ctypes.data_as(_dp),
                                              C.byref(k)))
        return out, k.value

    def coarse_pattern(self):
        """(block row offsets, block columns) of the coarse level
        (cs_coarse_pattern)."""
        self._need()
        nnz, D = C.c_uint64(), C.c_uint32()
        _check(self._lib.cs_coarse_pattern(self._h, None, None, 0, C.byref(nnz), C.byref(D)))
        off = np.zeros(D.value + 1, np.uint32)
        col = np.zeros(max(nnz.value, 1), np.uint32)
        _check(self._lib.cs_coarse_pattern(self._h, off.ctypes.data_as(_up), col.ctypes.data_as(_up), nnz.value,
                                           C.byref(nnz), C.byref(D)))
        return off, col[:nnz.value]

    def probe_two_level(self, x, diag, r):
        """The two-level preconditioner at (x, diag) applied to r
        (cs_probe_two_level): dict with z64, z32, raw, sym (nnz x 9) and dc
        (3D)."""
        self._need()
        off, col = self.coarse_pattern()
        nnz, D = len(col), len(off) - 1
        x, diag, r = (np.ascontiguousarray(a, np.float64) for a in (x, diag, r))
        out = dict(z64=np.zeros_like(r), z32=np.zeros_like(r), raw=np.zeros((nnz, 9)), sym=np.zeros((nnz, 9)),
                   dc=np.zeros(D * 3))
        _check(self._lib.cs_probe_two_level(self._h, *(a.ctypes.data_as(_dp) for a in (x, diag, r)),
                                            *(out[k].ctypes.data_as(_dp) for k in ("z64", "z32", "raw", "sym", "dc"))))
        return out

    def probe_active_set(self, x0, x, cap=1 << 20):
        self._need()
        x0 = np.ascontiguousarray(x0, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        ids = np.zeros((cap, 6), np.int32)
        vals = np.zeros((cap, 7))
        cnt = C.c_uint64()
        nc = C.c_uint64()
        _check(self._lib.cs_probe_active_set(self._h, x0.ctypes.data_as(_dp), x.ctypes.data_as(_dp),
                                             ids.ctypes.data_as(C.POINTER(C.c_int32)), vals.ctypes.data_as(_dp),
                                             cap, C.byref(cnt), C.byref(nc)))
        k = cnt.value
        if k > cap:
            return self.probe_active_set(x0, x, cap=int(k))
        return ids[:k].copy(), vals[:k].copy(), nc.value

    def probe_schwarz(self, x, diag, r):
        """buildSchwarzFactors + applySchwarz with this solver's domains:
        (z64, z32, flags64, flags32)."""
        self._need()
        x, diag, r = (np.ascontiguousarray(a, np.float64) for a in (x, diag, r))
        z64 = np.zeros_like(r)
        z32 = np.zeros_like(r)
        nd = C.c_uint32()
        f64 = np.zeros(len(r) // 3 + 1, np.uint8)
        f32 = np.zeros(len(r) // 3 + 1, np.uint8)
        _check(self._lib.cs_probe_schwarz(self._h, x.ctypes.data_as(_dp), diag.ctypes.data_as(_dp),
                                          r.ctypes.data_as(_dp), z64.ctypes.data_as(_dp), z32.ctypes.data_as(_dp),
                                          f64.ctypes.data_as(C.POINTER(C.c_uint8)),
                                          f32.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(nd)))
        return z64, z32, f64[:nd.value].copy(), f32[:nd.value].copy()

    def probe_check_trial(self, x0, xt):
        self._need()
        x0 = np.ascontiguousarray(x0, np.float64)
        xt = np.ascontiguousarray(xt, np.float64)
        v = C.c_int32()
        md = C.c_double()
        mt = C.c_double()
        _check(self._lib.cs_probe_check_trial(self._h, x0.ctypes.data_as(_dp), xt.ctypes.data_as(_dp), C.byref(v),
                                              C.byref(md), C.byref(mt)))
        return bool(v.value), md.value, mt.value

    def probe_contact_adjoint(self, x_step, x, lam=None):
        """(contact gradient at x for a step begun at x_step, and -(d grad /
        d x_step)^T lam) through cs_probe_contact_adjoint."""
        self._need()
        xs = np.ascontiguousarray(x_step, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        grad = np.zeros_like(x)
        lam = None if lam is None else np.ascontiguousarray(lam, np.float64)
        vjp = None if lam is None else np.zeros_like(x)
        lam_p = None if lam is None else lam.ctypes.data_as(_dp)
        _check(self._lib.cs_probe_contact_adjoint(self._h, xs.ctypes.data_as(_dp), x.ctypes.data_as(_dp), lam_p,
                                                  grad.ctypes.data_as(_dp),
                                                  vjp.ctypes.data_as(_dp) if vjp is not None else None))
        return grad, vjp

    def probe_candidates(self, x0, x, pad_is_dmin=False):
        """The cloth and cloth/static triangle pairs the contact system
        expands into candidates at x for a step begun at x0 (pad dHatM, or
        dMinM with pad_is_dmin), through cs_probe_candidates: two (k, 2)
        uint32 arrays in the traversal's order."""
        self._need()
        x0 = np.ascontiguousarray(x0, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        ns, nt = C.c_uint64(), C.c_uint64()
        _check(self._lib.cs_probe_candidates(self._h, x0.ctypes.data_as(_dp), x.ctypes.data_as(_dp),
                                             int(bool(pad_is_dmin)), None, 0, C.byref(ns), None, 0, C.byref(nt)))
        sp = np.zeros((max(ns.value, 1), 2), np.uint32)
        tp = np.zeros((max(nt.value, 1), 2), np.uint32)
        _check(self._lib.cs_probe_candidates(self._h, x0.ctypes.data_as(_dp), x.ctypes.data_as(_dp),
                                             int(bool(pad_is_dmin)), sp.ctypes.data_as(_up), ns.value,
                                             C.byref(ns), tp.ctypes.data_as(_up), nt.value, C.byref(nt)))
        return sp[:ns.value], tp[:nt.value]

    def _copy_out(self, velocities=True):
        sc = self.scene
        n3 = len(sc.positions)
        if not sc.positions.flags.c_contiguous or sc.positions.dtype != np.float64:
            raise ValueError("scene.positions must be a contiguous float64 array")
        _check(self._lib.cs_get_positions(self._h, sc.positions.ctypes.data_as(_dp), n3))
        if velocities:
            _check(self._lib.cs_get_velocities(self._h, sc.velocities.ctypes.data_as(_dp), n3))

    def close(self):
        if self._h is not None:
            self._lib.cs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipBatch:
    """Many independent HipSolvers stepped together (cs_batch_step): host
    threads take the instances largest first and step each on its own HIP
    stream, so small instances' kernels overlap on the GPU.  With packed=True
    (cs_batch_step_packed; mixed precision, Jacobi or block-Jacobi) the
    instances' inner PCG solves run together, one launch per PCG stage over
    all instances."""

    def __init__(self, solvers, threads=None, packed=False):
        self.solvers = list(solvers)
        self.packed = bool(packed)
        if len({id(s) for s in self.solvers}) != len(self.solvers):
            raise ValueError("a solver appears twice in the batch")
        # host threads: 16 measured fastest on the 64-sheet batch (4: 0.43 s,
        # 8: 0.35 s, 16: 0.32 s per batch step), bounded by the host's cores
        self.threads = int(threads) if threads else max(1, min(16, os.cpu_count() or 8))
        self._lib = load_library()

    def step(self, dt):
        for s in self.solvers:
            s._need()
        n = len(self.solvers)
        hs = (C.c_void_p * n)(*[s._h.value for s in self.solvers])
        status = (C.c_int32 * n)()
        if self.packed:
            rc = self._lib.cs_batch_step_packed(hs, n, float(dt), status)
        else:
            rc = self._lib.cs_batch_step(hs, n, float(dt), self.threads, status)
        # the instances that stepped are copied out even when another failed
        for s, st in zip(self.solvers, status):
            if st == _abi.CS_OK:
                s._copy_out()
        _check(rc)
        return list(status)

