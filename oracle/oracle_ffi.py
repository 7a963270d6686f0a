"""ctypes loader for the CPU oracle (libcloth_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product package never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libcloth_oracle.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)
_ip = C.POINTER(C.c_int32)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        from closim_amd import _abi
        sig = {
            "orc_build_grid": (None, [C.c_int, C.c_int, C.c_double, C.c_double, _fp, _fp, _up]),
            "orc_preprocess": (C.c_int, [C.c_uint32, C.c_uint32, _fp, _fp, _up, C.c_double, _fp, _fp, _fp,
                                         _up, _dp, _dp, _dp, C.c_int]),
            "orc_dihedral": (C.c_double, [_dp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_js_hypot3": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "orc_tri_energy_grad": (None, [_dp, _dp, _dp, _dp, C.c_double, C.POINTER(_abi.CsMaterial), _dp, _dp, _dp]),
            "orc_tri_hvp": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.POINTER(_abi.CsMaterial), _dp]),
            "orc_tri_hessian9": (None, [_dp, _dp, _dp, _dp, C.c_double, C.POINTER(_abi.CsMaterial), _dp]),
            "orc_hinge_energy": (C.c_double, [_dp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_eval_internal": (None, [C.c_void_p, _dp, _dp, _dp, _dp]),
            "orc_internal_energy": (C.c_double, [C.c_void_p, _dp]),
            "orc_membrane_hvp_fd": (None, [C.c_void_p, _dp, _dp, _dp]),
            "orc_membrane_hvp_analytic": (None, [C.c_void_p, _dp, _dp, _dp]),
            "orc_membrane_blocks": (None, [C.c_void_p, _dp, _dp]),
            "orc_closest_vt": (None, [_dp, _dp, _dp, _dp, _dp]),
            "orc_closest_ee": (None, [_dp, _dp, _dp, _dp, _dp]),
            "orc_solve_cubic01": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
            "orc_vt_ccd": (C.c_double, [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
            "orc_ee_ccd": (C.c_double, [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
            "orc_barrier_value": (C.c_double, [C.c_double, C.c_double]),
            "orc_barrier_grad": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_coulomb": (None, [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp]),
            "orc_self_pairs": (C.c_int64, [_dp, _dp, _up, C.c_uint32, C.c_double, _up, C.c_int64]),
            "orc_create": (C.c_void_p, [C.POINTER(_abi.CsSceneDesc), C.POINTER(_abi.CsOptions)]),
            "orc_destroy": (None, [C.c_void_p]),
            "orc_step": (C.c_int, [C.c_void_p, C.c_double]),
            "orc_pin_vertex": (C.c_int, [C.c_void_p, C.c_uint32, _dp]),
            "orc_set_static_mesh": (C.c_int, [C.c_void_p, _fp, C.c_uint32, _up, C.c_uint32]),
            "orc_unpin_vertex": (C.c_int, [C.c_void_p, C.c_uint32]),
            "orc_set_material": (None, [C.c_void_p, C.POINTER(_abi.CsMaterial)]),
            "orc_get_positions": (None, [C.c_void_p, _dp]),
            "orc_get_velocities": (None, [C.c_void_p, _dp]),
            "orc_set_state": (None, [C.c_void_p, _dp, _dp]),
            "orc_last_stats": (None, [C.c_void_p, C.POINTER(_abi.CsStepStats)]),
            "orc_last_error": (C.c_char_p, []),
            "orc_contact_probe": (C.c_int64, [C.c_void_p, _dp, _dp, C.c_double, _ip, _dp, C.c_int64,
                                              C.POINTER(C.c_int64)]),
            "orc_check_trial": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
            "orc_candidates": (C.c_int64, [C.c_void_p, _dp, _dp, C.c_int, _ip, C.c_int64]),
            "orc_contact_hvp": (C.c_int64, [C.c_void_p, _dp, _dp, _dp, _dp]),
            "orc_contact_grad_frozen": (C.c_int64, [C.c_void_p, _dp, _dp, _dp, _dp]),
            "orc_coarse_pattern": (C.c_int64, [C.c_void_p, C.c_int, C.c_int, C.c_int, _up, _up, C.c_int64]),
            "orc_coarse_values": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]),
            "orc_coarse_transfer": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]),
            "orc_coarse_exact": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
            "orc_coarse_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                           C.c_double, _dp, _dp, _dp]),
            "orc_two_level_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                              C.c_double, _dp, _dp]),
            "orc_schwarz_domains": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _up, _up]),
            "orc_schwarz_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                            C.POINTER(C.c_uint8)]),
            "orc_set_fault": (None, [C.c_int]),
            "orc_set_sum_mode": (None, [C.c_int]),
            "orc_armijo_trace": (C.c_int64, [C.c_void_p, _dp, C.c_int64]),
            "orc_block_factors": (None, [C.c_void_p, _dp, _dp, _dp, C.POINTER(C.c_uint8)]),
            "orc_apply_block_factors": (None, [_dp, _dp, _dp, C.c_uint32]),
            "orc_pcg_dense": (C.c_int, [_dp, C.c_uint32, _dp, _dp, _up, C.c_uint32, C.c_int, C.c_double, _dp,
                                        C.POINTER(C.c_int), _dp]),
            "orc_hinge_grad_fd": (C.c_double, [_dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
            "orc_static_pairs": (C.c_int64, [_dp, _dp, _up, C.c_uint32, _fp, C.c_uint32, _up, C.c_uint32,
                                             C.c_double, _up, C.c_int64]),
            "orc_kill_velocity": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
            "orc_commit_probe": (C.c_int64, [C.c_void_p, _dp, _dp, C.POINTER(C.c_int64), _dp, C.c_int64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def P(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def d3(v):
    return np.ascontiguousarray(v, np.float64)


class OracleSolver:
    """CpuSolver restatement behind the ClothSolver surface (cpu-solver.ts:7-49)."""

    def __init__(self, scene, opts=None):
        from closim_amd import _abi
        from closim_amd.scene import to_desc
        self.scene = scene
        self.n3 = scene.mesh.count * 3
        desc, keep = to_desc(scene)
        self._keep = keep
        opts = opts or _abi.options()
        self._opts = opts
        self.h = lib().orc_create(C.byref(desc), C.byref(opts))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.orc_destroy(h)
            self.h = None

    def step(self, dt):
        rc = lib().orc_step(self.h, dt)
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode(errors="replace"))
        lib().orc_get_positions(self.h, P(self.scene.positions))
        lib().orc_get_velocities(self.h, P(self.scene.velocities))

    def set_static_mesh(self, pos, idx):
        """ContactSystem.setStaticMesh (contact-assembly.ts:71-73)."""
        pos = np.ascontiguousarray(pos, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        rc = lib().orc_set_static_mesh(self.h, pos.ctypes.data_as(_fp), len(pos) // 3, idx.ctypes.data_as(_up),
                                       len(idx) // 3)
        assert rc == 0
        self.scene.contact.staticPos = pos

    def pinVertex(self, vid, position=None):
        if position is None:
            lib().orc_pin_vertex(self.h, vid, None)
        else:
            p = d3(position)
            lib().orc_pin_vertex(self.h, vid, P(p))
        pos = np.empty(self.n3)
        lib().orc_get_positions(self.h, P(pos))
        self.scene.positions[:] = pos
        self.scene.pinned[vid] = tuple(pos[vid * 3:vid * 3 + 3]) if position is None else tuple(position)

    def unpinVertex(self, vid):
        lib().orc_unpin_vertex(self.h, vid)
        self.scene.pinned.pop(vid, None)

    def setMaterial(self, mat):
        c = mat.to_c()
        lib().orc_set_material(self.h, C.byref(c))

    def getPositions(self):
        return self.scene.positions

    def getVelocities(self):
        return self.scene.velocities

    def set_state(self, x, v):
        x = d3(x); v = d3(v)
        lib().orc_set_state(self.h, P(x), P(v))
        self.scene.positions[:] = x
        self.scene.velocities[:] = v

    @property
    def lastStats(self):
        from closim_amd import _abi
        st = _abi.CsStepStats()
        lib().orc_last_stats(self.h, C.byref(st))
        return _abi.stats_to_dict(st)

    # ---- probes for component parity tests ----
    def eval_internal(self, x):
        x = d3(x)
        g = np.zeros(self.n3)
        e = C.c_double()
        ms = C.c_double()
        lib().orc_eval_internal(self.h, P(x), P(g), C.byref(e), C.byref(ms))
        return g, e.value, ms.value

    def internal_energy(self, x):
        return lib().orc_internal_energy(self.h, P(d3(x)))

    def membrane_hvp(self, x, p, analytic=False):
        x = d3(x); p = d3(p)
        out = np.zeros(self.n3)
        f = lib().orc_membrane_hvp_analytic if analytic else lib().orc_membrane_hvp_fd
        f(self.h, P(x), P(p), P(out))
        return out

    def membrane_blocks(self, x):
        x = d3(x)
        out = np.zeros(self.n3 * 3)
        lib().orc_membrane_blocks(self.h, P(x), P(out))
        return out

    def active_set(self, x0, x, cap=None, candidates_dmin=False):
        x0 = d3(x0); x = d3(x)
        cap = cap or 1 << 20
        ids = np.zeros((cap, 6), np.int32)
        vals = np.zeros((cap, 7))
        nc = C.c_int64()
        k = lib().orc_contact_probe(self.h, P(x0), P(x), 1.0 if candidates_dmin else 0.0,
                                    P(ids, C.c_int32), P(vals), cap, C.byref(nc))
        if k > cap:
            return self.active_set(x0, x, cap=int(k), candidates_dmin=candidates_dmin)
        return ids[:k].copy(), vals[:k].copy(), nc.value

    def contact_hvp(self, x0, x, p):
        """(applyHvp of the active set built at x after beginStep(x0), active
        contact count)."""
        out = np.zeros(self.n3)
        k = lib().orc_contact_hvp(self.h, P(d3(x0)), P(d3(x)), P(d3(p)), P(out))
        if k < 0:
            raise RuntimeError("scene has no contact system")
        return out, int(k)

    # ---- two-level Schwarz (mas.ts, coarse-csr.ts); dom = (target, min, max)
    def coarse_pattern(self, dom):
        D = self.schwarz_domains(*dom)[2]
        off = np.zeros(D + 1, np.uint32)
        nnz = lib().orc_coarse_pattern(self.h, *dom, P(off, C.c_uint32), None, 0)
        col = np.zeros(max(nnz, 1), np.uint32)
        lib().orc_coarse_pattern(self.h, *dom, P(off, C.c_uint32), P(col, C.c_uint32), nnz)
        return off, col[:nnz]

    def coarse_values(self, dom, x, diag):
        """(raw blocks, symmetrised blocks (nnz x 9), buildCoarseDiag (3D))."""
        off, col = self.coarse_pattern(dom)
        raw, sym = np.zeros((len(col), 9)), np.zeros((len(col), 9))
        dc = np.zeros((len(off) - 1) * 3)
        lib().orc_coarse_values(self.h, *dom, P(d3(x)), P(d3(diag)), P(raw), P(sym), P(dc))
        return raw, sym, dc

    def coarse_restrict(self, dom, r, D):
        out = np.zeros(D * 3)
        lib().orc_coarse_transfer(self.h, *dom, 0, P(d3(r)), P(out))
        return out

    def coarse_prolongate_add(self, dom, c, z):
        z = d3(z).copy()
        lib().orc_coarse_transfer(self.h, *dom, 1, P(d3(c)), P(z))
        return z

    def coarse_exact(self, dom, x, diag, D):
        """(dense Ac (3D x 3D), its Cholesky factor or Jacobi fallback, flag)."""
        ac, ch = np.zeros((D * 3, D * 3)), np.zeros((D * 3, D * 3))
        flag = lib().orc_coarse_exact(self.h, *dom, P(d3(x)), P(d3(diag)), P(ac), P(ch))
        return ac, ch, flag

    def coarse_solve(self, dom, x, diag, mode, rhs, iters=8, tol=1e-3):
        """mode 2 block Jacobi, 3 coarsePcg (iters, tol), 4 exact: (out, returned count or flag, residual)."""
        out = np.zeros(len(rhs))
        res = C.c_double(np.nan)
        k = lib().orc_coarse_solve(self.h, *dom, P(d3(x)), P(d3(diag)), mode, iters, tol, P(d3(rhs)), P(out),
                                   C.byref(res))
        return out, k, res.value

    def two_level_apply(self, dom, x, diag, r, mode, iters=8, omega=0.5):
        z = np.zeros(self.n3)
        lib().orc_two_level_apply(self.h, *dom, P(d3(x)), P(d3(diag)), mode, iters, omega, P(d3(r)), P(z))
        return z

    def contact_grad_frozen(self, x0, x, y):
        """(contact gradient at y of the active set frozen at x after
        beginStep(x0), active count)."""
        out = np.zeros(self.n3)
        k = lib().orc_contact_grad_frozen(self.h, P(d3(x0)), P(d3(x)), P(d3(y)), P(out))
        if k < 0:
            raise RuntimeError("scene has no contact system")
        return out, int(k)

    def candidates(self, x0, x, dmin=False, cap=None):
        """buildCandidates (contact-assembly.ts:87-121) at x after beginStep(x0),
        padded by dMinM (dmin) or dHatM: sorted (k, 6) int32 rows
        (kind, p, a, b, c, d)."""
        x0 = d3(x0); x = d3(x)
        cap = cap or 1 << 16
        ids = np.zeros((cap, 6), np.int32)
        k = lib().orc_candidates(self.h, P(x0), P(x), 1 if dmin else 0, P(ids, C.c_int32), cap)
        if k < 0:
            raise RuntimeError("scene has no contact system")
        if k > cap:
            return self.candidates(x0, x, dmin=dmin, cap=int(k))
        return ids[:k].copy()

    def schwarz_domains(self, target=16, minsz=8, maxsz=32):
        n = len(self.getPositions()) // 3
        dom = np.zeros(n, np.uint32)
        loc = np.zeros(n, np.uint32)
        cnt = lib().orc_schwarz_domains(self.h, target, minsz, maxsz, dom.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        loc.ctypes.data_as(C.POINTER(C.c_uint32)))
        return dom, loc, cnt

    def schwarz_apply(self, x, diag, r, target=16, minsz=8, maxsz=32):
        x = d3(x); diag = d3(diag); r = d3(r)
        z = np.zeros_like(r)
        flags = np.zeros(len(x) // 3 + 1, np.uint8)
        cnt = lib().orc_schwarz_apply(self.h, target, minsz, maxsz, P(x), P(diag), P(r), P(z),
                                      flags.ctypes.data_as(C.POINTER(C.c_uint8)))
        return z, flags[:cnt].copy()

    def armijo_trace(self):
        cap = 4096
        out = np.zeros((cap, 4))
        k = lib().orc_armijo_trace(self.h, P(out), cap)
        return out[:min(k, cap)].copy()

    def block_factors(self, x, diag):
        x = d3(x); diag = d3(diag)
        n = self.n3 // 3
        inv = np.zeros(9 * n)
        flags = np.zeros(n, np.uint8)
        lib().orc_block_factors(self.h, P(x), P(diag), P(inv), flags.ctypes.data_as(C.POINTER(C.c_uint8)))
        return inv, flags

    def kill_velocity(self, x0, x, v):
        v = d3(v).copy()
        k = lib().orc_kill_velocity(self.h, P(d3(x0)), P(d3(x)), P(v))
        return v, k

    def commit_table(self, x0, x):
        cap = 1 << 16
        keys = np.zeros((cap, 5), np.int64)
        vals = np.zeros((cap, 4))
        k = lib().orc_commit_probe(self.h, P(d3(x0)), P(d3(x)), P(keys, C.c_int64), P(vals), cap)
        return keys[:k].copy(), vals[:k].copy()

    def check_trial(self, x0, xt):
        md = C.c_double()
        mt = C.c_double()
        v = lib().orc_check_trial(self.h, P(d3(x0)), P(d3(xt)), C.byref(md), C.byref(mt))
        return bool(v), md.value, mt.value


def self_pairs(x0, x1, idx, pad):
    x0 = d3(x0); x1 = d3(x1)
    idx = np.ascontiguousarray(idx, np.uint32)
    m = len(idx) // 3
    cap = 1 << 16
    while True:
        out = np.zeros((cap, 2), np.uint32)
        k = lib().orc_self_pairs(P(x0), P(x1), P(idx, C.c_uint32), m, pad, P(out, C.c_uint32), cap)
        if k <= cap:
            return out[:k].copy()
        cap = int(k)


def set_sum_mode(faithful):
    lib().orc_set_sum_mode(1 if faithful else 0)


def apply_block_factors(inv, r):
    inv = d3(inv); r = d3(r)
    z = np.zeros_like(r)
    lib().orc_apply_block_factors(P(inv), P(r), P(z), len(r) // 3)
    return z


def pcg_dense(A, b, diag, pins=(), max_iters=100, tol=1e-6):
    A = d3(A); b = d3(b); diag = d3(diag)
    pins = np.ascontiguousarray(pins, np.uint32)
    x = np.zeros_like(b)
    it = C.c_int()
    res = C.c_double()
    ok = lib().orc_pcg_dense(P(A), len(b), P(b), P(diag), P(pins, C.c_uint32), len(pins), max_iters, tol, P(x),
                             C.byref(it), C.byref(res))
    return x, it.value, res.value, bool(ok)


def hinge_grad_fd(x12, rest, ln, asum, k):
    x12 = d3(x12)
    g = np.zeros(12)
    e = lib().orc_hinge_grad_fd(P(x12), rest, ln, asum, k, P(g))
    return g, e


def static_pairs(x0, x1, idx, spos, sidx, pad):
    x0 = d3(x0); x1 = d3(x1)
    idx = np.ascontiguousarray(idx, np.uint32)
    spos = np.ascontiguousarray(spos, np.float32)
    sidx = np.ascontiguousarray(sidx, np.uint32)
    cap = 1 << 16
    while True:
        out = np.zeros((cap, 2), np.uint32)
        k = lib().orc_static_pairs(P(x0), P(x1), P(idx, C.c_uint32), len(idx) // 3, P(spos, C.c_float),
                                   len(spos) // 3, P(sidx, C.c_uint32), len(sidx) // 3, pad,
                                   P(out, C.c_uint32), cap)
        if k <= cap:
            return out[:k].copy()
        cap = int(k)


def preprocess(pos, uv, idx, rho):
    pos = np.ascontiguousarray(pos, np.float32)
    uv = np.ascontiguousarray(uv, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    n = len(pos) // 3
    m = len(idx) // 3
    inv = np.zeros(4 * m, np.float32)
    areas = np.zeros(m, np.float32)
    masses = np.zeros(n, np.float32)
    mh = 3 * m
    hv = np.zeros(4 * mh, np.uint32)
    rest = np.zeros(mh); ln = np.zeros(mh); asum = np.zeros(mh)
    k = lib().orc_preprocess(n, m, P(pos, C.c_float), P(uv, C.c_float), P(idx, C.c_uint32), rho,
                             P(inv, C.c_float), P(areas, C.c_float), P(masses, C.c_float),
                             P(hv, C.c_uint32), P(rest), P(ln), P(asum), mh)
    if k < 0:
        raise ValueError(lib().orc_last_error().decode())
    return dict(inv_dm=inv, areas=areas, masses=masses, hinge_verts=hv[:4 * k].reshape(-1, 4),
                hinge_rest=rest[:k], hinge_len=ln[:k], hinge_area=asum[:k])


def build_grid(nx, ny, w, h):
    n = (nx + 1) * (ny + 1)
    pos = np.zeros(3 * n, np.float32)
    uv = np.zeros(2 * n, np.float32)
    idx = np.zeros(6 * nx * ny, np.uint32)
    lib().orc_build_grid(nx, ny, w, h, P(pos, C.c_float), P(uv, C.c_float), P(idx, C.c_uint32))
    return pos, uv, idx
