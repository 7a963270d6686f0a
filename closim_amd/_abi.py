"""ctypes mirror of include/cloth_hip.h (the C ABI of the HIP cloth step).

Only structure layouts and constants live here; `closim_amd.solver` binds the
functions.  Field order must match the header exactly.
"""
import ctypes as C

CS_OK = 0
CS_E_NOT_INIT = 1
CS_E_INVALID = 2
CS_E_NONFINITE = 3
CS_E_HIP = 4
CS_E_OVERFLOW = 5
CS_E_NOMEM = 6
CS_E_UNSUPPORTED = 7

CS_PREC_FP64 = 0
CS_PREC_MIXED = 1
CS_HVP_FD = 0
CS_HVP_ANALYTIC = 1
CS_PRECOND_JACOBI = 0
CS_PRECOND_BLOCK_JACOBI = 1
CS_PRECOND_SCHWARZ = 2
CS_COARSE_NONE = 0    # one-level Schwarz
CS_COARSE_JACOBI = 1  # + omega P Dc^-1 R (applyMasTwoLevel, mas.ts:126-143)
CS_COARSE_BLOCK = 2   # + P Bdd^-1 R (the GPU path's C0)
CS_COARSE_PCG = 3     # + P coarsePcg(Ac, R r, K) (C1, coarse-csr.ts:243-303)
CS_FLAG_TIMING = 1
CS_FLAG_POLISH_ROOT = 2
CS_FLAG_EAGER = 4
CS_FLAG_PSD_NEWTON = 8
CS_FLAG_DEVICE_NEWTON = 32
CS_FLAG_ARMIJO_SPECULATE = 128
CS_FLAG_BREAKDOWN_DESCENT = 64

CS_COPY_HOST_TO_DEVICE = 1
CS_COPY_DEVICE_TO_HOST = 2
CS_COPY_DEVICE_TO_DEVICE = 3

CS_KERNEL_HVP = 0
CS_KERNEL_GRADIENT = 1
CS_KERNEL_BROADPHASE = 2
CS_KERNEL_CCD = 3
CS_KERNEL_PRECOND = 4
CS_KERNEL_COUNT = 5
KERNEL_NAMES = ["hvp", "gradient", "broadphase", "ccd", "precond"]

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)


class CsMaterial(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "areal_density", "thickness", "stretch_warp", "stretch_weft",
        "stretch_coupling", "shear", "bend_warp", "bend_weft", "damping")]


class CsContactParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "d_hat", "kappa", "d_min", "friction_mu", "friction_eps")]


class CsSceneDesc(C.Structure):
    _fields_ = [
        ("n_verts", C.c_uint32), ("n_tris", C.c_uint32), ("n_hinges", C.c_uint32),
        ("positions", _dp), ("velocities", _dp), ("indices", _up),
        ("inv_dm", _fp), ("areas", _fp), ("masses", _fp), ("rest_positions", _fp),
        ("hinge_verts", _up), ("hinge_rest_angle", _dp), ("hinge_edge_len", _dp),
        ("hinge_area_sum", _dp),
        ("material", CsMaterial), ("gravity", C.c_double * 3),
        ("n_pins", C.c_uint32), ("pin_ids", _up), ("pin_pos", _dp),
        ("contact_enabled", C.c_int32), ("contact", CsContactParams),
        ("has_floor", C.c_int32), ("floor_y", C.c_double),
        ("n_static_verts", C.c_uint32), ("n_static_tris", C.c_uint32),
        ("static_pos", _fp), ("static_idx", _up),
    ]


class CsOptions(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int32), ("pcg_max_iters", C.c_int32),
        ("line_search_iters", C.c_int32), ("grad_tol", C.c_double),
        ("precision", C.c_int32), ("hvp_mode", C.c_int32), ("precond", C.c_int32),
        ("schwarz_target", C.c_int32), ("schwarz_min", C.c_int32), ("schwarz_max", C.c_int32),
        ("device", C.c_int32), ("flags", C.c_uint32),
        ("coarse", C.c_int32), ("coarse_iters", C.c_int32), ("coarse_omega", C.c_double),
    ]


class CsContactDiag(C.Structure):
    _fields_ = [
        ("candidate_pairs", C.c_uint64), ("active_pairs", C.c_uint64),
        ("vt_pairs", C.c_uint64), ("ee_pairs", C.c_uint64), ("floor_pairs", C.c_uint64),
        ("min_distance", C.c_double), ("min_toi", C.c_double),
        ("barrier_energy", C.c_double), ("friction_work", C.c_double),
        ("rejected_line_search_steps", C.c_uint64), ("ccd_failures", C.c_uint64),
        ("filter_rounds", C.c_uint64),
    ]


class CsStepStats(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int32), ("pcg_iters", C.c_int32),
        ("grad_norm", C.c_double), ("energy", C.c_double), ("max_strain", C.c_double),
        ("converged", C.c_int32), ("has_contact", C.c_int32),
        ("contact", CsContactDiag),
        ("newton_used", C.c_int32), ("line_search_trials", C.c_int32),
    ]


def options(newton_iters=10, pcg_max_iters=60, line_search_iters=10, grad_tol=1e-5,
            precision=CS_PREC_FP64, hvp_mode=None, precond=CS_PRECOND_JACOBI,
            schwarz=(16, 8, 32), device=0, flags=0, coarse=CS_COARSE_NONE, coarse_iters=8, coarse_omega=0.5):
    """StepperOptions defaults of the reference (newton.ts:30-33).

    hvp_mode defaults to the reference CpuSolver operator (FD) in FP64 and to
    the analytic HVP in mixed precision.
    """
    if hvp_mode is None:
        hvp_mode = CS_HVP_FD if precision == CS_PREC_FP64 else CS_HVP_ANALYTIC
    o = CsOptions()
    o.newton_iters = newton_iters
    o.pcg_max_iters = pcg_max_iters
    o.line_search_iters = line_search_iters
    o.grad_tol = grad_tol
    o.precision = precision
    o.hvp_mode = hvp_mode
    o.precond = precond
    o.schwarz_target, o.schwarz_min, o.schwarz_max = schwarz
    o.device = device
    o.flags = flags
    o.coarse = coarse
    o.coarse_iters = coarse_iters
    o.coarse_omega = coarse_omega
    return o


def stats_to_dict(st):
    d = {f: getattr(st, f) for f, _ in CsStepStats._fields_ if f != "contact"}
    d["contact"] = {f: getattr(st.contact, f) for f, _ in CsContactDiag._fields_} if st.has_contact else None
    return d
