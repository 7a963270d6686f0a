"""Device Newton control (CS_FLAG_DEVICE_NEWTON): the descent safeguard and
trust region decided on the device and the Armijo trials evaluated in
batches with one host read per batch (the reference's GPU path batches them
the same way, src/backend/webgpu/gpu-newton.ts:1889-2200).  The trials are
judged in the reference's order and rules (newton.ts:152-219), so the
trajectory, the StepStats and the contact diagnostics must equal the
sequential line search's bit for bit, in FP64 and in mixed precision, and
the FP64 one must still match the CPU oracle."""
import numpy as np
import pytest

import oracle_ffi as O
from closim_amd import _abi, scenes as S
from closim_amd.solver import HipSolver
from contact_scenes import crumpled_sheet, stacked_patches

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0

SCENES = {
    "strip_pull": (lambda: S.strip_pull(9, 29, 0.1, 0.3), dict(grad_tol=1e-6), 12),
    "floor_patch": (S.floor_patch, dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI), 30),
    "stacked_patches": (stacked_patches, dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI), 20),
    "crumpled_sheet": (crumpled_sheet, dict(precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(16, 8, 32)), 12),
}


def drive(scene, solver, k):
    if "bottom" in (getattr(scene, "meta", None) or {}):
        for i, p in S.strip_pull_targets(scene, k):
            solver.pinVertex(i, p)


@pytest.mark.parametrize("speculate", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_device_control_equals_sequential_line_search(name, precision, speculate):
    # one trial per host read, or speculative batches of 1, 2, 4, 8
    make, kw, steps = SCENES[name]
    sa, sb = make(), make()
    a = HipSolver(precision, **kw)
    spec = _abi.CS_FLAG_ARMIJO_SPECULATE if speculate else 0
    b = HipSolver(precision, **kw, flags=_abi.CS_FLAG_DEVICE_NEWTON | spec)
    a.initialize(sa)
    b.initialize(sb)
    trials = 0
    for k in range(steps):
        drive(sa, a, k)
        drive(sb, b, k)
        a.step(H)
        b.step(H)
        assert np.array_equal(sa.positions, sb.positions), k
        assert np.array_equal(sa.velocities, sb.velocities), k
        assert a.lastStats == b.lastStats, k
        trials += b.lastStats["line_search_trials"]
    print(f"{name} {precision}: {steps} steps, {trials} line-search trials", flush=True)
    assert trials > steps  # the scenes exercise rejected trials


def test_device_control_fp64_matches_oracle():
    sc, so = stacked_patches(), stacked_patches()
    g = HipSolver("fp64", precond=_abi.CS_PRECOND_BLOCK_JACOBI, flags=_abi.CS_FLAG_DEVICE_NEWTON)
    g.initialize(sc)
    o = O.OracleSolver(so, _abi.options(precond=_abi.CS_PRECOND_BLOCK_JACOBI))
    for k in range(20):
        g.step(H)
        o.step(H)
        assert np.array_equal(sc.positions, so.positions), k
        st, ref = g.lastStats, o.lastStats
        for key in ("newton_used", "pcg_iters", "line_search_trials"):
            assert st[key] == ref[key], (k, key)
        for key in ("active_pairs", "rejected_line_search_steps", "ccd_failures", "min_distance", "min_toi"):
            assert st["contact"][key] == ref["contact"][key], (k, key)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_breakdown_descent_gate_host_equals_device(precision):
    # CS_FLAG_BREAKDOWN_DESCENT, the reference WebGPU Newton's gate
    # (gpu-solver.ts:1168-1188, newton-control.wgsl:58-73): a PCG stopped on
    # non-positive curvature also takes the Jacobi-scaled steepest descent.
    # The host loop reads the PCG latch, the device control kernel reads it
    # on the device: the same trajectory bit for bit.  On the 64x64 fold the
    # compressed membrane makes PCG break down, so the gated trajectory leaves
    # CpuSolver's.
    kw = dict(precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(64, 1, 64))
    gate = _abi.CS_FLAG_BREAKDOWN_DESCENT
    scs = [S.crumple_fold(63) for _ in range(3)]
    a = HipSolver(precision, **kw, flags=gate)
    b = HipSolver(precision, **kw, flags=gate | _abi.CS_FLAG_DEVICE_NEWTON)
    c = HipSolver(precision, **kw)
    for g, sc in zip((a, b, c), scs):
        g.initialize(sc)
    moved = 0.0
    for k in range(12):
        for g in (a, b, c):
            g.step(H)
        assert np.array_equal(scs[0].positions, scs[1].positions), k
        assert np.array_equal(scs[0].velocities, scs[1].velocities), k
        assert a.lastStats == b.lastStats, k
        moved = max(moved, float(np.max(np.abs(scs[0].positions - scs[2].positions))))
    print(f"fold64 {precision}: gated vs CpuSolver gate max position difference {moved:.3e} m", flush=True)
    assert moved > 0
    for g in (a, b, c):
        g.close()
