"""FP64 HipSolver vs the CPU oracle on identical seeded scenes.

The FP64 GPU path runs the reference's own operator (FD membrane HVP, Jacobi
PCG) with the same per-element arithmetic and correctly rounded reductions,
so the bar is 1e-9 relative per step (the tolerance the cloth-step contract
states for the plain FP64 path); in practice the trajectories are identical.
"""
import numpy as np
import pytest

import oracle_ffi as O
from closim_amd import _abi, scenes as S
from closim_amd.solver import HipSolver

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0
REL_TOL = 1e-9


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run_pair(make_scene, steps, opts_kw=None, drive=None):
    opts_kw = opts_kw or {}
    sc_g = make_scene()
    sc_o = make_scene()
    g = HipSolver("fp64", **opts_kw)
    g.initialize(sc_g)
    o = O.OracleSolver(sc_o, _abi.options(**opts_kw))
    errs = []
    for k in range(steps):
        if drive:
            drive(g, sc_g, k)
            drive(o, sc_o, k)
        g.step(H)
        o.step(H)
        e = rel_err(g.getPositions(), o.getPositions())
        ev = rel_err(g.getVelocities(), o.getVelocities())
        sg, so = g.lastStats, o.lastStats
        print(f"step {k}: pos rel err {e:.3e} vel {ev:.3e} pcg {sg['pcg_iters']}/{so['pcg_iters']} "
              f"newton {sg['newton_used']}/{so['newton_used']}", flush=True)
        errs.append(e)
        assert e <= REL_TOL, (k, e)
        assert sg["pcg_iters"] == so["pcg_iters"] and sg["newton_used"] == so["newton_used"]
        assert sg["line_search_trials"] == so["line_search_trials"]
    return errs, g, o


def test_small_strip_fp64_matches_oracle():
    errs, g, o = run_pair(S.small_strip, 20)
    assert max(errs) <= REL_TOL


def test_strip_pull_fp64_matches_oracle_per_step():
    # Strip-pull workload: 20x100 vertices, top pinned, bottom driven 5 cm
    # down over 60 steps, Newton tol 1e-6, Jacobi PCG.
    def drive(solver, scene, k):
        for i, p in S.strip_pull_targets(scene, k):
            solver.pinVertex(i, p)
    errs, g, o = run_pair(S.strip_pull, 60, dict(grad_tol=1e-6), drive)
    assert max(errs) <= REL_TOL
    st = g.lastStats
    assert st["grad_norm"] == o.lastStats["grad_norm"] or abs(st["grad_norm"] - o.lastStats["grad_norm"]) <= 1e-9 * o.lastStats["grad_norm"]


def test_strip_fp64_analytic_and_block_jacobi_match_oracle():
    kw = dict(hvp_mode=_abi.CS_HVP_ANALYTIC, precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    errs, g, o = run_pair(S.small_strip, 10, kw)
    assert max(errs) <= REL_TOL


def test_strip_pull_fp64_schwarz_matches_oracle_per_step():
    # Additive Schwarz (schwarz.ts): 16-vertex domains on the driven strip.
    def drive(solver, scene, k):
        for i, p in S.strip_pull_targets(scene, k):
            solver.pinVertex(i, p)
    kw = dict(grad_tol=1e-6, precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(16, 8, 32))
    errs, g, o = run_pair(S.strip_pull, 30, kw, drive)
    assert max(errs) <= REL_TOL


def test_small_strip_fp64_schwarz_large_domains_matches_oracle():
    # 64-vertex domains (192x192 local Cholesky) with the analytic HVP.
    kw = dict(precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(64, 1, 64), hvp_mode=_abi.CS_HVP_ANALYTIC)
    errs, g, o = run_pair(S.small_strip, 10, kw)
    assert max(errs) <= REL_TOL


def _schwarz_case(sizes, seed, stretch=1.02):
    sc = S.small_strip()
    rng = np.random.default_rng(seed)
    x = sc.positions * stretch + 1e-4 * rng.standard_normal(sc.positions.shape)
    masses = np.repeat(np.asarray(sc.mesh.masses, np.float64), 3)
    diag = masses * 3600.0 + 50.0 * (1 + rng.random(len(x)))
    r = rng.standard_normal(len(x))
    g = HipSolver("fp64", precond=_abi.CS_PRECOND_SCHWARZ, schwarz=sizes)
    g.initialize(S.small_strip())
    o = O.OracleSolver(S.small_strip())
    return g, o, x, diag, r


@pytest.mark.parametrize("sizes", [(16, 8, 32), (64, 1, 64), (1, 0, 1), (91, 1, 91)])
def test_schwarz_apply_component_parity(sizes):
    # buildSchwarzFactors + applySchwarz: the FP64 apply is the oracle's bit for
    # bit; the mixed solve's FP32 factors and apply agree to FP32 accuracy
    g, o, x, diag, r = _schwarz_case(sizes, 11)
    z64, z32, f64, f32 = g.probe_schwarz(x, diag, r)
    zo, fo = o.schwarz_apply(x, diag, r, *sizes)
    assert np.array_equal(f64, fo) and np.all(f64 == 1)
    assert np.array_equal(z64, zo)
    assert np.array_equal(f32, fo)
    err = np.max(np.abs(z32 - z64)) / np.max(np.abs(z64))
    print(f"domains {len(f64)}, FP32 apply rel err {err:.2e}")
    assert err < 1e-3


def test_schwarz_apply_jacobi_fallback_parity():
    g, o, x, diag, r = _schwarz_case((16, 8, 32), 12)
    diag = -np.abs(diag)
    diag[::7] = 0.0
    z64, z32, f64, f32 = g.probe_schwarz(x, diag, r)
    zo, fo = o.schwarz_apply(x, diag, r, 16, 8, 32)
    assert np.all(fo == 0) and np.array_equal(f64, fo) and np.array_equal(f32, fo)
    assert np.array_equal(z64, zo)
    assert np.allclose(z32, zo, rtol=1e-6, atol=0)


@pytest.mark.parametrize("newton,ls", [(0, 10), (1, 10), (3, 1), (10, 0)])
def test_iteration_caps_match_oracle(newton, ls):
    # the reference's failure-injection corners (maxNewton = 0 / 1, a single
    # line-search trial, none at all): the step must still be the oracle's,
    # on a contact scene where the line search rejects trials
    run_pair(S.floor_patch, 8, dict(newton_iters=newton, line_search_iters=ls))


def test_nonfinite_state_is_reported():
    # newton.ts:129: a non-finite gradient throws "non-finite gradient"
    from closim_amd.solver import ClothError
    sc = S.small_strip()
    g = HipSolver("fp64")
    g.initialize(sc)
    x = sc.positions.copy()
    free = next(v for v in range(sc.mesh.count) if v not in sc.pinned)
    x[free * 3 + 1] = np.inf
    g.set_state(x, sc.velocities)
    with pytest.raises(ClothError) as e:
        g.step(H)
    assert e.value.code == _abi.CS_E_NONFINITE
    assert "non-finite gradient" in str(e.value)
    g.close()
