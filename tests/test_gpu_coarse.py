"""The device coarse level of the two-level Schwarz preconditioner against
the oracle's restatement of mas.ts / coarse-csr.ts: the block pattern, the
raw and symmetrised blocks and Dc bit for bit (the same sums in the same
order), the two-level apply z = Schwarz1(r) + P C(R r) bit for bit for the
single-sweep corrections (damped Jacobi, block Jacobi) and to 1e-12 for the
coarse PCG (its reductions are trees on the device, left to right in the
reference), the FP32 apply of the mixed solve to FP32 accuracy, and the
preconditioner inside the solver's PCG."""
import numpy as np
import pytest

import oracle_ffi as O
from closim_amd import _abi, scenes as S
from closim_amd.solver import ClothError, HipSolver

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0


def cases():
    return [("strip", S.small_strip, (16, 8, 32)), ("fold64", lambda: S.crumple_fold(63), (64, 1, 64)),
            ("drape48", lambda: S.sphere_drape(47), (32, 8, 64))]


@pytest.mark.parametrize("name,make,dom", cases(), ids=[c[0] for c in cases()])
@pytest.mark.parametrize("mode", [_abi.CS_COARSE_JACOBI, _abi.CS_COARSE_BLOCK, _abi.CS_COARSE_PCG])
def test_two_level_apply_matches_oracle(name, make, dom, mode):
    sc = make()
    g = HipSolver("fp64", precond=_abi.CS_PRECOND_SCHWARZ, schwarz=dom, coarse=mode)
    g.initialize(sc)
    o = O.OracleSolver(make())
    rng = np.random.default_rng(7)
    x = sc.positions * 1.01 + rng.standard_normal(len(sc.positions)) * 1e-4
    diag = np.repeat(np.asarray(sc.mesh.masses, np.float64) * 3600.0 + 2.0, 3)
    r = rng.standard_normal(len(x))
    off, col = g.coarse_pattern()
    ooff, ocol = o.coarse_pattern(dom)
    assert np.array_equal(off, ooff) and np.array_equal(col, ocol)
    got = g.probe_two_level(x, diag, r)
    raw, sym, dc = o.coarse_values(dom, x, diag)
    assert np.array_equal(got["raw"], raw) and np.array_equal(got["sym"], sym) and np.array_equal(got["dc"], dc)
    want = o.two_level_apply(dom, x, diag, r, mode, iters=8, omega=0.5)
    err = np.max(np.abs(got["z64"] - want)) / np.max(np.abs(want))
    e32 = np.max(np.abs(got["z32"] - want)) / np.max(np.abs(want))
    print(f"[coarse] {name} mode {mode}: {len(off) - 1} domains, {len(col)} blocks, FP64 rel {err:.1e}, "
          f"FP32 rel {e32:.1e}", flush=True)
    if mode == _abi.CS_COARSE_PCG:
        assert err <= 1e-12
    else:
        assert np.array_equal(got["z64"], want)
    assert e32 <= 1e-4
    g.close()


def test_coarse_options_are_validated():
    sc = S.small_strip()
    for kw in (dict(coarse=_abi.CS_COARSE_PCG),  # needs Schwarz
               dict(precond=_abi.CS_PRECOND_SCHWARZ, coarse=7),
               dict(precond=_abi.CS_PRECOND_SCHWARZ, coarse=_abi.CS_COARSE_PCG, coarse_iters=-1)):
        g = HipSolver("fp64", **kw)
        with pytest.raises(ClothError) as e:
            g.initialize(sc)
        assert e.value.code == _abi.CS_E_INVALID


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_coarse_pcg_cuts_pcg_iterations_on_a_stiff_fold(precision):
    # the 64x64 fold with 10x stiffer membrane and 64-vertex domains: the
    # same steps with the one-level and the two-level (C1, 8 coarse
    # iterations) preconditioner take fewer PCG iterations in total (neither
    # reaches its Newton tolerance within the 10-iteration cap, so the states
    # they stop at differ)
    from dataclasses import replace

    def make():
        sc = S.crumple_fold(63)
        m = sc.material
        sc.material = replace(m, stretchWarp=m.stretchWarp * 10, stretchWeft=m.stretchWeft * 10, shear=m.shear * 10)
        return sc

    kw = dict(precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(64, 1, 64), pcg_max_iters=200)
    runs = {}
    for coarse in (_abi.CS_COARSE_NONE, _abi.CS_COARSE_PCG):
        sc = make()
        g = HipSolver(precision, coarse=coarse, **kw)
        g.initialize(sc)
        its = 0
        for _ in range(3):
            g.step(H)
            its += g.lastStats["pcg_iters"]
        assert np.all(np.isfinite(sc.positions))
        runs[coarse] = (its, sc.positions.copy(), g.lastStats["grad_norm"])
        g.close()
    (i1, x1, g1), (i2, x2, g2) = runs[_abi.CS_COARSE_NONE], runs[_abi.CS_COARSE_PCG]
    d = np.max(np.abs(x1 - x2))
    print(f"[coarse] {precision}: PCG iterations one-level {i1}, two-level {i2}; |grad| {g1:.2e} / {g2:.2e}; "
          f"max position difference {d:.2e} m", flush=True)
    assert i2 < i1
