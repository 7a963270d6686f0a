"""Device contact system vs the oracle: primitives, candidate/active
multisets (bit-exact, in the reference's active-list order), trial validity
and full contact trajectories."""
import math

import numpy as np
import pytest

import oracle_ffi as O
from closim_amd import _abi, scenes as S
from closim_amd.solver import HipSolver, probe_ccd, probe_closest
from contact_scenes import ALL, crumpled_sheet, stacked_patches

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0


def oracle_closest(kind, P):
    out = np.zeros(6)
    f = O.lib().orc_closest_vt if kind == 0 else O.lib().orc_closest_ee
    f(*[O.P(O.d3(P[s])) for s in range(4)], O.P(out))
    return out


def random_prims(rng, count):
    x0 = rng.standard_normal(count * 12) * 0.01
    x1 = x0 + rng.standard_normal(count * 12) * 0.004
    kinds = rng.integers(0, 2, count).astype(np.int32)
    ids = np.arange(count * 4, dtype=np.int32).reshape(count, 4)
    return x0, x1, kinds, ids


def test_closest_points_bit_exact_vs_oracle():
    rng = np.random.default_rng(21)
    x0, _, kinds, ids = random_prims(rng, 4000)
    got = probe_closest(x0, kinds, ids)
    X = x0.reshape(-1, 3)
    for i in range(len(kinds)):
        ref = oracle_closest(kinds[i], X[ids[i]])
        assert np.array_equal(got[i], ref[:3]), (i, got[i], ref[:3])


def test_ccd_known_answers_and_bit_exact_vs_oracle():
    # tests/ccd.test.ts:12-63 known answers
    kats = [
        ([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 0.01], [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, -0.01], 0, [3, 0, 1, 2]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 0.01], [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 0.005], 0, [3, 0, 1, 2]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 5e-5], [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 5e-5], 0, [3, 0, 1, 2]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0, 5, 5, 0.01], [0, 0, 0, 1, 0, 0, 0, 1, 0, 5, 5, -0.01], 0, [3, 0, 1, 2]),
        ([-1, 0, 0, 1, 0, 0, 0, 0.01, -1, 0, 0.01, 1], [-1, 0, 0, 1, 0, 0, 0, -0.01, -1, 0, -0.01, 1], 1, [0, 1, 2, 3]),
        ([-1, 0, 0, 1, 0, 0, -1, 0.5, -1, 1, 0.5, 1], [-1, 0, 0, 1, 0, 0, -1, 0.4, -1, 1, 0.4, 1], 1, [0, 1, 2, 3]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.3, 0.3, 3e-4], [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.3, 0.3, 5e-5], 0, [3, 0, 1, 2]),
    ]
    expect = [lambda t: abs(t - 0.5) < 0.05, lambda t: t == math.inf, lambda t: t == 0, lambda t: t == math.inf,
              lambda t: abs(t - 0.5) < 0.1, lambda t: t == math.inf, lambda t: t == 1.0]
    for (a, b, k, ids), ok in zip(kats, expect):
        t = probe_ccd(np.array(a, float), np.array(b, float), [k], [ids], 1e-4)[0]
        assert ok(t), (a, b, t)
    rng = np.random.default_rng(5)
    x0, x1, kinds, ids = random_prims(rng, 20000)
    got = probe_ccd(x0, x1, kinds, ids, 1e-3)
    for i in range(len(kinds)):
        f = O.lib().orc_vt_ccd if kinds[i] == 0 else O.lib().orc_ee_ccd
        ref = f(O.P(x0), O.P(x1), *[int(v) for v in ids[i]], 1e-3)
        assert got[i] == ref or (math.isnan(got[i]) and math.isnan(ref)), (i, got[i], ref)
    assert np.isfinite(got).sum() > 100  # the sample exercises real impacts


def test_moving_collider_trajectory_matches_oracle():
    """A moving rigid collider: the static plate of static-collider.test.ts,
    once the patch rests on it, slides 0.2 mm in x and 0.1 mm in z per step,
    attached before every step through ContactSystem.setStaticMesh
    (contact-assembly.ts:71-73), the way avatar-collider.ts:205-230 attaches
    a sampled avatar mesh.  (The static mesh is not swept within a step, so
    motion that closes the gap by more than it is between steps would start
    a step inside d_min, in the reference as here.)  The FP64 HipSolver
    follows the oracle bit for bit, the contacts hold on every step and the
    patch stays above the plate by more than d_min."""
    sg, so = S.static_plate_patch(), S.static_plate_patch()
    g = HipSolver("fp64")
    g.initialize(sg)
    o = O.OracleSolver(so, _abi.options())
    base = sg.contact.staticPos.reshape(-1, 3).astype(np.float64)
    idx = sg.contact.staticIdx.copy()
    for k in range(45):
        pos = (base + np.array([2e-4, 0.0, 1e-4]) * max(0, k - 14)).astype(np.float32).reshape(-1)
        sg.contact.setStaticMesh(pos, idx)
        o.set_static_mesh(pos, idx)
        g.step(H)
        o.step(H)
        sgs, sos = g.lastStats, o.lastStats
        assert np.array_equal(sg.positions, so.positions), k
        assert np.array_equal(sg.velocities, so.velocities), k
        for key in ("newton_used", "pcg_iters", "line_search_trials"):
            assert sgs[key] == sos[key], (k, key)
        assert sgs["contact"]["active_pairs"] == sos["contact"]["active_pairs"]
        assert sgs["contact"]["min_distance"] == sos["contact"]["min_distance"]
        if k >= 14:
            assert sgs["contact"]["active_pairs"] > 0, k
            assert sgs["contact"]["min_distance"] > sg.contact.params.dMinM, k
    top = float(pos.reshape(-1, 3)[:, 1].max())
    assert float(np.min(sg.positions[1::3])) - top > sg.contact.params.dMinM


def test_static_mesh_topology_is_fixed():
    sg = S.static_plate_patch()
    g = HipSolver("fp64")
    g.initialize(sg)
    from closim_amd.solver import ClothError
    sg.contact.setStaticMesh(sg.contact.staticPos, np.array([0, 1, 2], np.uint32))
    with pytest.raises(ClothError, match="fixed at creation"):
        g.step(H)


def test_device_ccd_queries_match_oracle_on_expanded_candidates():
    """cs_ccd_queries (device arrays) on the VT/EE prims of a crumpled sheet's
    candidate pairs, bit-exact against the oracle's vtCCD / eeCCD; an id out
    of range is rejected."""
    from closim_amd.solver import broadphase_self_pairs, ccd_queries, expand_pairs
    from closim_amd.mesh import build_grid
    rng = np.random.default_rng(11)
    pos, uv, idx = build_grid(15, 13, 0.1, 0.08)
    X = pos.astype(np.float64).reshape(-1, 3)
    X[:, 1] = 0.01 * np.sin(60 * X[:, 0]) * np.cos(50 * X[:, 2])
    x0 = X.reshape(-1).copy()
    x1 = x0 + rng.normal(0, 0.003, x0.shape)
    pairs = broadphase_self_pairs(x0, x1, idx, 1e-3)
    kinds, ids = expand_pairs(pairs, idx)
    assert len(kinds) == 15 * len(pairs) > 1000
    got, ms = ccd_queries(x0, x1, kinds, ids, 1e-4, timing=True)
    assert ms > 0
    for i in range(len(kinds)):
        f = O.lib().orc_vt_ccd if kinds[i] == 0 else O.lib().orc_ee_ccd
        ref = f(O.P(x0), O.P(x1), *[int(v) for v in ids[i]], 1e-4)
        assert got[i] == ref or (math.isnan(got[i]) and math.isnan(ref)), (i, got[i], ref)
    assert (got < math.inf).sum() > 10
    bad = ids.copy()
    bad[3, 2] = len(x0) // 3
    with pytest.raises(Exception, match="out of range"):
        ccd_queries(x0, x1, kinds, bad, 1e-4)


@pytest.mark.parametrize("name", sorted(ALL))
def test_active_set_bit_exact_and_ordered(name):
    scene = ALL[name]()
    g = HipSolver("fp64")
    g.initialize(scene)
    o = O.OracleSolver(ALL[name]())
    rng = np.random.default_rng(7)
    x0 = scene.positions.copy()
    x = x0 + rng.normal(0, 2e-4, x0.shape)
    gi, gv, gc = g.probe_active_set(x0, x)
    oi, ov, oc = o.active_set(x0, x)
    print(f"{name}: {len(gi)} active, {gc} candidates", flush=True)
    assert gc == oc
    assert len(gi) == len(oi) and len(gi) > 0
    assert np.array_equal(gi, oi)
    assert np.array_equal(gv, ov)


@pytest.mark.parametrize("name", sorted(ALL))
def test_check_trial_matches_oracle(name):
    scene = ALL[name]()
    g = HipSolver("fp64")
    g.initialize(scene)
    o = O.OracleSolver(ALL[name]())
    rng = np.random.default_rng(11)
    x0 = scene.positions.copy()
    seen = set()
    for k in range(12):
        scale = [1e-5, 1e-4, 5e-4, 2e-3][k % 4]
        xt = x0 + rng.normal(0, scale, x0.shape)
        vg, dg, tg = g.probe_check_trial(x0, xt)
        vo, do, to = o.check_trial(x0, xt)
        assert vg == vo, (k, vg, vo)
        if vg:
            assert dg == do and tg == to
        seen.add(vg)
    assert seen == {True, False}


def run_pair(make, steps, opts=None):
    sg, so = make(), make()
    g = HipSolver("fp64", **(opts or {}))
    g.initialize(sg)
    o = O.OracleSolver(so, _abi.options(**(opts or {})))
    for k in range(steps):
        g.step(H)
        o.step(H)
        a, b = g.getPositions(), o.getPositions()
        e = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        st_g, st_o = g.lastStats, o.lastStats
        print(f"step {k}: rel err {e:.3e} active {st_g['contact']['active_pairs']}/{st_o['contact']['active_pairs']}"
              f" pcg {st_g['pcg_iters']}/{st_o['pcg_iters']}", flush=True)
        assert e <= 1e-9, (k, e)
        for key in ("pcg_iters", "newton_used", "line_search_trials"):
            assert st_g[key] == st_o[key], (k, key)
        for key in ("active_pairs", "vt_pairs", "ee_pairs", "floor_pairs", "candidate_pairs",
                    "rejected_line_search_steps", "ccd_failures"):
            assert st_g["contact"][key] == st_o["contact"][key], (k, key)
        assert np.array_equal(g.getVelocities(), o.getVelocities())
    return g, o


def test_floor_patch_trajectory_matches_oracle():
    run_pair(S.floor_patch, 40)


def test_head_on_patches_trajectory_matches_oracle():
    run_pair(S.head_on_patches, 30)


def test_static_plate_trajectory_matches_oracle():
    run_pair(S.static_plate_patch, 40)


def test_crumpled_fold_trajectory_matches_oracle():
    run_pair(crumpled_sheet, 8)


def test_stacked_patches_block_jacobi_matches_oracle():
    run_pair(stacked_patches, 8, dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI, hvp_mode=_abi.CS_HVP_ANALYTIC))


def test_floor_patch_schwarz_matches_oracle():
    run_pair(S.floor_patch, 20, dict(precond=_abi.CS_PRECOND_SCHWARZ))


def test_crumpled_fold_schwarz_matches_oracle():
    run_pair(crumpled_sheet, 6, dict(precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(64, 1, 64)))
