"""Pin the oracle's implicit step against the reference's integration tests
(tests/smoke.test.ts, floor-contact, determinism, static-collider,
self-contact) and the host mesh mirror against the oracle's mesh.ts
restatement (bit-exact)."""
import math

import numpy as np
import pytest

import oracle_ffi as O
from closim_amd import _abi, mesh as M, scenes as S
from closim_amd.scene import DEFAULT_MATERIAL

H = 1.0 / 60.0


@pytest.mark.parametrize("dims", [(12, 6, 0.2, 0.1), (19, 99, 0.2, 1.0), (31, 17, 1.0, 0.5)])
def test_mesh_mirror_bit_exact_vs_oracle(dims):
    pos, uv, idx = M.build_grid(*dims)
    opos, ouv, oidx = O.build_grid(*dims)
    assert np.array_equal(pos, opos) and np.array_equal(uv, ouv) and np.array_equal(idx, oidx)
    # strip-pull style re-pose exercises non-flat positions for the hinges
    P = pos.reshape(-1, 3)
    P2 = np.stack([P[:, 0], -P[:, 2], 0.01 * np.sin(30 * P[:, 0])], 1).astype(np.float32).reshape(-1)
    for p in (pos, P2):
        m = M.preprocess(p, uv, idx, 0.2)
        o = O.preprocess(p, uv, idx, 0.2)
        for k in ("inv_dm", "areas", "masses", "hinge_verts", "hinge_rest", "hinge_len", "hinge_area"):
            a = getattr(m, k)
            assert np.array_equal(a, o[k]), k


def test_grid_mesh_counts_and_rest_angles():
    m = S.strip_pull().mesh
    assert (m.count, m.tri_count) == (2000, 3762)
    # interior edges of a 19x99 grid: 3*m edges total over 2 tris each + boundary
    assert m.hinge_count == 5525
    assert np.allclose(m.hinge_rest, math.pi, atol=1e-7)
    assert abs(m.masses.astype(float).sum() - 0.2 * 0.2 * 1.0) < 1e-6


def run_oracle(scene, steps, opts=None, per_step=None):
    s = O.OracleSolver(scene, opts)
    stats = []
    for k in range(steps):
        if per_step:
            per_step(s, k)
        s.step(H)
        assert np.all(np.isfinite(s.getPositions()))
        stats.append(s.lastStats)
    return s, stats


def test_strip_smoke_sags_without_overstretch():
    # tests/smoke.test.ts:16-48
    scene = S.small_strip()
    s, _ = run_oracle(scene, 60)
    pos = s.getPositions()
    tip = [j * 13 + 12 for j in range(7)]
    assert np.mean(pos[np.array(tip) * 3 + 1]) < -0.01
    _, _, max_strain = s.eval_internal(pos)
    assert max_strain < 0.05
    for i, p in scene.pinned.items():
        assert abs(pos[i * 3] - p[0]) < 1e-9 and abs(pos[i * 3 + 1] - p[1]) < 1e-9


def test_floor_contact_settles_without_penetration():
    # tests/floor-contact.test.ts:11-49
    scene = S.floor_patch()
    s, stats = run_oracle(scene, 90)
    fwork = sum(st["contact"]["friction_work"] for st in stats)
    pos = s.getPositions().reshape(-1, 3)
    vel = s.getVelocities().reshape(-1, 3)
    assert pos[:, 1].min() > -5e-4 and pos[:, 1].min() < 0.01
    assert np.linalg.norm(vel, axis=1).mean() < 0.5
    assert stats[-1]["contact"]["active_pairs"] > 0
    assert fwork <= 1e-9
    assert math.isfinite(stats[-1]["energy"])


def determinism_run():
    pos, uv, idx = S.offset_mesh(5, 5, 0.1, 0.1, 0, 0.04, 0.01)
    mesh = M.preprocess(pos, uv, idx, 0.15)
    from closim_amd.scene import create_scene, ContactSystem, DEFAULT_CONTACT_PARAMS
    from dataclasses import replace
    scene = create_scene(mesh, DEFAULT_MATERIAL)
    c = ContactSystem(replace(DEFAULT_CONTACT_PARAMS, frictionMu=0.3))
    c.setFloor(0.0)
    scene.contact = c
    s, _ = run_oracle(scene, 30)
    return s.getPositions().copy()


def test_determinism_bit_identical():
    # tests/determinism.test.ts:26-32
    a = determinism_run()
    b = determinism_run()
    assert np.array_equal(a, b)


def min_cloth_static(x, n, sp, si):
    from test_oracle_collision import closest_vt
    X = x.reshape(-1, 3)
    SP = sp.reshape(-1, 3).astype(np.float64)
    T = si.reshape(-1, 3)
    return min(closest_vt(X[v], SP[t[0]], SP[t[1]], SP[t[2]])[2] for v in range(n) for t in T)


def test_static_collider_drape_without_penetration():
    # tests/static-collider.test.ts:35-69
    scene = S.static_plate_patch()
    s, stats = run_oracle(scene, 60)
    assert max(st["contact"]["active_pairs"] for st in stats) > 0
    assert max(st["contact"]["vt_pairs"] for st in stats) > 0
    c = scene.contact
    d = min_cloth_static(s.getPositions(), scene.mesh.count, c.staticPos, c.staticIdx)
    assert 5e-5 < d < 2e-3


def audit_min_distance(x, mesh):
    """auditMinDistance() tests/helpers.ts:47-102 (brute force, adjacency excluded)."""
    from test_oracle_collision import closest_vt, closest_ee
    X = x.reshape(-1, 3)
    T = mesh.indices.reshape(-1, 3)
    m = len(T)
    best = math.inf
    for a in range(m):
        for b in range(a + 1, m):
            if set(T[a]) & set(T[b]):
                continue
            for p in T[a]:
                best = min(best, closest_vt(X[p], X[T[b][0]], X[T[b][1]], X[T[b][2]])[2])
            for p in T[b]:
                best = min(best, closest_vt(X[p], X[T[a][0]], X[T[a][1]], X[T[a][2]])[2])
    edges = set()
    for t in T:
        for e in range(3):
            u, v = int(t[e]), int(t[(e + 1) % 3])
            edges.add((min(u, v), max(u, v)))
    edges = sorted(edges)
    for i in range(len(edges)):
        for j in range(i + 1, len(edges)):
            a, b = edges[i]
            c, d = edges[j]
            if len({a, b, c, d}) < 4:
                continue
            best = min(best, closest_ee(X[a], X[b], X[c], X[d])[2])
    return best


def test_head_on_patches_no_penetration():
    # tests/self-contact.test.ts:11-38
    scene = S.head_on_patches()
    s, stats = run_oracle(scene, 60)
    assert max(st["contact"]["active_pairs"] for st in stats) > 0
    assert audit_min_distance(s.getPositions(), scene.mesh) > 1e-4
    assert math.isfinite(stats[-1]["energy"])


def test_strip_pull_pins_exact():
    scene = S.strip_pull()
    opts = _abi.options(grad_tol=1e-6)

    def drive(s, k):
        for i, p in S.strip_pull_targets(scene, k):
            s.pinVertex(i, p)
    s, stats = run_oracle(scene, 3, opts, drive)
    for st in stats:
        assert st["newton_used"] >= 1
    pos = s.getPositions()
    for i, p in S.strip_pull_targets(scene, 2):
        assert tuple(pos[i * 3:i * 3 + 3]) == p
