"""Pins the oracle's additive-Schwarz preconditioner (schwarz.ts:44-309).

The partition is checked through its invariants (every vertex in exactly one
domain, local ids ranked by vertex id, each domain connected in the mesh
graph, the size bounds growth and merging guarantee).  The apply is checked
against an independent dense solve: the global membrane Hessian is assembled
in numpy from the element 9x9 blocks (orc_tri_hessian9, pinned by FD columns
in test_oracle_elements.py), restricted to each domain's dofs, shifted by the
Jacobi diagonal and solved with numpy.linalg.  A planted transposed index in
the back substitution must fail that check; a diagonal that makes the local
matrices indefinite must select the Jacobi fallback.
"""
from collections import deque

import numpy as np
import pytest

import oracle_ffi as O
from closim_amd import scenes as S

L = O.lib()
ORC_FAULT_SCHWARZ_BACK_ROW = 5
H = 1.0 / 60.0


def deformed(scene, seed, stretch=1.02, amp=1e-4):
    """A stretched, slightly perturbed state: the membrane Hessian in tension
    is positive definite, so every local Cholesky succeeds."""
    rng = np.random.default_rng(seed)
    return scene.positions * stretch + amp * rng.standard_normal(scene.positions.shape)


def global_hessian(scene, x):
    m = scene.mesh
    n3 = m.count * 3
    A = np.zeros((n3, n3))
    c = scene.material.to_c()
    idx = m.indices.reshape(-1, 3)
    for t, (a, b, d) in enumerate(idx):
        out = np.zeros(81)
        xs = [np.ascontiguousarray(x[v * 3:v * 3 + 3]) for v in (a, b, d)]
        inv = np.asarray(m.inv_dm[t * 4:t * 4 + 4], np.float64)
        L.orc_tri_hessian9(*[O.P(O.d3(q)) for q in (*xs, inv)], float(m.areas[t]), O.C.byref(c), O.P(out))
        h = out.reshape(9, 9)
        dofs = np.concatenate([np.arange(v * 3, v * 3 + 3) for v in (a, b, d)])
        A[np.ix_(dofs, dofs)] += h
    return A


def jacobi_like_diag(scene, x, A):
    masses = np.repeat(np.asarray(scene.mesh.masses, np.float64), 3)
    return masses / (H * H) + np.diag(A)


def adjacency(mesh):
    adj = [set() for _ in range(mesh.count)]
    for a, b, c in mesh.indices.reshape(-1, 3):
        adj[a] |= {b, c}; adj[b] |= {a, c}; adj[c] |= {a, b}
    return adj


@pytest.mark.parametrize("make,sizes", [
    (S.small_strip, (16, 8, 32)),
    (S.small_strip, (5, 3, 10)),
    (S.strip_pull, (16, 8, 32)),
    (S.strip_pull, (64, 1, 64)),
])
def test_partition_invariants(make, sizes):
    scene = make()
    o = O.OracleSolver(scene)
    dom, loc, cnt = o.schwarz_domains(*sizes)
    n = scene.mesh.count
    target, mn, _ = sizes
    assert dom.max() == cnt - 1 and len(np.unique(dom)) == cnt
    adj = adjacency(scene.mesh)
    for d in range(cnt):
        mem = np.flatnonzero(dom == d)
        # localOf is the rank of the vertex id inside the sorted member list
        assert np.array_equal(loc[mem], np.arange(len(mem)))
        # domains grow by BFS and merge into adjacent domains: connected
        seen = {int(mem[0])}
        q = deque([int(mem[0])])
        ms = set(int(v) for v in mem)
        while q:
            u = q.popleft()
            for w in adj[u]:
                if w in ms and w not in seen:
                    seen.add(w); q.append(w)
        assert seen == ms, d
        if n >= mn:
            assert len(mem) >= mn
        assert len(mem) < target + max(mn, 1)


def test_partition_target_covers_mesh_is_one_domain():
    scene = S.small_strip()
    n = scene.mesh.count
    dom, loc, cnt = O.OracleSolver(scene).schwarz_domains(n, 1, n)
    assert cnt == 1 and np.array_equal(loc, np.arange(n))


def test_partition_singletons():
    scene = S.small_strip()
    dom, loc, cnt = O.OracleSolver(scene).schwarz_domains(1, 0, 1)
    assert cnt == scene.mesh.count and np.all(loc == 0) and len(np.unique(dom)) == cnt


def schwarz_reference(A, diag, dom, r):
    z = np.zeros_like(r)
    K = A + np.diag(diag)
    for d in range(dom.max() + 1):
        dofs = np.concatenate([np.arange(v * 3, v * 3 + 3) for v in np.flatnonzero(dom == d)])
        z[dofs] = np.linalg.solve(K[np.ix_(dofs, dofs)], r[dofs])
    return z


@pytest.mark.parametrize("sizes", [(16, 8, 32), (1, 0, 1), (91, 1, 91), (7, 3, 14)])
def test_apply_equals_dense_domain_solves(sizes):
    scene = S.small_strip()
    o = O.OracleSolver(scene)
    x = deformed(scene, 3)
    A = global_hessian(scene, x)
    diag = jacobi_like_diag(scene, x, A)
    r = np.random.default_rng(4).standard_normal(len(x))
    z, flags = o.schwarz_apply(x, diag, r, *sizes)
    assert np.all(flags == 1)
    dom, _, _ = o.schwarz_domains(*sizes)
    ref = schwarz_reference(A, diag, dom, r)
    err = np.max(np.abs(z - ref)) / np.max(np.abs(ref))
    assert err < 1e-11, err


def test_apply_detects_planted_back_substitution_transpose():
    scene = S.small_strip()
    o = O.OracleSolver(scene)
    x = deformed(scene, 5)
    A = global_hessian(scene, x)
    diag = jacobi_like_diag(scene, x, A)
    r = np.random.default_rng(6).standard_normal(len(x))
    dom, _, _ = o.schwarz_domains(16, 8, 32)
    ref = schwarz_reference(A, diag, dom, r)
    L.orc_set_fault(ORC_FAULT_SCHWARZ_BACK_ROW)
    try:
        z, _ = o.schwarz_apply(x, diag, r, 16, 8, 32)
    finally:
        L.orc_set_fault(0)
    assert np.max(np.abs(z - ref)) / np.max(np.abs(ref)) > 1e-3


def test_apply_jacobi_fallback_on_indefinite_domains():
    scene = S.small_strip()
    o = O.OracleSolver(scene)
    x = deformed(scene, 7)
    n3 = len(x)
    diag = np.full(n3, -1.0)
    diag[::7] = 0.0           # 1/dd only where dd > 1e-12, else 0
    diag[1::5] = 3.0
    r = np.random.default_rng(8).standard_normal(n3)
    z, flags = o.schwarz_apply(x, diag, r, 16, 8, 32)
    assert np.all(flags == 0)
    # the fallback stores 1/dd and multiplies (schwarz.ts:244, 272-276)
    expect = r * np.where(diag > 1e-12, 1.0 / np.where(diag > 1e-12, diag, 1.0), 0.0)
    assert np.array_equal(z, expect)
