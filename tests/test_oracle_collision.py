"""Pin the oracle's collision primitives: closest points, cubic roots, CCD,
barrier, friction and the broadphase candidate set."""
import math

import numpy as np
import pytest

import oracle_ffi as O

L = O.lib()
DMIN = 1e-4
ORC_FAULT_CCD_NO_FALLBACK = 2
ORC_FAULT_BARRIER_SIGN = 3
ORC_FAULT_EE_TRANSPOSE = 4


def vt_ccd(x0, x1, p=3, a=0, b=1, c=2):
    return L.orc_vt_ccd(O.P(O.d3(x0)), O.P(O.d3(x1)), p, a, b, c, DMIN)


def ee_ccd(x0, x1, a=0, b=1, c=2, d=3):
    return L.orc_ee_ccd(O.P(O.d3(x0)), O.P(O.d3(x1)), a, b, c, d, DMIN)


# ---- known answers: tests/ccd.test.ts:12-63 ------------------------------
def test_vt_ccd_tunnelling_toi_half():
    toi = vt_ccd([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 0.01], [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, -0.01])
    assert 0 <= toi < 1 and abs(toi - 0.5) < 0.05


def test_vt_ccd_stop_short_is_safe():
    assert vt_ccd([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 0.01],
                  [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 0.005]) == math.inf


def test_vt_ccd_resting_contact_toi_zero():
    base = [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.2, 0.2, 0.00005]
    assert vt_ccd(base, base) == 0


def test_vt_ccd_outside_triangle_is_safe():
    assert vt_ccd([0, 0, 0, 1, 0, 0, 0, 1, 0, 5.0, 5.0, 0.01],
                  [0, 0, 0, 1, 0, 0, 0, 1, 0, 5.0, 5.0, -0.01]) == math.inf


def test_ee_ccd_crossing_toi_half():
    toi = ee_ccd([-1, 0, 0, 1, 0, 0, 0, 0.01, -1, 0, 0.01, 1], [-1, 0, 0, 1, 0, 0, 0, -0.01, -1, 0, -0.01, 1])
    assert 0 <= toi < 1 and abs(toi - 0.5) < 0.1


def test_ee_ccd_parallel_is_safe():
    assert ee_ccd([-1, 0, 0, 1, 0, 0, -1, 0.5, -1, 1, 0.5, 1], [-1, 0, 0, 1, 0, 0, -1, 0.4, -1, 1, 0.4, 1]) == math.inf


def test_vt_ccd_grazing_needs_the_sampled_fallback():
    # The vertex descends from 3e-4 to 5e-5 above the face without crossing
    # its plane: V(t) has no root in [0,1], so only the 17-sample fallback
    # (ccd-vt.ts:99-114) sees the approach within d_min, at t = 1.
    x0 = [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.3, 0.3, 3e-4]
    x1 = [0, 0, 0, 1, 0, 0, 0, 1, 0, 0.3, 0.3, 5e-5]
    assert vt_ccd(x0, x1) == 1.0
    L.orc_set_fault(ORC_FAULT_CCD_NO_FALLBACK)
    try:
        assert vt_ccd(x0, x1) == math.inf  # the known answer catches the planted omission
    finally:
        L.orc_set_fault(0)


def closest_vt(p, a, b, c):
    out = np.zeros(6)
    L.orc_closest_vt(*[O.P(O.d3(v)) for v in (p, a, b, c)], O.P(out))
    return out


def closest_ee(a, b, c, d):
    out = np.zeros(6)
    L.orc_closest_ee(*[O.P(O.d3(v)) for v in (a, b, c, d)], O.P(out))
    return out


def test_closest_vt_matches_dense_sampling():
    rng = np.random.default_rng(7)
    g = np.linspace(0, 1, 201)
    S, T = np.meshgrid(g, g)
    inside = S + T <= 1
    S, T = S[inside], T[inside]
    for _ in range(60):
        p, a, b, c = rng.standard_normal((4, 3))
        s, t, dist, rx, ry, rz = closest_vt(p, a, b, c)
        q = a + s * (b - a) + t * (c - a)
        assert abs(np.linalg.norm(p - q) - dist) < 1e-12
        assert np.allclose([rx, ry, rz], p - q, atol=1e-12)
        Q = a[None] + S[:, None] * (b - a)[None] + T[:, None] * (c - a)[None]
        brute = np.linalg.norm(Q - p[None], axis=1).min()
        assert dist <= brute + 1e-12 and dist > brute - 2e-2 * (1 + brute)


def check_ee_sampling(rng, cases=60):
    g = np.linspace(0, 1, 401)
    worst = 0.0
    for _ in range(cases):
        a, b, c, d = rng.standard_normal((4, 3))
        s, t, dist, *_ = closest_ee(a, b, c, d)
        A = a[None] + g[:, None] * (b - a)[None]
        Cc = c[None] + g[:, None] * (d - c)[None]
        brute = np.linalg.norm(A[:, None, :] - Cc[None, :, :], axis=2).min()
        worst = max(worst, dist - brute)
    return worst


def test_closest_ee_matches_dense_sampling():
    # the oracle distance may not exceed the sampled minimum
    assert check_ee_sampling(np.random.default_rng(9)) <= 1e-12


def test_closest_ee_check_detects_planted_transpose():
    L.orc_set_fault(ORC_FAULT_EE_TRANSPOSE)
    try:
        assert check_ee_sampling(np.random.default_rng(9)) > 1e-6
    finally:
        L.orc_set_fault(0)


def test_cubic_roots_match_companion_matrix():
    rng = np.random.default_rng(3)
    roots = np.zeros(3)
    checked = 0
    for _ in range(400):
        r = np.sort(rng.uniform(-0.5, 1.5, 3))
        scale = 10 ** rng.uniform(-3, 3)
        c3, c2, c1, c0 = scale * np.poly(r)
        k = L.orc_solve_cubic01(c3, c2, c1, c0, O.P(roots))
        got = np.sort(roots[:k])
        ref = np.sort(np.clip([x.real for x in np.roots([c3, c2, c1, c0]) if abs(x.imag) < 1e-9
                               and -1e-9 <= x.real <= 1 + 1e-9], 0, 1))
        # roots closer than 1e-9 are deduplicated by the solver (cubic.ts:48)
        if len(ref) > 1 and np.min(np.diff(ref)) < 1e-6:
            continue
        assert len(got) == len(ref)
        assert np.allclose(got, ref, atol=1e-8)
        checked += 1
    assert checked > 300
    # quadratic / linear degenerations (cubic.ts:51-66)
    assert L.orc_solve_cubic01(0, 1, -1, 0.21, O.P(roots)) == 2
    assert np.allclose(roots[:2], [0.3, 0.7])
    assert L.orc_solve_cubic01(0, 0, 2, -1, O.P(roots)) == 1 and roots[0] == 0.5
    assert L.orc_solve_cubic01(0, 0, 0, 1, O.P(roots)) == 0


# ---- barrier / friction: tests/barrier-check.test.ts, friction.test.ts ----
def test_barrier_zero_beyond_dhat_and_blows_up():
    assert L.orc_barrier_value(0.005, 0.002) == 0
    assert L.orc_barrier_value(0.002, 0.002) == 0
    assert L.orc_barrier_value(0.001, 0.002) > 0
    assert L.orc_barrier_value(1e-9, 0.002) > L.orc_barrier_value(0.001, 0.002)


def barrier_fd_err():
    worst = 0.0
    for d in np.linspace(1e-4, 1.9e-3, 40):
        eps = 1e-9
        num = 50 * (L.orc_barrier_value(d + eps, 2e-3) - L.orc_barrier_value(d - eps, 2e-3)) / (2 * eps)
        ana = L.orc_barrier_grad(d, 2e-3, 50, 1)
        worst = max(worst, abs(num - ana) / (abs(num) + 1e-12))
    return worst


def test_barrier_gradient_matches_fd():
    assert barrier_fd_err() < 1e-6


def test_barrier_gradient_check_detects_planted_sign_error():
    L.orc_set_fault(ORC_FAULT_BARRIER_SIGN)
    try:
        assert barrier_fd_err() > 1e-6
    finally:
        L.orc_set_fault(0)


def test_coulomb_invariants():
    f = np.zeros(3)
    n = np.array([0.0, 1.0, 0.0])
    L.orc_coulomb(O.P(O.d3([0.01, 0.002, -0.004])), O.P(n), 2.5, 0.3, 1e-4, O.P(f))
    mag = np.linalg.norm(f)
    assert abs(f @ n) <= 1e-9 * (1 + mag)
    assert mag <= 0.3 * 2.5 + 1e-12
    L.orc_coulomb(O.P(O.d3([0, 0, 0])), O.P(n), 2.5, 0.3, 1e-4, O.P(f))
    assert np.linalg.norm(f) < 1e-9


# ---- broadphase: candidate set == O(m^2) brute force ----------------------
def brute_pairs(x0, x1, idx, pad):
    T = idx.reshape(-1, 3)
    X0 = x0.reshape(-1, 3)[T]
    X1 = x1.reshape(-1, 3)[T]
    lo = np.minimum(X0.min(1), X1.min(1)) - pad
    hi = np.maximum(X0.max(1), X1.max(1)) + pad
    m = len(T)
    out = []
    for a in range(m):
        ov = np.all(lo[a][None] <= hi[a + 1:], 1) & np.all(hi[a][None] >= lo[a + 1:], 1)
        for b in np.flatnonzero(ov) + a + 1:
            if len(set(T[a]) & set(T[b])):
                continue
            out.append((a, b))
    return np.array(out, np.uint32).reshape(-1, 2)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_self_pairs_equal_brute_force(seed):
    from closim_amd.mesh import build_grid
    rng = np.random.default_rng(seed)
    pos, uv, idx = build_grid(9, 7, 0.1, 0.08)
    x0 = pos.astype(np.float64)
    X = x0.reshape(-1, 3)
    X[:, 1] = 0.02 * np.sin(40 * X[:, 0]) * np.cos(30 * X[:, 2])   # crumple
    x0 = X.reshape(-1).copy()
    x1 = x0 + rng.normal(0, 0.004, x0.shape)
    for pad in (1e-4, 2e-3, 1e-2):
        got = O.self_pairs(x0, x1, idx, pad)
        ref = brute_pairs(x0, x1, idx, pad)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("name", ["stacked", "floor", "crumpled"])
def test_frozen_contact_hvp_is_the_derivative_of_the_frozen_contact_gradient(name):
    # applyHvp (contact-assembly.ts:371-392) is kappa b''(d) (J.p) J^T for the
    # frozen contacts, and the frozen distance is linear in x (frozen normal
    # and weights), so without friction it is exactly the derivative of the
    # frozen contact gradient: central differences along p agree to
    # truncation (b is smooth inside dHat) and the planted sign is caught
    from dataclasses import replace
    from closim_amd import _abi
    from closim_amd.scene import ContactSystem
    import contact_scenes as CS
    sc = {"stacked": lambda: CS.stacked_patches(mu=0.0), "floor": lambda: CS.floor_close(),
          "crumpled": lambda: CS.crumpled_sheet()}[name]()
    c = sc.contact
    params = replace(c.params, frictionMu=0.0)
    nc = ContactSystem(params)
    if c.floorY is not None:
        nc.setFloor(c.floorY)
    sc.contact = nc
    o = O.OracleSolver(sc, _abi.options())
    x = sc.positions.copy()
    rng = np.random.default_rng(11)
    p = rng.standard_normal(len(x))
    hvp, k = o.contact_hvp(x, x, p)
    assert k > 0 and np.any(hvp != 0)
    eps = 1e-9
    gp, _ = o.contact_grad_frozen(x, x, x + eps * p)
    gm, _ = o.contact_grad_frozen(x, x, x - eps * p)
    fd = (gp - gm) / (2 * eps)
    err = np.max(np.abs(fd - hvp)) / np.max(np.abs(hvp))
    print(f"{name}: {k} active, FD rel err {err:.1e}")
    assert err < 1e-5
    assert np.max(np.abs(fd + hvp)) / np.max(np.abs(hvp)) > 1.0  # a flipped sign fails
