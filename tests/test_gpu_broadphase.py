"""Standalone device broadphase (cs_broadphase_self_pairs: linear BVH with a
Karras hierarchy) vs the oracle's TriBvh.build + selfPairs (bvh.ts:19-124):
the sorted candidate-pair list must be identical, on small crumpled sheets
(also pinned by O(m^2) brute force), on the sphere drape (256x256, 130,050
triangles) and on the self-contact fold (512x512, 522,242 triangles) at their
first step, and on degenerate inputs."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from closim_amd import _abi, scenes as S
from closim_amd.devmem import DeviceArray
from closim_amd.mesh import build_grid
from closim_amd.solver import broadphase_self_pairs, load_library

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0


def brute_pairs(x0, x1, idx, pad):
    T = idx.reshape(-1, 3)
    X0 = x0.reshape(-1, 3)[T]
    X1 = x1.reshape(-1, 3)[T]
    lo = np.minimum(X0.min(1), X1.min(1)) - pad
    hi = np.maximum(X0.max(1), X1.max(1)) + pad
    out = []
    for a in range(len(T)):
        ov = np.all(lo[a][None] <= hi[a + 1:], 1) & np.all(hi[a][None] >= lo[a + 1:], 1)
        for b in np.flatnonzero(ov) + a + 1:
            if not set(T[a]) & set(T[b]):
                out.append((a, b))
    return np.array(out, np.uint32).reshape(-1, 2)


def predictor(scene):
    x0 = scene.positions.copy()
    x1 = (x0.reshape(-1, 3) + H * scene.velocities.reshape(-1, 3) + H * H * np.array([0.0, -9.81, 0.0]))
    return x0, x1.reshape(-1)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_small_crumpled_sheets_equal_oracle_and_brute_force(seed):
    rng = np.random.default_rng(seed)
    pos, uv, idx = build_grid(9, 7, 0.1, 0.08)
    X = pos.astype(np.float64).reshape(-1, 3)
    X[:, 1] = 0.02 * np.sin(40 * X[:, 0]) * np.cos(30 * X[:, 2])
    x0 = X.reshape(-1).copy()
    x1 = x0 + rng.normal(0, 0.004, x0.shape)
    for pad in (1e-4, 2e-3, 1e-2):
        got = broadphase_self_pairs(x0, x1, idx, pad)
        ref = O.self_pairs(x0, x1, idx, pad)
        assert np.array_equal(got, ref)
        assert np.array_equal(got, brute_pairs(x0, x1, idx, pad))


@pytest.mark.parametrize("name", ["sphere_drape_256", "crumple_fold_512"])
def test_config_size_pair_lists_equal_oracle(name):
    scene, pad = (S.sphere_drape(255), 1e-3) if name == "sphere_drape_256" else (S.crumple_fold(511), 5e-4)
    x0, x1 = predictor(scene)
    idx = scene.mesh.indices
    print(f"[broadphase] {name}: {len(idx) // 3} triangles, pad {pad}", flush=True)
    got, bms, qms = broadphase_self_pairs(x0, x1, idx, pad, timing=True)
    ref = O.self_pairs(x0, x1, idx, pad)
    print(f"[broadphase] {name}: {len(got)} pairs (oracle {len(ref)}), build {bms:.3f} ms, query {qms:.3f} ms",
          flush=True)
    assert len(ref) > 0
    assert np.array_equal(got, ref)
    assert bms > 0 and qms > 0


def test_random_soup_with_shared_and_duplicate_vertices():
    rng = np.random.default_rng(7)
    n, m = 900, 2000
    x0 = rng.uniform(0, 0.2, 3 * n)
    x0[3 * 5:3 * 40] = 0.05  # 35 coincident vertices
    x1 = x0 + rng.normal(0, 0.003, x0.shape)
    idx = rng.integers(0, n, 3 * m).astype(np.uint32)
    for pad in (0.0, 1e-3):
        got = broadphase_self_pairs(x0, x1, idx, pad)
        assert np.array_equal(got, O.self_pairs(x0, x1, idx, pad))


def test_all_triangles_at_one_point():
    # every Morton code equal: the hierarchy splits by index only
    m = 300
    idx = np.arange(3 * m, dtype=np.uint32)
    x0 = np.full(9 * m, 0.25)
    got = broadphase_self_pairs(x0, x0, idx, 1e-4)
    ref = O.self_pairs(x0, x0, idx, 1e-4)
    assert len(ref) == m * (m - 1) // 2
    assert np.array_equal(got, ref)


def test_tiny_and_nonfinite_inputs():
    pos, uv, idx = build_grid(3, 3, 0.1, 0.1)
    x0 = pos.astype(np.float64)
    x1 = x0.copy()
    x1[1::3] += 0.01
    for k in (0, 1, 2, 3):
        sub = idx[:3 * k]
        assert np.array_equal(broadphase_self_pairs(x0, x1, sub, 0.05), O.self_pairs(x0, x1, sub, 0.05))
    xn = x1.copy()
    xn[3 * 4 + 1] = np.nan
    xn[3 * 7:3 * 8] = np.inf
    assert np.array_equal(broadphase_self_pairs(x0, xn, idx, 0.05), O.self_pairs(x0, xn, idx, 0.05))


def test_overflow_reports_count_and_bad_index_is_rejected():
    L = load_library()
    pos, uv, idx = build_grid(9, 7, 0.1, 0.08)
    x0 = pos.astype(np.float64)
    x0[1::3] = 0.02 * np.sin(40 * x0[0::3])
    ref = O.self_pairs(x0, x0, idx, 2e-2)
    assert len(ref) > 4
    d0, di = DeviceArray(x0), DeviceArray(idx)
    out = DeviceArray(dtype=np.uint32, shape=(4, 2))
    cnt = C.c_uint64()
    rc = L.cs_broadphase_self_pairs(0, d0.ptr, d0.ptr, len(x0) // 3, di.ptr, len(idx) // 3, 2e-2, out.ptr, 4,
                                    C.byref(cnt), None, None)
    assert rc == _abi.CS_E_OVERFLOW and cnt.value == len(ref)
    bad = idx.copy()
    bad[10] = len(x0) // 3
    db = DeviceArray(bad)
    rc = L.cs_broadphase_self_pairs(0, d0.ptr, d0.ptr, len(x0) // 3, db.ptr, len(idx) // 3, 2e-2, out.ptr, 4,
                                    C.byref(cnt), None, None)
    assert rc == _abi.CS_E_INVALID
    assert b"out of range" in L.cs_last_error()
