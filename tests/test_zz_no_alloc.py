"""cs_broadphase_self_pairs_noalloc (the allocation-free twin of
cs_broadphase_self_pairs) returns exactly what the allocating entry point
returns, on two calls in a row (the second on a smaller mesh, so it runs over
the previous call's scratch contents), and rejects a mesh beyond its fixed
scratch with a clear error."""
import ctypes as C

import numpy as np
import pytest

from closim_amd import _abi, scenes as S
from closim_amd.devmem import DeviceArray
from closim_amd.solver import load_library

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0
MAX_TRIS = 1 << 19


def predictor(scene):
    x0 = scene.positions.copy()
    x1 = x0.reshape(-1, 3) + H * scene.velocities.reshape(-1, 3) + H * H * np.array([0.0, -9.81, 0.0])
    return x0, x1.reshape(-1)


def both(x0, x1, idx, pad):
    """(pairs, count) of the allocating and of the allocation-free entry point."""
    L = load_library()
    n, m = len(x0) // 3, len(idx) // 3
    d0, d1, di = DeviceArray(x0), DeviceArray(x1), DeviceArray(idx)
    cap = 4 * m
    res = []
    for name in ("cs_broadphase_self_pairs", "cs_broadphase_self_pairs_noalloc"):
        out = DeviceArray(dtype=np.uint32, shape=(cap, 2))
        cnt = C.c_uint64()
        args = [0, d0.ptr, d1.ptr, n, di.ptr, m, pad, out.ptr, cap, C.byref(cnt)]
        if name == "cs_broadphase_self_pairs":
            args += [None, None]
        rc = getattr(L, name)(*args)
        assert rc == _abi.CS_OK, (name, L.cs_last_error())
        res.append((out.numpy(cnt.value), cnt.value))
        out.free()
    return res


def test_noalloc_matches_allocating_entry_on_two_calls():
    fold = S.crumple_fold(511)       # 522,242 triangles (the limit is 524,288)
    drape = S.sphere_drape(255)      # 130,050 triangles
    for name, scene, pad in (("fold 512x512", fold, 5e-4), ("drape 256x256", drape, 1e-3)):
        x0, x1 = predictor(scene)
        idx = np.ascontiguousarray(scene.mesh.indices, np.uint32)
        print(f"[no-alloc] {name}: {len(idx) // 3} triangles", flush=True)
        (a, ca), (b, cb) = both(x0, x1, idx, pad)
        assert ca == cb and ca > 0
        assert np.array_equal(a, b)
        print(f"[no-alloc] {name}: {cb} pairs identical", flush=True)


def test_noalloc_rejects_meshes_beyond_its_scratch():
    L = load_library()
    cnt = C.c_uint64(5)
    p = C.c_void_p(8)
    rc = L.cs_broadphase_self_pairs_noalloc(0, p, p, 3, p, MAX_TRIS + 1, 1e-3, p, 1, C.byref(cnt))
    assert rc == _abi.CS_E_INVALID and cnt.value == 0
    assert b"at most 524288 triangles" in L.cs_last_error()
