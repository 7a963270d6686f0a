"""C ABI: the library loads on the CPU, exports exactly what include/ declares
and rejects bad arguments before touching a device."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from closim_amd import _abi
from closim_amd.solver import load_library, LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char\s*\*|const char\s*\*)\s*\**\s*(cs_\w+)\s*\(",
                                    txt, re.M))
    return names


def test_header_declares_the_clothsolver_surface():
    names = declared_symbols()
    for n in ("cs_create", "cs_step", "cs_set_material", "cs_pin_vertex", "cs_unpin_vertex",
              "cs_get_positions", "cs_get_velocities", "cs_last_stats", "cs_last_error", "cs_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = declared_symbols() - exported
    assert not missing, missing
    for n in declared_symbols():
        assert hasattr(L, n)


def test_struct_sizes_match_header(tmp_path):
    src = str(tmp_path / "abi_sizes.c")
    exe = src[:-2]
    with open(src, "w") as f:
        f.write('#include <stdio.h>\n#include "cloth_hip.h"\nint main(){printf("%zu %zu %zu %zu %zu %zu\\n",'
                'sizeof(cs_material),sizeof(cs_contact_params),sizeof(cs_scene_desc),sizeof(cs_options),'
                'sizeof(cs_contact_diag),sizeof(cs_step_stats));return 0;}\n')
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    sizes = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    py = [C.sizeof(t) for t in (_abi.CsMaterial, _abi.CsContactParams, _abi.CsSceneDesc, _abi.CsOptions,
                                _abi.CsContactDiag, _abi.CsStepStats)]
    assert sizes == py


def test_defaults_match_reference_stepper_options():
    L = load_library()
    o = _abi.CsOptions()
    L.cs_options_default(C.byref(o))
    assert (o.newton_iters, o.pcg_max_iters, o.line_search_iters, o.grad_tol) == (10, 60, 10, 1e-5)
    assert o.precision == _abi.CS_PREC_FP64 and o.hvp_mode == _abi.CS_HVP_FD and o.precond == _abi.CS_PRECOND_JACOBI


def test_invalid_arguments_are_rejected_without_a_device():
    L = load_library()
    h = C.c_void_p()
    assert L.cs_create(None, None, C.byref(h)) == _abi.CS_E_INVALID
    d = _abi.CsSceneDesc()
    assert L.cs_create(C.byref(d), None, C.byref(h)) == _abi.CS_E_INVALID
    assert b"missing" in L.cs_last_error()
    assert L.cs_step(None, 1 / 60) == _abi.CS_E_NOT_INIT
    assert L.cs_last_error() == b"CpuSolver not initialized"
    assert L.cs_get_positions(None, None, 0) == _abi.CS_E_NOT_INIT
    assert L.cs_abi_version() == 2


@pytest.mark.parametrize("make,sizes", [
    ("small_strip", (16, 8, 32)),
    ("strip_pull", (16, 8, 32)),
    ("strip_pull", (64, 1, 64)),
    ("crumpled", (24, 8, 48)),
])
def test_schwarz_partition_matches_oracle(make, sizes):
    """buildSchwarzDomains (schwarz.ts:44-140): the library's host partition is
    the oracle's, vertex by vertex (no device needed)."""
    import oracle_ffi as O
    from closim_amd import scenes as S
    from closim_amd.solver import schwarz_partition
    from contact_scenes import crumpled_sheet
    mk = crumpled_sheet if make == "crumpled" else getattr(S, make)
    dom, loc, cnt = schwarz_partition(mk(), *sizes)
    o = O.OracleSolver(mk())
    od, ol, oc = o.schwarz_domains(*sizes)
    assert cnt == oc and cnt > 1
    assert np.array_equal(dom, od)
    assert np.array_equal(loc, ol)
    sizes_seen = np.bincount(dom)
    assert sizes_seen.max() < sizes[0] + sizes[1]  # a merged remainder has fewer than min vertices


def test_batch_arguments_are_rejected_without_a_device():
    """cs_batch_step / HipBatch: null or repeated solvers and a bad dt fail
    before any instance steps (no device needed)."""
    from closim_amd.solver import HipBatch, load_library
    L = load_library()
    assert L.cs_batch_step(None, 0, 1 / 60, 1, None) == _abi.CS_OK  # an empty batch steps nothing
    arr = (C.c_void_p * 2)(None, None)
    assert L.cs_batch_step(arr, 2, 1 / 60, 2, None) == _abi.CS_E_NOT_INIT
    assert L.cs_batch_step(arr, 2, -1.0, 2, None) == _abi.CS_E_INVALID
    assert L.cs_batch_step(None, 2, 1 / 60, 2, None) == _abi.CS_E_INVALID
    s = object()
    with pytest.raises(ValueError):
        HipBatch([s, s])


def test_batch_drapes_workload_is_seeded():
    """The batched drape workload: 64 instances with the stated size and
    parameter ranges, identical for the same seed."""
    from closim_amd import scenes as S
    a, b = S.batch_drapes(64), S.batch_drapes(64)
    assert len(a) == 64
    n = np.array([sc.mesh.count for sc in a])
    assert n.min() >= 900 and n.max() <= 110000
    for x, y in zip(a, b):
        assert np.array_equal(x.positions, y.positions)
        m = x.material
        assert 1e5 <= m.stretchWarp <= 1e7 and 1e5 <= m.stretchWeft <= 1e7 and 5e4 <= m.shear <= 5e6
        assert 1e-6 <= m.bendWarp <= 1e-4 and m.bendWarp == m.bendWeft and m.stretchCoupling == 0.0
        assert x.contact.floorY == 0.0 and x.contact.params.frictionMu == 0.3


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus 2 exits with a message before touching
    a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _fan(rim):
    """A disk: one centre vertex shared by `rim` triangles."""
    from closim_amd.mesh import preprocess
    from closim_amd.scene import create_scene
    from closim_amd.scenes import STRIP_MATERIAL
    a = np.linspace(0, 2 * np.pi, rim, endpoint=False)
    pos = np.concatenate([[0.0, 0.0, 0.0], np.stack([np.cos(a), np.zeros(rim), np.sin(a)], 1).ravel()])
    uv = np.concatenate([[0.0, 0.0], np.stack([np.cos(a), np.sin(a)], 1).ravel()])
    idx = np.array([[0, 1 + k, 1 + (k + 1) % rim] for k in range(rim)], np.uint32).ravel()
    return create_scene(preprocess(pos, uv, idx, STRIP_MATERIAL.arealDensityKgM2), STRIP_MATERIAL)


def test_broadphase_arguments_are_rejected_without_a_device():
    L = load_library()
    cnt = C.c_uint64(7)
    assert L.cs_broadphase_self_pairs(0, None, None, 3, None, 1, 1e-3, None, 0, None, None, None) == _abi.CS_E_INVALID
    assert L.cs_broadphase_self_pairs(0, None, None, 3, None, 1, 1e-3, None, 0, C.byref(cnt), None,
                                      None) == _abi.CS_E_INVALID
    assert cnt.value == 0 and b"null" in L.cs_last_error()
    p = C.c_void_p(8)
    assert L.cs_broadphase_self_pairs(0, p, p, 3, p, 1, -1.0, p, 1, C.byref(cnt), None, None) == _abi.CS_E_INVALID
    assert L.cs_broadphase_self_pairs(0, p, p, 3, p, 1, float("nan"), p, 1, C.byref(cnt), None,
                                      None) == _abi.CS_E_INVALID
    assert L.cs_broadphase_self_pairs(0, p, p, 3, p, 1, 1e-3, None, 4, C.byref(cnt), None, None) == _abi.CS_E_INVALID
