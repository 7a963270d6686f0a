"""The JS HipSolver over the Node-API addon on the GPU: the strip (FP64,
with a pin taken at the current position mid-run) against the CPU oracle
(1e-9 relative) and against the Python binding (identical), and a floor
patch in mixed precision with block-Jacobi against the Python binding
(identical: the same library on the same inputs)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
from closim_amd import _abi, scenes as S
from closim_amd.solver import HipSolver
from node_scene import write_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(shutil.which("node") is None, reason="no node in this image")]
H = 1.0 / 60.0


def run_node(tmp_path, scene, opts, steps, pin_step=-1, pin_vertex=-1):
    write_scene(scene, tmp_path / "scene.json")
    with open(tmp_path / "opts.json", "w") as f:
        json.dump(opts, f)
    out = tmp_path / "out.json"
    print(f"[node] {steps} steps, options {opts}", flush=True)
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "addon_gpu.js"), str(tmp_path / "scene.json"),
                        str(tmp_path / "opts.json"), str(steps), str(pin_step), str(pin_vertex), str(out)],
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    with open(out) as f:
        return json.load(f)


def test_strip_through_node_matches_oracle_and_python(tmp_path):
    steps, pin_step, pin_v = 4, 2, 40
    res = run_node(tmp_path, S.small_strip(), {"precision": "fp64"}, steps, pin_step, pin_v)
    so, sp = S.small_strip(), S.small_strip()
    o = O.OracleSolver(so, _abi.options())
    g = HipSolver("fp64")
    g.initialize(sp)
    for k in range(steps):
        if k == pin_step:
            o.pinVertex(pin_v)
            g.pinVertex(pin_v)
        o.step(H)
        g.step(H)
    xn = np.array(res["positions"])
    ref = o.getPositions()
    assert float(np.max(np.abs(xn - ref)) / np.max(np.abs(ref))) <= 1e-9
    assert np.array_equal(xn, sp.positions) and np.array_equal(np.array(res["velocities"]), sp.velocities)
    st = res["stats"][-1]
    assert st["newtonIters"] == 10 and st["newtonUsed"] == g.lastStats["newton_used"]
    assert st["pcgIters"] == g.lastStats["pcg_iters"] and st["contact"] is None
    pins = dict((int(k), v) for k, v in res["pinned"])
    assert pin_v in pins and np.allclose(xn[pin_v * 3:pin_v * 3 + 3], pins[pin_v], rtol=0, atol=0)


def test_floor_patch_mixed_through_node_matches_python(tmp_path):
    opts = {"precision": "mixed", "preconditioner": "block-jacobi"}
    res = run_node(tmp_path, S.floor_patch(), opts, 6)
    sp = S.floor_patch()
    g = HipSolver("mixed", precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    g.initialize(sp)
    for _ in range(6):
        g.step(H)
    assert np.array_equal(np.array(res["positions"]), sp.positions)
    c = res["stats"][-1]["contact"]
    assert c is not None and c["activePairs"] == g.lastStats["contact"]["active_pairs"]
