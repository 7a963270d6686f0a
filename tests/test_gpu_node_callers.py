"""The reference's two callers of a ClothSolver (the frame-driven simulation
session, sim-session.ts:116-152, and the fitting run, garment/assembly.ts:
554-640) driven over the JS HipSolver on the GPU (tests/node/callers_gpu.js):
the fixed-step accumulator takes the expected number of steps per frame, the
published positions are the live array, the fitting run's trajectory equals
the Python binding's, and a blown-up state makes step() throw the
reference's non-finite error, which the run reports as a failure."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from closim_amd import _abi, scenes as S
from closim_amd.solver import HipSolver
from node_scene import write_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(shutil.which("node") is None, reason="no node in this image")]


def test_session_and_fitting_callers_over_node(tmp_path):
    write_scene(S.floor_patch(), tmp_path / "scene.json")
    opts = {"precision": "mixed", "preconditioner": "block-jacobi"}
    with open(tmp_path / "opts.json", "w") as f:
        json.dump(opts, f)
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "callers_gpu.js"), str(tmp_path / "scene.json"),
                        str(tmp_path / "opts.json"), str(tmp_path / "out.json")], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    with open(tmp_path / "out.json") as f:
        res = json.load(f)
    s = res["session"]
    # 1/60 -> 1 step; 1/30 -> 2; a 2 s stall is clamped to 0.25 s and capped
    # at 8 steps (backlog dropped); four 5 ms frames make 1 step; then 1
    assert s["perFrame"] == [1, 2, 8, 0, 0, 0, 1, 1]
    assert s["steps"] == 13 and len(s["published"]) == 13
    assert s["stats"]["contact"] is not None
    # the fitting run: 15 steps without failure, the same trajectory as the Python binding
    f = res["fitting"]
    assert f["taken"] == 15 and f["failure"] is None
    sc = S.floor_patch()
    g = HipSolver("mixed", precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    g.initialize(sc)
    for _ in range(15):
        g.step(1 / 60)
    assert np.array_equal(np.array(f["positions"]), sc.positions)
    b = res["blowup"]
    assert b["taken"] == 0 and b["failure"]["code"] == "nan-inf-state"
    assert "non-finite gradient" in b["failure"]["message"]
