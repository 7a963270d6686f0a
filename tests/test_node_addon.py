"""The Node-API addon and the JS HipSolver (closim_amd/node/) under the
image's Node 12, without a device: argument rejection, the ClothSolver
contract before initialize, and the host-only entry points
(buildSchwarzDomains, the HVP chunk layout) equal to the Python binding's."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from closim_amd import scenes as S
from closim_amd.solver import hvp_tiles, load_library, schwarz_partition
from node_scene import write_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "closim_amd", "_build", "hip_solver_addon.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None, reason="no node in this image")


def test_addon_checks_on_the_cpu(tmp_path):
    assert os.path.exists(ADDON), "addon not built (closim_amd/node/build_addon.sh)"
    sc = S.strip_pull(9, 29, 0.1, 0.3)
    write_scene(sc, tmp_path / "scene.json")
    dom, loc, cnt = schwarz_partition(sc, 16, 8, 32)
    tl = hvp_tiles(sc)
    with open(tmp_path / "expected.json", "w") as f:
        json.dump(dict(schwarz=dict(count=int(cnt), domainOf=[int(v) for v in dom], localOf=[int(v) for v in loc]),
                       tiles=dict(tiles=int(len(tl["hdr"])), **{k: [int(v) for v in tl[k].ravel()]
                                                                 for k in ("hdr", "loc", "lidx", "off", "ent")})), f)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")  # no device even on a GPU host
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "addon_cpu.js"), str(tmp_path / "scene.json"),
                        str(tmp_path / "expected.json")], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "addon cpu checks OK" in r.stdout
