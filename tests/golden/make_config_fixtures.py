"""Writes the config-size golden fixtures (tests/golden/*.json) from the CPU
oracle, which is too slow to run on the GPU host per test (about 50 s per
Newton iteration of the 512x512 fold):

* sphere_drape_256_step0: the 256x256 sphere + floor drape, block-Jacobi,
  the reference's caps (10 Newton, 60 PCG), first step;
* crumple_fold_512_step0: the 512x512 self-contact fold, additive Schwarz on
  64-vertex domains, first step with the Newton cap at 2.

Each fixture holds the options, the step's StepStats, SHA-256 digests of the
positions and velocities (little-endian float64) and every 97th coordinate,
so a test can check the FP64 HipSolver bit for bit and, if the digests
differ, to 1e-9 relative on the sample.  Run from the repository root:
python3 tests/golden/make_config_fixtures.py [name ...]"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle_ffi as O  # noqa: E402
from closim_amd import _abi, scenes as S  # noqa: E402

H = 1.0 / 60.0
STRIDE = 97

CASES = {
    "sphere_drape_256_step0": dict(scene="sphere_drape", args=[255],
                                   options=dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI), steps=1),
    "crumple_fold_512_step0": dict(scene="crumple_fold", args=[511],
                                   options=dict(precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(64, 1, 64), newton_iters=2),
                                   steps=1),
}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


def make(name):
    c = CASES[name]
    scene = getattr(S, c["scene"])(*c["args"])
    o = O.OracleSolver(scene, _abi.options(**c["options"]))
    t0 = time.time()
    stats = []
    for _ in range(c["steps"]):
        o.step(H)
        stats.append(o.lastStats)
    x, v = o.getPositions(), o.getVelocities()
    doc = dict(name=name, scene=c["scene"], args=c["args"], options={k: list(v_) if isinstance(v_, tuple) else v_
                                                                     for k, v_ in c["options"].items()},
               steps=c["steps"], h=H, stats=stats, positions_sha256=digest(x), velocities_sha256=digest(v),
               stride=STRIDE, positions_sample=[float(q) for q in x[::STRIDE]],
               velocities_sample=[float(q) for q in v[::STRIDE]], oracle_seconds=round(time.time() - t0, 1))
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), name + ".json"), "w") as f:
        json.dump(doc, f, indent=1, default=float)
    print(name, "written", doc["oracle_seconds"], "s", stats[-1]["newton_used"], stats[-1]["pcg_iters"], flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or list(CASES):
        make(n)
