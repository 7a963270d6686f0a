"""The FP64 HipSolver at the workloads' full size against the CPU oracle,
through the golden fixtures tests/golden/make_config_fixtures.py wrote (the
oracle takes minutes per step at this size): the first step of the 256x256
sphere drape (block-Jacobi, the reference's caps) and of the 512x512
self-contact fold (additive Schwarz, Newton capped at 2) must give the same
StepStats and bit-identical positions and velocities (SHA-256), and 1e-9
relative on every 97th coordinate.  Then the resulting state is audited for
interpenetration: every vertex-triangle and edge-edge primitive of the
non-adjacent triangle pairs within d_hat (the standalone broadphase, which
finds every pair closer than twice its pad) is farther apart than d_min,
the bound checkTrial (contact-assembly.ts:489-553) keeps, and every vertex
is above the floor by more than d_min.  (auditMinDistance, tests/helpers.ts:
47-102, also scans edge pairs inside one vertex's ring, which the contact
system never pairs; those are not audited here.)"""
import hashlib
import json
import os

import numpy as np
import pytest

from closim_amd import _abi, scenes as S
from closim_amd.solver import HipSolver, broadphase_self_pairs, expand_pairs, probe_closest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["sphere_drape_256_step0", "crumple_fold_512_step0"])
def test_first_step_at_full_size_matches_oracle_fixture(name):
    with open(os.path.join(HERE, "golden", name + ".json")) as f:
        doc = json.load(f)
    opts = {k: tuple(v) if isinstance(v, list) else v for k, v in doc["options"].items()}
    scene = getattr(S, doc["scene"])(*doc["args"])
    g = HipSolver("fp64", **opts)
    g.initialize(scene)
    print(f"[config] {name}: {scene.mesh.count} vertices, {scene.mesh.tri_count} triangles", flush=True)
    for _ in range(doc["steps"]):
        g.step(doc["h"])
    st, ref = g.lastStats, doc["stats"][-1]
    print(f"[config] {name}: newton {st['newton_used']}/{ref['newton_used']} pcg {st['pcg_iters']}/"
          f"{ref['pcg_iters']} trials {st['line_search_trials']}/{ref['line_search_trials']}", flush=True)
    for k in ("newton_used", "pcg_iters", "line_search_trials", "converged"):
        assert st[k] == ref[k], k
    assert st["grad_norm"] == ref["grad_norm"] and st["energy"] == ref["energy"]
    for k in ("candidate_pairs", "active_pairs", "vt_pairs", "ee_pairs", "floor_pairs", "min_distance",
              "rejected_line_search_steps"):
        assert st["contact"][k] == ref["contact"][k], k
    x, v = scene.positions, scene.velocities
    s = doc["stride"]
    xs, vs = np.array(doc["positions_sample"]), np.array(doc["velocities_sample"])
    assert np.max(np.abs(x[::s] - xs)) <= 1e-9 * np.max(np.abs(xs))
    assert np.max(np.abs(v[::s] - vs)) <= 1e-9 * max(np.max(np.abs(vs)), 1e-300)
    assert digest(x) == doc["positions_sha256"] and digest(v) == doc["velocities_sha256"]
    # interpenetration audit of the state the step produced
    p = scene.contact.params
    idx = np.ascontiguousarray(scene.mesh.indices, np.uint32)
    pairs = broadphase_self_pairs(x, x, idx, p.dHatM)
    kinds, ids = expand_pairs(pairs, idx)
    dmin = float(probe_closest(x, kinds, ids)[:, 2].min()) if len(kinds) else np.inf
    print(f"[config] {name}: {len(pairs)} triangle pairs within d_hat, {len(kinds)} primitives, "
          f"min distance {dmin:.3e} (d_min {p.dMinM})", flush=True)
    assert dmin > p.dMinM
    if scene.contact.floorY is not None:
        assert float(np.min(x[1::3])) - scene.contact.floorY > p.dMinM
