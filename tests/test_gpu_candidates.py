"""The device contact system's candidate list against the oracle's
buildCandidates (contact-assembly.ts:87-121), as a multiset, not a count:
after beginStep at x0, the cloth/cloth triangle pairs the device traversal
finds within the swept, padded boxes x0 -> x (pad dHatM for the active set,
dMinM for checkTrial) and its cloth/static pairs are expanded here the way
expandTriPair (self-collision.ts:23-41) and the static loop of
buildCandidates do, sorted, and compared row for row with the oracle's
sorted candidates.  States: small contact scenes after a few steps, and the
256x256 sphere drape and 512x512 fold at their initial state and after one
FP64 step (which is bit-identical to the oracle's by the golden fixtures)."""
import numpy as np
import pytest

from closim_amd import _abi, scenes as S
from closim_amd.solver import HipSolver
from oracle_ffi import OracleSolver
from contact_scenes import ALL

pytestmark = pytest.mark.gpu

TE = ((0, 1), (1, 2), (2, 0))


def expand(self_pairs, static_pairs, idx, sidx):
    """buildCandidates' expansion of triangle pairs into (kind, p, a, b, c, d)
    rows: per cloth pair 3 + 3 vertex-triangle rows and the edge-edge rows of
    edges sharing no vertex; per cloth/static pair 3 vertex-triangle rows
    (static vertices s as -(s + 2), -2 in the last slot) and 9 edge-edge
    rows.  Returned sorted lexicographically."""
    T = np.asarray(idx, np.int64).reshape(-1, 3)
    rows = []
    if len(self_pairs):
        A, B = T[self_pairs[:, 0]], T[self_pairs[:, 1]]
        m1 = np.full(len(A), -1, np.int64)
        for P, Q in ((A, B), (B, A)):
            for c in range(3):
                rows.append(np.stack([np.zeros_like(m1), P[:, c], Q[:, 0], Q[:, 1], Q[:, 2], m1], 1))
        for ea in TE:
            for eb in TE:
                a, b, c, d = A[:, ea[0]], A[:, ea[1]], B[:, eb[0]], B[:, eb[1]]
                keep = (a != c) & (a != d) & (b != c) & (b != d)
                rows.append(np.stack([np.ones_like(m1), m1, a, b, c, d], 1)[keep])
    if len(static_pairs):
        Sx = -(np.asarray(sidx, np.int64).reshape(-1, 3) + 2)
        A, B = T[static_pairs[:, 0]], Sx[static_pairs[:, 1]]
        m1 = np.full(len(A), -1, np.int64)
        for c in range(3):
            rows.append(np.stack([np.zeros_like(m1), A[:, c], B[:, 0], B[:, 1], B[:, 2], np.full_like(m1, -2)], 1))
        for ea in TE:
            for eb in TE:
                rows.append(np.stack([np.ones_like(m1), m1, A[:, ea[0]], A[:, ea[1]], B[:, eb[0]], B[:, eb[1]]], 1))
    if not rows:
        return np.zeros((0, 6), np.int64)
    R = np.concatenate(rows)
    return R[np.lexsort(R.T[::-1])]


def compare(scene, g, o, x0, x, tag):
    idx = np.asarray(scene.mesh.indices, np.int64)
    c = scene.contact
    sidx = np.asarray(c.staticIdx, np.int64) if c.staticIdx is not None else np.zeros(0, np.int64)
    for dmin in (False, True):
        sp, st = g.probe_candidates(x0, x, pad_is_dmin=dmin)
        mine = expand(sp, st, idx, sidx)
        ref = o.candidates(x0, x, dmin=dmin).astype(np.int64)
        print(f"[cand] {tag} pad={'dMin' if dmin else 'dHat'}: {len(sp)} cloth pairs, {len(st)} static pairs, "
              f"{len(mine)} candidates (oracle {len(ref)})", flush=True)
        assert len(mine) == len(ref)
        assert np.array_equal(mine, ref)
        if len(sp):
            assert np.all(sp[:, 0] < sp[:, 1])
            # no cloth pair shares a vertex: the traversal drops adjacent triangles
            A, B = idx.reshape(-1, 3)[sp[:, 0]], idx.reshape(-1, 3)[sp[:, 1]]
            assert not np.any((A[:, :, None] == B[:, None, :]).any(axis=(1, 2)))
        assert len(np.unique(sp, axis=0)) == len(sp) and len(np.unique(st, axis=0)) == len(st)


@pytest.mark.parametrize("name", ["stacked", "plate", "crumpled", "stacked_large", "sphere_touch"])
def test_candidates_match_oracle_small_scenes(name):
    scene = ALL[name]()
    g = HipSolver("fp64")
    g.initialize(scene)
    o = OracleSolver(ALL[name](), _abi.options())
    compare(scene, g, o, scene.positions.copy(), scene.positions.copy(), f"{name} rest")
    for k in range(4):
        x0 = scene.positions.copy()
        g.step(1 / 60)
        compare(scene, g, o, x0, scene.positions.copy(), f"{name} step {k + 1}")


@pytest.mark.parametrize("make,opts", [
    (lambda: S.sphere_drape(255), dict(precond=1)),
    (lambda: S.crumple_fold(511), dict(precond=2, schwarz=(64, 1, 64), newton_iters=2)),
], ids=["sphere_drape_256", "crumple_fold_512"])
def test_candidates_match_oracle_full_size(make, opts):
    scene = make()
    g = HipSolver("fp64", **opts)
    g.initialize(scene)
    o = OracleSolver(make(), _abi.options(**opts))
    x0 = scene.positions.copy()
    compare(scene, g, o, x0, x0, "rest")
    g.step(1 / 60)
    compare(scene, g, o, x0, scene.positions.copy(), "after one step")
