"""cs_batch_step: many independent solvers stepped together on host threads,
one HIP stream per instance.  Each instance's trajectory must be the one it
has when stepped alone (the instances share no state), bit for bit, in both
precisions; a failing instance reports its own status and leaves the others
stepped."""
import numpy as np
import pytest

from closim_amd import _abi, scenes as S
from closim_amd.solver import ClothError, HipBatch, HipSolver

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0


def instances():
    # uneven sizes, with and without contact
    return [S.small_strip(), S.floor_patch(nx=4), S.floor_patch(nx=9, mu=0.0), S.small_strip(),
            S.static_plate_patch(), S.floor_patch(nx=6)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_instances_stepped_alone(precision):
    kw = dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    sb, sa = instances(), instances()
    gb = [HipSolver(precision, **kw) for _ in sb]
    ga = [HipSolver(precision, **kw) for _ in sa]
    for g, sc in zip(gb, sb):
        g.initialize(sc)
    for g, sc in zip(ga, sa):
        g.initialize(sc)
    B = HipBatch(gb, threads=4)
    for t in range(6):
        assert B.step(H) == [0] * len(gb)
        for g in ga:
            g.step(H)
        for k, (a, b) in enumerate(zip(sa, sb)):
            assert np.array_equal(a.positions, b.positions), (t, k)
            assert np.array_equal(a.velocities, b.velocities), (t, k)
    for g in gb + ga:
        g.close()


def test_batch_reports_the_failing_instance():
    # instance 1 (no contact) starts from a non-finite position: its Newton
    # gradient is non-finite and its step fails on the host check
    scenes = [S.floor_patch(nx=4), S.small_strip(), S.floor_patch(nx=6)]
    gs = [HipSolver("fp64") for _ in scenes]
    for g, sc in zip(gs, scenes):
        g.initialize(sc)
    ref = [sc.positions.copy() for sc in scenes]
    bad = scenes[1].positions.copy()
    free = next(v for v in range(scenes[1].mesh.count) if v not in scenes[1].pinned)
    bad[free * 3] = np.nan  # pins would overwrite a pinned vertex
    gs[1].set_state(bad, scenes[1].velocities)
    B = HipBatch(gs, threads=3)
    with pytest.raises(ClothError) as e:
        B.step(H)
    assert e.value.code == _abi.CS_E_NONFINITE
    assert "instance 1" in str(e.value)
    # the healthy instances stepped (they moved under gravity)
    assert not np.array_equal(gs[0].getPositions(), ref[0])
    assert not np.array_equal(gs[2].getPositions(), ref[2])
    for g in gs:
        g.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_of_eight_uneven_drapes(precision):
    # The first 8 instances of the 64-sheet uneven drape workload (1e3..1e5
    # vertices, per-instance materials, each over its own sphere and a floor),
    # stepped as one batch and one at a time: identical trajectories.
    sb, sa = S.batch_drapes(8), S.batch_drapes(8)
    print(f"[batch] vertices per instance: {[sc.mesh.count for sc in sb]}", flush=True)
    gb = [HipSolver(precision, precond=_abi.CS_PRECOND_BLOCK_JACOBI) for _ in sb]
    ga = [HipSolver(precision, precond=_abi.CS_PRECOND_BLOCK_JACOBI) for _ in sa]
    for g, sc in zip(gb + ga, sb + sa):
        g.initialize(sc)
    B = HipBatch(gb, threads=8)
    for t in range(2):
        assert B.step(H) == [0] * len(gb)
        for g in ga:
            g.step(H)
        for k, (a, b) in enumerate(zip(sa, sb)):
            assert np.array_equal(a.positions, b.positions), (t, k)
            assert np.array_equal(a.velocities, b.velocities), (t, k)
    for g in gb + ga:
        g.close()


def test_smallest_uneven_drape_matches_oracle():
    # FP64 first steps of the smallest of those 8 instances against the oracle.
    import oracle_ffi as O
    scenes = S.batch_drapes(8)
    k = int(np.argmin([sc.mesh.count for sc in scenes]))
    sg, so = scenes[k], S.batch_drapes(8)[k]
    g = HipSolver("fp64", precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    g.initialize(sg)
    o = O.OracleSolver(so, _abi.options(precond=_abi.CS_PRECOND_BLOCK_JACOBI))
    for t in range(2):
        g.step(H)
        o.step(H)
        print(f"[batch] instance {k} ({sg.mesh.count} vertices) step {t}: newton {g.lastStats['newton_used']} "
              f"pcg {g.lastStats['pcg_iters']}", flush=True)
        assert g.lastStats["pcg_iters"] == o.lastStats["pcg_iters"]
        assert np.array_equal(sg.positions, so.positions) and np.array_equal(sg.velocities, so.velocities)
    g.close()


def test_contact_free_step_after_a_freed_contact_solver():
    # A solver freed after steps with active contacts hands its device memory
    # back to the runtime, and the next solver's buffers may reuse it.  The
    # FP64 PCG's per-vertex contact sums must not enter a solve whose active
    # set is empty: the smallest uneven drape's first step (no contact yet)
    # still matches the oracle bit for bit.
    import oracle_ffi as O
    scenes = S.batch_drapes(8)
    k = int(np.argmin([sc.mesh.count for sc in scenes]))
    ga = HipSolver("fp64", precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    ga.initialize(scenes[k])
    active = 0
    for _ in range(12):
        ga.step(H)
        active = max(active, ga.lastStats["contact"]["active_pairs"])
    ga.close()
    assert active > 0
    sg, so = S.batch_drapes(8)[k], S.batch_drapes(8)[k]
    g = HipSolver("fp64", precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    g.initialize(sg)
    o = O.OracleSolver(so, _abi.options(precond=_abi.CS_PRECOND_BLOCK_JACOBI))
    g.step(H)
    o.step(H)
    print(f"[batch] after a freed solver with {active} active contacts: pcg {g.lastStats['pcg_iters']} "
          f"(oracle {o.lastStats['pcg_iters']})", flush=True)
    assert g.lastStats["contact"]["active_pairs"] == 0
    assert g.lastStats["pcg_iters"] == o.lastStats["pcg_iters"]
    assert np.array_equal(sg.positions, so.positions) and np.array_equal(sg.velocities, so.velocities)
    g.close()


def test_packed_batch_matches_instances_stepped_alone():
    # cs_batch_step_packed: the instances' PCG systems solved together at every
    # Newton iteration (one launch per PCG stage over all instances), each with
    # its own arithmetic and reduction trees: trajectories and PCG counts equal
    # to stepping every instance alone, bit for bit
    kw = dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI)
    sb, sa = instances(), instances()
    gb = [HipSolver("mixed", **kw) for _ in sb]
    ga = [HipSolver("mixed", **kw) for _ in sa]
    for g, sc in zip(gb + ga, sb + sa):
        g.initialize(sc)
    B = HipBatch(gb, packed=True)
    for t in range(6):
        assert B.step(H) == [0] * len(gb)
        for g in ga:
            g.step(H)
        for k, (a, b) in enumerate(zip(sa, sb)):
            assert gb[k].lastStats["pcg_iters"] == ga[k].lastStats["pcg_iters"], (t, k)
            assert np.array_equal(a.positions, b.positions), (t, k)
            assert np.array_equal(a.velocities, b.velocities), (t, k)
    for g in gb + ga:
        g.close()


def test_packed_batch_of_eight_uneven_drapes():
    sb, sa = S.batch_drapes(8), S.batch_drapes(8)
    gb = [HipSolver("mixed", precond=_abi.CS_PRECOND_BLOCK_JACOBI) for _ in sb]
    ga = [HipSolver("mixed", precond=_abi.CS_PRECOND_BLOCK_JACOBI) for _ in sa]
    for g, sc in zip(gb + ga, sb + sa):
        g.initialize(sc)
    B = HipBatch(gb, packed=True)
    for t in range(2):
        assert B.step(H) == [0] * len(gb)
        for g in ga:
            g.step(H)
        its = [g.lastStats["pcg_iters"] for g in gb]
        print(f"[packed] step {t}: PCG iterations {its}", flush=True)
        for k, (a, b) in enumerate(zip(sa, sb)):
            assert np.array_equal(a.positions, b.positions), (t, k)
    for g in gb + ga:
        g.close()


def test_packed_batch_rejects_fp64_and_schwarz_and_survives_a_failing_instance():
    for prec, kw in (("fp64", {}), ("mixed", dict(precond=_abi.CS_PRECOND_SCHWARZ))):
        g = HipSolver(prec, **kw)
        g.initialize(S.small_strip())
        with pytest.raises(ClothError) as e:
            HipBatch([g], packed=True).step(H)
        assert e.value.code == _abi.CS_E_INVALID
        g.close()
    # instance 1 fails its first Newton check (non-finite gradient) and leaves
    # the packed solves; the others keep stepping
    scenes = [S.floor_patch(nx=4), S.small_strip(), S.floor_patch(nx=6)]
    gs = [HipSolver("mixed", precond=_abi.CS_PRECOND_BLOCK_JACOBI) for _ in scenes]
    for g, sc in zip(gs, scenes):
        g.initialize(sc)
    ref = [sc.positions.copy() for sc in scenes]
    bad = scenes[1].positions.copy()
    free = next(v for v in range(scenes[1].mesh.count) if v not in scenes[1].pinned)
    bad[free * 3] = np.nan
    gs[1].set_state(bad, scenes[1].velocities)
    with pytest.raises(ClothError) as e:
        HipBatch(gs, packed=True).step(H)
    assert e.value.code == _abi.CS_E_NONFINITE and "instance 1" in str(e.value)
    assert not np.array_equal(gs[0].getPositions(), ref[0])
    assert not np.array_equal(gs[2].getPositions(), ref[2])
    for g in gs:
        g.close()
