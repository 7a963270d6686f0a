"""Host logic of the material-parameter adjoint (no device): parameter
packing, finite-difference steps, the trajectory loss and its gradient, and
the seeded material workload."""
import numpy as np

from closim_amd import adjoint as A
from closim_amd.scenes import STRIP_MATERIAL


def test_material_round_trip_and_kb():
    th = np.array([2e6, 3e5, 1e4, 6e4, 3e-6])
    m = A.material_with(STRIP_MATERIAL, th)
    assert m.bendWarp == m.bendWeft == th[4]  # kb = (bendWarp + bendWeft) / 2 (fem.ts:44)
    assert np.array_equal(A.theta_of(m), th)
    assert m.thickness == STRIP_MATERIAL.thickness and m.arealDensityKgM2 == STRIP_MATERIAL.arealDensityKgM2


def test_fd_step_scales_and_zero_parameters():
    th = np.array([2e6, 3e5, 0.0, 6e4, 0.0])
    assert A.fd_step(th, 0, 1e-3) == 2e3
    assert A.fd_step(th, 2, 1e-3) == 2e3  # a zero modulus steps by the largest membrane modulus
    assert A.fd_step(th, 4, 1e-3) == 1e-3  # a zero kb steps by rel_eps


def test_trajectory_loss_gradient_is_its_derivative():
    rng = np.random.default_rng(1)
    xs, tg = rng.standard_normal((4, 30)), rng.standard_normal((4, 30))
    L = A.trajectory_loss(xs, tg)
    assert np.isclose(L, np.sum((xs - tg) ** 2) / (4 * 10))
    dL = 2.0 * (xs - tg) / (4 * 10)  # the form loss_and_grad passes to the adjoint
    d = rng.standard_normal(xs.shape)
    e = 1e-6
    fd = (A.trajectory_loss(xs + e * d, tg) - A.trajectory_loss(xs - e * d, tg)) / (2 * e)
    assert abs(fd - np.sum(dL * d)) <= 1e-8 * abs(fd)


def test_material_problem_is_seeded_within_ranges():
    sc, tt, tg = A.material_problem(15, seed=3)
    sc2, tt2, tg2 = A.material_problem(15, seed=3)
    assert np.array_equal(tt, tt2) and np.array_equal(tg, tg2)
    assert 1e5 <= tt[0] <= 1e7 and 1e5 <= tt[1] <= 1e7 and 5e4 <= tt[3] <= 5e6 and 1e-6 <= tt[4] <= 1e-4
    assert tt[2] == 0.0
    assert np.all(np.abs(tg / np.where(tt == 0, 1, tt) - 1)[[0, 1, 3, 4]] <= 0.2)
    assert sc.mesh.count == 16 * 16 and set(sc.pinned) == {0, 15}
