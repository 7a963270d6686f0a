"""Step adjoint (cs_step_adjoint) against central finite differences.

The adjoint has no counterpart in the reference; it is pinned by the
mathematics: with converged forward solves, the implicit-function gradient
of a loss must equal its central finite differences.  The sheet is the
adjoint workload's shape at a small size: a grid pinned at two corners of one
edge, swinging under gravity, started from a stretched, bent state with
random velocities so that every material parameter moves the step's result.

The forward solves are converged to the root (`converged_solver`: root
polish past the Armijo test's energy round-off, |grad| ~ 1e-12 here), and the
central differences use a relative step of 1e-3: their truncation error is
O(1e-6) relative, while their noise, the root's round-off divided by the
step, is below 1e-5 at that step (at 1e-6 it reaches 1e-3 on these scenes).
"""
import numpy as np
import pytest

from closim_amd import adjoint as A
from closim_amd.scenes import STRIP_MATERIAL

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0
REL = 1e-4
REL_EPS = 1e-3
THETA = np.array([1e5, 6e4, 2e4, 3e4, 2e-5])


def rel_err(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def deformed_state(sc, seed=5):
    """The grid stretched along x and curved out of plane, by a displacement
    that vanishes on the pinned edge (z = 0), with random velocities on the
    free vertices."""
    P = sc.positions.reshape(-1, 3)
    side = float(P[:, 0].max())
    x, z = P[:, 0], P[:, 2]
    d = np.stack([0.02 * x * z / side,
                  3e-3 * np.sin(np.pi * x / side) * np.sin(0.5 * np.pi * z / side),
                  -0.01 * z * z / side], 1)
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(P.shape) * 0.02
    for p in sc.pinned:
        v[p] = 0
    return (P + d).reshape(-1), v.reshape(-1)


def make(nx=8, side=0.2, precision="fp64"):
    sc = A.corner_pinned_sheet(nx, side, material=A.material_with(STRIP_MATERIAL, THETA))
    g = A.converged_solver(sc, precision)
    g.set_state(*deformed_state(sc))
    return sc, g


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_step_parameter_and_state_gradients_match_fd(precision):
    # the mixed forward (FP32 inner solves) still converges the FP64 Newton
    # gradient to the root; the adjoint itself runs in FP64
    sc, g = make(precision=precision)
    x0, v0 = sc.positions.copy(), sc.velocities.copy()
    rng = np.random.default_rng(3)
    wx = rng.standard_normal(len(x0))
    wv = rng.standard_normal(len(x0))
    theta = THETA.copy()

    def loss(th, x=None, v=None):
        g.setMaterial(A.material_with(STRIP_MATERIAL, th))
        g.set_state(x0 if x is None else x, v0 if v is None else v)
        g.step(H)
        return float(wx @ sc.positions + 1e-2 * wv @ sc.velocities), sc.positions.copy()

    _, x1 = loss(theta)
    st = g.lastStats
    print(f"forward: newton {st['newton_used']} |g| {st['grad_norm']:.2e}")
    dx, dv, dth, it, rr = g.step_adjoint(H, x0, v0, x1, wx, 1e-2 * wv)
    print(f"adjoint PCG {it} iterations, rel residual {rr:.1e}")
    assert rr <= 1e-10
    for j in range(5):
        e = A.fd_step(theta, j, REL_EPS)
        tp, tm = theta.copy(), theta.copy()
        tp[j] += e
        tm[j] -= e
        fd = (loss(tp)[0] - loss(tm)[0]) / (2 * e)
        print(f"{A.PARAMS[j]}: adjoint {dth[j]:.9e} fd {fd:.9e} rel {rel_err(dth[j], fd):.1e}")
        assert rel_err(dth[j], fd) < REL, j
    # initial state along random directions (free vertices only move)
    free = np.ones(len(x0))
    for p in sc.pinned:
        free[p * 3:p * 3 + 3] = 0
    for which in ("x", "v"):
        d = rng.standard_normal(len(x0)) * free
        e = 1e-5 if which == "x" else 1e-3  # metres, metres per second
        if which == "x":
            fd = (loss(theta, x=x0 + e * d)[0] - loss(theta, x=x0 - e * d)[0]) / (2 * e)
            ad = float(dx @ d)
        else:
            fd = (loss(theta, v=v0 + e * d)[0] - loss(theta, v=v0 - e * d)[0]) / (2 * e)
            ad = float(dv @ d)
        print(f"d{which}0: adjoint {ad:.9e} fd {fd:.9e} rel {rel_err(ad, fd):.1e}")
        assert rel_err(ad, fd) < REL


def test_rollout_with_checkpoints_matches_fd():
    # six steps, checkpoints every two; loss = mean squared distance to the
    # trajectory of perturbed parameters
    sc, g = make()
    theta_true = THETA * np.array([1.3, 0.8, 1.5, 1.2, 2.0])
    r = A.MaterialRollout(g, steps=6, h=H, checkpoint_every=2)
    targets = r.run(theta_true)
    L, grad, fd = A.fd_check(r, THETA, targets, rel_eps=REL_EPS)
    print("loss", L, "adjoint solves", r.adjoint_iters)
    for j in range(5):
        print(f"{A.PARAMS[j]}: adjoint {grad[j]:.9e} fd {fd[j]:.9e} rel {rel_err(grad[j], fd[j]):.1e}")
        assert rel_err(grad[j], fd[j]) < REL, j


def test_torch_autograd_path_matches_rollout_backward():
    torch = pytest.importorskip("torch")
    # 4 steps x 64 vertices: the loss's normalisation 1/256 is exact, so the
    # torch gradient of the loss and loss_and_grad's are the same numbers
    sc, g = make(7, 0.15)
    r = A.MaterialRollout(g, steps=4, h=H, checkpoint_every=3)
    targets = torch.from_numpy(r.run(THETA * 1.1))
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    xs = A.rollout_fn().apply(th, r)
    loss = ((xs - targets) ** 2).sum() / (xs.shape[0] * (xs.shape[1] // 3))
    loss.backward()
    L2, g2 = A.loss_and_grad(r, THETA, targets.numpy())
    assert abs(float(loss.detach()) - L2) <= 1e-14 * abs(L2)
    assert np.array_equal(th.grad.numpy(), g2)


def test_one_step_gradients_with_floor_contact_match_fd():
    # A frictionless patch settled on the floor (every vertex inside the
    # barrier zone), then stretched and sheared 2% in its plane: the membrane pulls it
    # back during the step while the barrier holds it up.  The active set
    # stays the same within the differences' step, so the loss is smooth
    # there; the adjoint differentiates the barrier with that set frozen.
    from closim_amd import scenes as S
    sc = S.floor_patch(mu=0.0)
    theta = A.theta_of(sc.material)
    theta[2] = 0.3 * theta[1]
    g = A.converged_solver(sc, "fp64")
    g.setMaterial(A.material_with(sc.material, theta))
    for _ in range(40):
        g.step(H)
    P = sc.positions.reshape(-1, 3).copy()
    c = P.mean(0)
    P[:, 0] = c[0] + 1.02 * (P[:, 0] - c[0]) + 0.02 * (P[:, 2] - c[2])  # stretch and shear
    P[:, 2] = c[2] + 1.02 * (P[:, 2] - c[2])
    x0, v0 = P.reshape(-1), np.zeros(len(sc.velocities))
    rng = np.random.default_rng(7)
    wx = rng.standard_normal(len(x0))

    def loss(th):
        g.setMaterial(A.material_with(sc.material, th))
        g.set_state(x0, v0)
        g.step(H)
        return float(wx @ sc.positions), g.lastStats["contact"]["active_pairs"]

    L0, act0 = loss(theta)
    assert act0 > 0
    x1 = sc.positions.copy()
    dx, dv, dth, it, rr = g.step_adjoint(H, x0, v0, x1, wx, np.zeros_like(wx))
    print(f"active {act0}, |g| {g.lastStats['grad_norm']:.1e}, adjoint PCG {it} iterations, rel residual {rr:.1e}")
    assert rr <= 1e-10
    for j in range(4):  # the membrane moduli (the patch is flat: kb does not act)
        e = A.fd_step(theta, j, REL_EPS)
        tp, tm = theta.copy(), theta.copy()
        tp[j] += e
        tm[j] -= e
        (lp, ap), (lm, am) = loss(tp), loss(tm)
        fd = (lp - lm) / (2 * e)
        print(f"{A.PARAMS[j]}: adjoint {dth[j]:.9e} fd {fd:.9e} rel {rel_err(dth[j], fd):.1e} active {ap}/{am}")
        assert ap == act0 and am == act0
        assert rel_err(dth[j], fd) < REL, j


def test_material_adjoint_at_128x128_first_steps_match_fd():
    # The 128x128 material-adjoint workload (two pinned corners, seeded true
    # and guessed parameters), its first 2 steps: targets from the true
    # parameters, the trajectory loss's gradient at the guess by checkpointed
    # step adjoints against central differences in every parameter.
    sc, theta_true, theta_guess = A.material_problem(127, 1.0, seed=0)
    g = A.converged_solver(sc, "fp64")
    r = A.MaterialRollout(g, 2, H, checkpoint_every=1)
    targets = r.run(theta_true)
    L, grad, fd = A.fd_check(r, theta_guess, targets, rel_eps=1e-3)
    print(f"[c3] {sc.mesh.count} vertices, loss {L:.6e}, forward |grad| up to {r.max_grad_norm:.1e}, "
          f"adjoint iterations {[it for it, _ in r.adjoint_iters]}", flush=True)
    for j in range(5):
        if theta_guess[j] == 0 and abs(fd[j]) < 1e-300:
            continue
        print(f"[c3] {A.PARAMS[j]}: adjoint {grad[j]:.9e} fd {fd[j]:.9e} rel {rel_err(grad[j], fd[j]):.1e}",
              flush=True)
        assert rel_err(grad[j], fd[j]) < REL, A.PARAMS[j]
    g.close()
