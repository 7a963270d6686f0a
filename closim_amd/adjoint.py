"""Material-parameter gradients of cloth rollouts through the step adjoint.

`cs_step_adjoint` differentiates one implicit step by the implicit function
theorem (the step's Newton root x1 of g = M (x - yhat)/h^2 + grad E(x; theta)
+ ...; H lambda = dL/dx1 + (damp/h) dL/dv1; dL/dtheta_j = -lambda .
dg/dtheta_j).  This module chains it over a rollout:

* `MaterialRollout` runs T steps of a HipSolver from a fixed initial state
  under material parameters theta = (stretchWarp, stretchWeft,
  stretchCoupling, shear, kb), keeping the state at every k-th step
  (checkpoints); its `backward` walks the segments in reverse, recomputes
  each segment's states from its checkpoint and applies the step adjoint
  step by step, accumulating dL/dtheta.
* `RolloutFn` wraps that as a torch.autograd.Function of theta, so a loss on
  the positions (for instance the mean squared distance to a target
  trajectory) differentiates with `backward()`.
* `fd_check` compares the adjoint gradient with central finite differences.

kb is the hinge stiffness the reference's energy uses, (bendWarp +
bendWeft)/2 (src/physics/fem.ts:44); the rollout sets both to kb.  The
forward solves must be converged tightly (small gradTol, enough Newton and
PCG iterations, root polish; `converged_solver`) for the implicit-function
gradient to be the derivative of what the rollout computes.
"""
from dataclasses import replace

import numpy as np

from . import _abi
from .scene import ClothMaterial
from .solver import HipSolver

PARAMS = ("stretchWarp", "stretchWeft", "stretchCoupling", "shear", "kb")


def material_with(base: ClothMaterial, theta) -> ClothMaterial:
    c00, c11, c01, g, kb = (float(t) for t in theta)
    return replace(base, stretchWarp=c00, stretchWeft=c11, stretchCoupling=c01, shear=g, bendWarp=kb, bendWeft=kb)


def theta_of(material: ClothMaterial) -> np.ndarray:
    return np.array([material.stretchWarp, material.stretchWeft, material.stretchCoupling, material.shear,
                     0.5 * (material.bendWarp + material.bendWeft)])


class MaterialRollout:
    """T steps of `solver` (initialized on its scene, pins set) from the
    scene's current state, differentiable in the material parameters."""

    def __init__(self, solver: HipSolver, steps: int, h: float, checkpoint_every: int = 5,
                 adjoint_tol: float = 1e-10, adjoint_iters: int = 4000, progress=None, strict=True):
        self.solver = solver
        # strict: raise when an adjoint solve ends above 100x its tolerance (a
        # Hessian that is not positive definite along the search, typically
        # at a forward step that did not reach its root)
        self.strict = strict
        self.progress = progress  # optional callable(text) after every checkpoint segment
        self.steps = int(steps)
        self.h = float(h)
        self.k = max(1, int(checkpoint_every))
        self.tol = adjoint_tol
        self.iters = adjoint_iters
        sc = solver.scene
        self.x_init = sc.positions.copy()
        self.v_init = sc.velocities.copy()
        self.base = replace(sc.material)
        self.checkpoints = {}
        self.adjoint_iters = []
        self.max_grad_norm = 0.0  # largest final Newton gradient norm of the steps run so far

    def _step(self):
        self.solver.step(self.h)
        self.max_grad_norm = max(self.max_grad_norm, self.solver.lastStats["grad_norm"])
        sc = self.solver.scene
        return sc.positions.copy(), sc.velocities.copy()

    def run(self, theta, keep_all=False):
        """Positions after every step, shape (T, 3n); keeps a checkpoint
        (x, v) before every k-th step."""
        self.theta = np.asarray(theta, np.float64).copy()
        self.solver.setMaterial(material_with(self.base, self.theta))
        self.solver.set_state(self.x_init, self.v_init)
        x, v = self.x_init.copy(), self.v_init.copy()
        self.checkpoints = {}
        out = np.empty((self.steps, len(x)))
        for t in range(self.steps):
            if t % self.k == 0:
                self.checkpoints[t] = (x.copy(), v.copy())
            x, v = self._step()
            out[t] = x
            if self.progress and (t + 1) % self.k == 0:
                self.progress(f"forward step {t + 1}/{self.steps}")
        return out

    def backward(self, dL_dxs, dL_dvs=None):
        """dL/dtheta for gradients with respect to the positions (and
        optionally velocities) after every step, shape (T, 3n)."""
        n3 = dL_dxs.shape[1]
        gx = np.zeros(n3)
        gv = np.zeros(n3)
        gtheta = np.zeros(5)
        self.adjoint_iters = []
        self.solver.setMaterial(material_with(self.base, self.theta))
        starts = sorted(self.checkpoints)
        for s in reversed(starts):
            e = min(s + self.k, self.steps)
            # recompute the segment's states from its checkpoint
            xs, vs = [self.checkpoints[s][0]], [self.checkpoints[s][1]]
            self.solver.set_state(*self.checkpoints[s])
            for _ in range(s, e):
                x, v = self._step()
                xs.append(x)
                vs.append(v)
            for t in reversed(range(s, e)):
                gx = gx + dL_dxs[t]
                if dL_dvs is not None:
                    gv = gv + dL_dvs[t]
                j = t - s
                dx, dv, dth, it, rr = self.solver.step_adjoint(self.h, xs[j], vs[j], xs[j + 1], gx, gv, self.tol,
                                                               self.iters)
                self.adjoint_iters.append((it, rr))
                if self.strict and not rr <= 100 * self.tol:
                    raise RuntimeError(f"adjoint solve of step {t} ended at relative residual {rr:.2e} "
                                       f"after {it} iterations (forward |grad| up to {self.max_grad_norm:.2e})")
                gtheta += dth
                gx, gv = dx, dv
            if self.progress:
                self.progress(f"backward segment from step {s}")
        self.dL_dx0, self.dL_dv0 = gx, gv
        return gtheta


def rollout_fn():
    """torch.autograd.Function over a MaterialRollout (torch imported lazily)."""
    import torch

    class RolloutFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, rollout):
            xs = rollout.run(theta.detach().cpu().numpy())
            ctx.rollout = rollout
            return torch.from_numpy(xs)

        @staticmethod
        def backward(ctx, grad_xs):
            g = ctx.rollout.backward(grad_xs.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g), None

    return RolloutFn


def trajectory_loss(xs, targets):
    """Mean over steps and vertices of the squared position error."""
    d = xs - targets
    return float(np.sum(d * d) / (xs.shape[0] * (xs.shape[1] // 3)))


def loss_and_grad(rollout: MaterialRollout, theta, targets):
    xs = rollout.run(theta)
    L = trajectory_loss(xs, targets)
    dL = 2.0 * (xs - targets) / (xs.shape[0] * (xs.shape[1] // 3))
    return L, rollout.backward(dL)


def fd_step(theta, j, rel_eps):
    """Central-difference step for parameter j: rel_eps |theta_j|, or for a
    zero parameter rel_eps times the largest modulus of its group (the four
    membrane moduli share units; kb has its own)."""
    theta = np.asarray(theta, np.float64)
    if theta[j] != 0:
        return rel_eps * abs(theta[j])
    scale = np.max(np.abs(theta[:4])) if j < 4 else 1.0
    return rel_eps * (scale if scale > 0 else 1.0)


def fd_check(rollout: MaterialRollout, theta, targets, rel_eps=1e-4, params=range(5)):
    """Central differences of the trajectory loss in each parameter (step
    `fd_step`) beside the adjoint gradient: returns (loss, adjoint gradient,
    FD gradient over `params`).  The truncation error of a central difference
    is O(rel_eps^2); its noise is the forward solves' convergence floor
    divided by the step, so rel_eps should be as large as the truncation
    allows (1e-4 by default)."""
    theta = np.asarray(theta, np.float64)
    L, g = loss_and_grad(rollout, theta, targets)
    fd = np.full(5, np.nan)
    for j in params:
        e = fd_step(theta, j, rel_eps)
        tp, tm = theta.copy(), theta.copy()
        tp[j] += e
        tm[j] -= e
        fd[j] = (trajectory_loss(rollout.run(tp), targets) - trajectory_loss(rollout.run(tm), targets)) / (2 * e)
    return L, g, fd


def corner_pinned_sheet(nx=127, side=1.0, material=None):
    """The adjoint workload: an (nx+1)^2 sheet in the horizontal plane pinned
    at the two corners of one edge, swinging down under gravity."""
    from .mesh import build_grid, preprocess
    from .scene import create_scene
    from .scenes import STRIP_MATERIAL
    mat = material or STRIP_MATERIAL
    pos, uv, idx = build_grid(nx, nx, side, side)
    mesh = preprocess(pos, uv, idx, mat.arealDensityKgM2)
    scene = create_scene(mesh, mat, (0.0, -9.81, 0.0))
    for i in (0, nx):
        scene.pinned[i] = tuple(float(v) for v in scene.positions[i * 3:i * 3 + 3])
    scene.meta = dict(name=f"corner_pinned_sheet_{nx + 1}x{nx + 1}")
    return scene


def converged_solver(scene, precision="fp64", grad_tol=1e-13, newton_iters=200, pcg_max_iters=600, **kw):
    """A HipSolver whose steps converge to the root: a small grad_tol, enough
    Newton and PCG iterations, a positive semidefinite Newton operator
    (CS_FLAG_PSD_NEWTON) so that Newton descends through buckling, and root
    polish (CS_FLAG_POLISH_ROOT) past the point where the Armijo test stops
    resolving energy decreases."""
    g = HipSolver(precision, hvp_mode=_abi.CS_HVP_ANALYTIC, grad_tol=grad_tol, newton_iters=newton_iters,
                  pcg_max_iters=pcg_max_iters, flags=_abi.CS_FLAG_POLISH_ROOT | _abi.CS_FLAG_PSD_NEWTON, **kw)
    g.initialize(scene)
    return g


def material_problem(nx=127, side=1.0, seed=0, guess_spread=0.2):
    """The material-parameter workload: an (nx+1)^2 sheet of side `side`
    pinned at two corners; true parameters drawn log-uniformly (C00, C11 in
    [1e5, 1e7], G in [5e4, 5e6], kb in [1e-6, 1e-4], C01 = 0) from a seeded
    generator, and an initial guess that scales each drawn parameter by
    1 + U[-spread, spread].  Returns (scene, theta_true, theta_guess)."""
    from .scenes import STRIP_MATERIAL
    rng = np.random.default_rng(seed)

    def logu(lo, hi):
        return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))

    theta_true = np.array([logu(1e5, 1e7), logu(1e5, 1e7), 0.0, logu(5e4, 5e6), logu(1e-6, 1e-4)])
    theta_guess = theta_true * (1 + rng.uniform(-guess_spread, guess_spread, 5))
    scene = corner_pinned_sheet(nx, side, material=material_with(STRIP_MATERIAL, theta_true))
    return scene, theta_true, theta_guess

