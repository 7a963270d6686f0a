"""The lagged friction's share of the step adjoint's dL/dx_prev (the friction
slip is measured from the step's start positions, contact-assembly.ts:
344-367): cs_probe_contact_adjoint's vjp = -(d g_contact / d x_step)^T lam
against central differences of lam . g_contact(x; x_step) along random
directions of x_step, on the floor patch with friction (mu 0.3), in the
smooth regime (slip below the friction regularisation eps = 1e-4 m) and in
the sliding one (slip above it).  Floor contacts depend on x alone, so the
active set is the same for every x_step of the differences."""
import numpy as np
import pytest

from closim_amd import scenes as S
from closim_amd.solver import HipSolver

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0


@pytest.mark.parametrize("slip", [2e-5, 4e-4])
def test_friction_start_position_vjp_matches_differences(slip):
    sc = S.floor_patch(mu=0.3)
    g = HipSolver("fp64")
    g.initialize(sc)
    for _ in range(30):
        g.step(H)
    x = sc.positions.copy()
    rng = np.random.default_rng(23)
    d = np.zeros_like(x).reshape(-1, 3)
    d[:, 0] = slip * (1 + 0.3 * rng.standard_normal(len(d)))
    d[:, 2] = 0.5 * slip * rng.standard_normal(len(d))
    xs = x - d.reshape(-1)  # the step started behind x: slip u = x - x_step
    lam = rng.standard_normal(len(x))
    grad, vjp = g.probe_contact_adjoint(xs, x, lam)
    assert np.any(grad != 0) and np.any(vjp != 0)
    worst = 0.0
    def central(q, eps):
        gp, _ = g.probe_contact_adjoint(xs + eps * q, x)
        gm, _ = g.probe_contact_adjoint(xs - eps * q, x)
        return -(lam @ gp - lam @ gm) / (2 * eps)

    for t in range(4):
        q = rng.standard_normal(len(x))
        # Richardson extrapolation of two central differences (error O(eps^4)):
        # in the sliding regime the slip's third derivative scales as 1/|u|^3
        eps = 1e-3 * slip
        fd = (4 * central(q, eps / 2) - central(q, eps)) / 3
        ad = float(vjp @ q)
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        print(f"slip {slip:g}: vjp.q {ad:.9e} fd {fd:.9e} rel {rel:.1e}", flush=True)
        worst = max(worst, rel)
    assert worst < 1e-5
