"""Monotone Armijo energy decrease on the GPU (newton.ts:184-219).

Every accepted Newton iterate satisfies eNew <= E + 1e-4 alpha g.dx with
g.dx < 0, so the step's energy never rises from one Newton iteration to the
next.  A step's Newton iterates do not depend on the iteration cap, so the
step taken from one saved state with caps 1, 2, ..., K reports the energy
after each of its first K iterations (`lastStats["energy"]`, the energy of
the last accepted iterate); the sequence must be non-increasing.  Checked in
FP64 and mixed precision, with the host line search and with the device
Newton control (batched Armijo trials), on a sphere drape with friction and
on a self-contact fold, both in contact.
"""
import numpy as np
import pytest

from closim_amd import _abi, scenes as S
from closim_amd.solver import HipSolver

pytestmark = pytest.mark.gpu
H = 1.0 / 60.0
K = 8

CASES = {
    # 48x48 sheet on the sphere (mu 0.3), block-Jacobi
    "drape48": (lambda: S.sphere_drape(47), dict(precond=_abi.CS_PRECOND_BLOCK_JACOBI), 40),
    # 64x64 crumpled fold (self-contact, mu 0.4), additive Schwarz
    "fold64": (lambda: S.crumple_fold(63), dict(precond=_abi.CS_PRECOND_SCHWARZ, schwarz=(64, 1, 64)), 6),
}


def _state_in_contact(make, kw, precision, warm):
    sc = make()
    g = HipSolver(precision, **kw)
    g.initialize(sc)
    for _ in range(warm):
        g.step(H)
    x, v = g.getPositions().copy(), g.getVelocities().copy()
    g.close()
    return x, v


@pytest.mark.parametrize("device_newton", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_energy_never_rises_over_newton_iterations(case, precision, device_newton):
    make, kw, warm = CASES[case]
    x, v = _state_in_contact(make, kw, precision, warm)
    flags = _abi.CS_FLAG_DEVICE_NEWTON if device_newton else 0
    energies, used, active = [], [], []
    for k in range(1, K + 1):
        g = HipSolver(precision, **kw, newton_iters=k, grad_tol=1e-12, flags=flags)
        g.initialize(make())
        g.set_state(x, v)
        g.step(H)
        st = g.lastStats
        energies.append(st["energy"])
        used.append(st["newton_used"])
        active.append(st["contact"]["active_pairs"])
        g.close()
    print(f"{case} {precision} device_newton={device_newton}: newton used {used}, active pairs {active}, "
          f"energies {['%.12e' % e for e in energies]}", flush=True)
    assert all(np.isfinite(energies))
    assert max(active) > 0
    assert used[-1] >= 2
    for k in range(1, K):
        assert energies[k] <= energies[k - 1], (k, energies[k - 1], energies[k])
    assert energies[-1] < energies[0]
