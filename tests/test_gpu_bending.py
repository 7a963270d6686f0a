"""Analytic hinge bending on the device (bend.h; the reference differences
hingeEnergy, bending.ts:19-42): the gradient against a complex-step
derivative of the closed-form energy (exact to rounding) and against the
oracle's central-difference gradient (its truncation), the Hessian-vector
product against central differences of the analytic gradient, symmetry
p.H q = q.H p, and the flat hinge of a flat sheet (theta = pi: gradient 0,
a finite Hessian where acos' is infinite)."""
import math

import numpy as np
import pytest

import oracle_ffi as O
from closim_amd.solver import probe_hinges
from test_oracle_pinning import _complex_step_grad, _random_hinges

pytestmark = pytest.mark.gpu


def _batch(hinges):
    X = np.array([h[0] for h in hinges])
    rest = np.array([h[1] for h in hinges])
    ln = np.array([h[2] for h in hinges])
    asum = np.array([h[3] for h in hinges])
    return X, rest, ln, asum


def test_gradient_matches_complex_step_and_reference_fd():
    rng = np.random.default_rng(31)
    hinges = _random_hinges(rng, 60)
    k = 3e-5
    X, rest, ln, asum = _batch(hinges)
    g, _ = probe_hinges(X, np.zeros_like(X), rest, ln, asum, k)
    worst_cs = worst_fd = 0.0
    for i, (x, r, l, a, _) in enumerate(hinges):
        cs = _complex_step_grad(x, r, l, a, k)
        scale = np.max(np.abs(cs))
        worst_cs = max(worst_cs, np.max(np.abs(g[i] - cs)) / scale)
        fd, _ = O.hinge_grad_fd(x, r, l, a, k)
        worst_fd = max(worst_fd, np.max(np.abs(g[i] - fd)) / scale)
    print(f"analytic vs complex step {worst_cs:.2e}, vs reference FD {worst_fd:.2e}")
    assert worst_cs < 1e-9
    assert worst_fd < 1e-5  # the FD's truncation and round-off (eps 1e-7)


def test_hvp_matches_differences_and_is_symmetric():
    rng = np.random.default_rng(32)
    hinges = _random_hinges(rng, 40)
    k = 2e-5
    X, rest, ln, asum = _batch(hinges)
    P = rng.standard_normal(X.shape)
    Q = rng.standard_normal(X.shape)
    _, hp = probe_hinges(X, P, rest, ln, asum, k)
    _, hq = probe_hinges(X, Q, rest, ln, asum, k)
    eps = 1e-6
    gp, _ = probe_hinges(X + eps * P, P, rest, ln, asum, k)
    gm, _ = probe_hinges(X - eps * P, P, rest, ln, asum, k)
    fd = (gp - gm) / (2 * eps)
    rel = np.max(np.abs(hp - fd), 1) / np.max(np.abs(fd), 1)
    print(f"hvp vs central differences: worst {rel.max():.2e}")
    assert rel.max() < 1e-6
    a = np.sum(Q * hp, 1)
    b = np.sum(P * hq, 1)
    assert np.all(np.abs(a - b) <= 1e-10 * (np.abs(a) + np.abs(b)))


def test_flat_hinge():
    # two triangles of a flat sheet (rest angle pi): no force, and the
    # Hessian is that of 0.5 k w phi^2 along the fold direction
    X = np.array([[0, 0, 0, 1, 0, 0, 0.3, 0, 1.0, 0.6, 0, -0.8]], float)
    P = np.zeros_like(X)
    P[0, 7] = 1.0  # lift v2 out of the plane
    g, hp = probe_hinges(X, P, math.pi, 1.0, 0.9, 1e-4)
    assert np.all(g == 0)
    assert np.all(np.isfinite(hp)) and abs(hp[0, 7]) > 0
    # the same lift by central differences of the (analytic) gradient
    eps = 1e-7
    gp, _ = probe_hinges(X + eps * P, P, math.pi, 1.0, 0.9, 1e-4)
    gm, _ = probe_hinges(X - eps * P, P, math.pi, 1.0, 0.9, 1e-4)
    assert np.allclose(hp, (gp - gm) / (2 * eps), rtol=1e-6, atol=1e-12 * np.abs(hp).max())
