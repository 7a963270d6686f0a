"""Pin the oracle's JavaScript Math restatements (fdlibm / V8 algorithms)
against the C library: at most one ulp apart everywhere, exact at the
special values the reference relies on."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle_ffi as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _mathlib():
    # Small test shim exposing the oracle's jsmath.h routines (built from the
    # oracle directory by its Makefile target `mathprobe`).
    import subprocess
    so = os.path.join(O.HERE, "_build", "libjsmath_probe.so")
    subprocess.run(["make", "-s", "-C", O.HERE, "mathprobe"], check=True)
    L = C.CDLL(so)
    f = L.orc_math_eval
    f.restype = None
    f.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64]
    g = L.orc_cr_dot
    g.restype = C.c_double
    g.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64]
    return L


FN = {"log": 0, "acos": 1, "cbrt": 2, "cos": 3}
REF = {"log": math.log, "acos": math.acos, "cbrt": lambda v: math.copysign(abs(v) ** (1 / 3), v) if False else None,
       "cos": math.cos}


def evalf(L, name, xs):
    xs = np.ascontiguousarray(xs, np.float64)
    out = np.zeros_like(xs)
    L.orc_math_eval(FN[name], O.P(xs), O.P(out), len(xs))
    return out


def ulps(a, b):
    a = np.asarray(a); b = np.asarray(b)
    ia = a.view(np.int64); ib = b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-(2**63)) - ia, ia)
    ib = np.where(ib < 0, np.int64(-(2**63)) - ib, ib)
    return np.abs(ia - ib)


def test_fdlibm_restatements_within_one_ulp_of_libm():
    L = _mathlib()
    rng = np.random.default_rng(11)
    x = np.concatenate([10 ** rng.uniform(-300, 300, 20000), rng.uniform(1e-4, 4, 20000),
                        1 + rng.uniform(-1e-6, 1e-6, 2000), [1.0, 2.0, 0.5, 1e-308, 5e-324]])
    assert ulps(evalf(L, "log", x), np.log(x)).max() <= 1
    c = np.concatenate([rng.uniform(-1, 1, 20000), -1 + 10 ** rng.uniform(-16, -1, 5000),
                        1 - 10 ** rng.uniform(-16, -1, 5000), [-1.0, 1.0, 0.0, 0.5, -0.5]])
    ref = np.array([math.acos(v) for v in c])
    assert ulps(evalf(L, "acos", c), ref).max() <= 1
    y = np.concatenate([rng.uniform(-10, 10, 20000), 10 ** rng.uniform(-300, 300, 5000),
                        -(10 ** rng.uniform(-300, 300, 5000)), [8.0, 27.0, -8.0, 1e-310]])
    assert ulps(evalf(L, "cbrt", y), np.cbrt(y)).max() <= 1
    a = np.concatenate([rng.uniform(-7, 7, 20000), rng.uniform(0, 5.3, 20000), [0.0, math.pi, math.pi / 2]])
    ref = np.array([math.cos(v) for v in a])
    assert ulps(evalf(L, "cos", a), ref).max() <= 1
    # special values
    assert evalf(L, "acos", [-1.0])[0] == math.pi and evalf(L, "acos", [1.0])[0] == 0.0
    assert evalf(L, "log", [1.0])[0] == 0.0 and evalf(L, "log", [0.0])[0] == -math.inf
    assert math.isnan(evalf(L, "log", [-1.0])[0]) and math.isnan(evalf(L, "acos", [1.5])[0])
    assert evalf(L, "cbrt", [-27.0])[0] == -3.0 and evalf(L, "cbrt", [0.0])[0] == 0.0


def test_correctly_rounded_dot_is_order_independent():
    L = _mathlib()
    rng = np.random.default_rng(4)
    from fractions import Fraction
    for n in (10, 1000, 50000):
        a = rng.standard_normal(n) * 10 ** rng.uniform(-5, 5, n)
        b = rng.standard_normal(n)
        exact = float(sum(Fraction(x) * Fraction(y) for x, y in zip(a.tolist(), b.tolist())))
        got = L.orc_cr_dot(O.P(a), O.P(b), n)
        perm = rng.permutation(n)
        got2 = L.orc_cr_dot(O.P(np.ascontiguousarray(a[perm])), O.P(np.ascontiguousarray(b[perm])), n)
        assert got == exact and got2 == exact
