"""The product's PLIS state machine (host drive) against the reference's own LuksanPLIS compiled from
potential/LuksanPLIS.cpp under oracle/_ref: identical iterates, bit for bit, on the problem families of
potential/tests/luksan_plis (chained Rosenbrock, stiff quadratic + quartic) and on a tiny stealthy relax."""
import ctypes as C

import numpy as np
import pytest

from modules_amd import _lib
from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")


def product_minimize(fun, x0, stopval=-1e10, maxeval=0, mf=40):
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    n = len(x)
    trace = []

    @_lib.OBJECTIVE
    def obj(nn, xp, gp, data):
        xv = np.ctypeslib.as_array(xp, shape=(nn,)).copy()
        f, g = fun(xv)
        np.ctypeslib.as_array(gp, shape=(nn,))[:] = g
        trace.append((xv, f))
        return f

    minf = C.c_double(0)
    ev = C.c_int(0)
    st = _lib.lib().ccgpu_lbfgs_minimize(n, C.cast(obj, C.c_void_p), None, x.ctypes.data_as(_lib.dp),
                                         C.byref(minf), stopval, maxeval, mf, C.byref(ev))
    return x, minf.value, st, ev.value, trace


def rosenbrock(x):
    f = 0.0
    g = np.zeros_like(x)
    for i in range(len(x) - 1):
        a = x[i + 1] - x[i] * x[i]
        b = 1.0 - x[i]
        f += 100.0 * a * a + b * b
        g[i] += -400.0 * x[i] * a - 2.0 * b
        g[i + 1] += 200.0 * a
    return f, g


def stiff(x):
    w = np.logspace(0, 4, len(x))
    f = float(np.sum(w * x * x) + 0.25 * np.sum(x ** 4))
    return f, 2 * w * x + x ** 3


@pytest.mark.parametrize("n,mf,seed", [(2, 5, 0), (10, 3, 1), (10, 40, 2), (25, 7, 3)])
@pytest.mark.parametrize("fun", [rosenbrock, stiff])
def test_bitwise_equal_to_reference(fun, n, mf, seed):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-1.5, 1.5, size=n)
    xr, fr, sr, er, tr = po.ref_plis(fun, x0, mf=mf, maxeval=3000)
    xp, fp, sp, ep, tp = product_minimize(fun, x0, mf=mf, maxeval=3000)
    assert (sr, er) == (sp, ep)
    assert fr == fp
    assert np.array_equal(xr, xp)
    assert len(tr) == len(tp)
    for (a, fa), (b, fb) in zip(tr, tp):
        assert np.array_equal(a, b) and fa == fb


def steep_cosh(x):
    # the unit first step overshoots to an overflowing objective: bisection below the huge-value bound
    f = float(np.sum(np.cosh(40.0 * x)))
    return f, 40.0 * np.sinh(40.0 * x)


def rastrigin(x):
    # negative curvature along many directions: extrapolation, interpolation and failed searches (restarts)
    f = float(10.0 * len(x) + np.sum(x * x - 10.0 * np.cos(2 * np.pi * x)))
    return f, 2 * x + 20.0 * np.pi * np.sin(2 * np.pi * x)


def quartic(x):
    # a flat minimum: ends on the relative x or f change tests
    w = np.arange(1, len(x) + 1, dtype=float)
    return float(np.sum(w * x ** 4)), 4 * w * x ** 3


@pytest.mark.parametrize("fun,n,mf,seed", [(steep_cosh, 3, 4, 5), (steep_cosh, 8, 40, 6), (rastrigin, 6, 5, 7),
                                           (rastrigin, 12, 40, 8), (quartic, 5, 1, 9), (quartic, 9, 10, 10),
                                           (rosenbrock, 6, 1, 11)])
def test_bitwise_equal_on_hard_paths(fun, n, mf, seed):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-1.0, 1.0, size=n)
    with np.errstate(over="ignore", invalid="ignore"):
        xr, fr, sr, er, tr = po.ref_plis(fun, x0, mf=mf, maxeval=5000)
        xp, fp, sp, ep, tp = product_minimize(fun, x0, mf=mf, maxeval=5000)
    assert (sr, er) == (sp, ep)
    assert fr == fp or (np.isnan(fr) and np.isnan(fp))
    assert np.array_equal(xr, xp, equal_nan=True)
    assert len(tr) == len(tp)
    for (a, fa), (b, fb) in zip(tr, tp):
        assert np.array_equal(a, b, equal_nan=True) and (fa == fb or (np.isnan(fa) and np.isnan(fb)))


def test_stopval_and_maxeval_codes():
    x0 = np.array([-1.2, 1.0, -1.2, 1.0])
    for kw in ({"stopval": 1e-3}, {"maxeval": 17}):
        r = po.ref_plis(rosenbrock, x0, **kw)
        p = product_minimize(rosenbrock, x0, **kw)
        assert r[2] == p[2] and r[3] == p[3] and np.array_equal(r[0], p[0])


def test_stealthy_objective_bitwise():
    # pass 2 (rescale 1) of a tiny stealthy relax, objective from the oracle
    N, d = 40, 2
    s = po.make_config(2, N, d)
    a = np.eye(d) * np.sqrt(N)
    k, _ = po.get_ks(a, po.getK(0.3, 1.0, d))
    v = np.ones(len(k))

    def fun(x):
        ss = (x.reshape(N, d)) % 1.0
        E, F = po.total_energy_forces(ss, a, k, v, nthreads=1)
        return E, -(F @ a.T).ravel()

    r = po.ref_plis(fun, s.ravel(), mf=40, maxeval=400)
    p = product_minimize(fun, s.ravel(), mf=40, maxeval=400)
    assert r[2] == p[2] and r[3] == p[3] and r[1] == p[1] and np.array_equal(r[0], p[0])
