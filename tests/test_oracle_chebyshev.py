"""Pins the oracle's Jacobi + Chebyshev global step (pdSolver.cu:193-199,
pdUtil.cu:195-226) against the closed form of Chebyshev semi-iteration.

On the scene of `cheb_case` the step is a linear iteration with known
contraction factors, so the position error after n local/global iterations
is P_n(0.9 g + 0.1) times the initial one, P_n evaluated through cos/acos
and cosh/acosh (see cheb_case).  A wrong switch index, a damping other than
0.9, a wrong rho or a wrong omega recurrence changes P_n by far more than the
tolerance; the last test shows that for each such mistake.
"""
import numpy as np
import pytest

import cheb_case as C
import oracle as O

ITERS = [1, 2, 5, 10, 11, 12, 13, 15, 20, 30, 50, 100]


def run(n, fp64=True, rho=C.RHO):
    s = C.scene()
    p = O.make_params(solver=O.SOLVER_JACOBI, num_iterations=n, dt=C.DT, gravity=0.0, rho=rho)
    o = O.PdOracle(s, p, fp64=fp64, V0=C.velocity())
    ex = o.export()
    o.step(1)
    X = o.get("X")
    o.close()
    return X, ex


def weight():
    dt_inv = 1.0 / C.DT
    return 1e6 * (dt_inv * dt_inv)


def test_scene_keeps_the_rotation_exact():
    # F = diag(x1, y2, z3), descending and positive on every iterate: R = I,
    # exactly in FP64; in FP32 the SVD's normalisations leave a few ulps
    for d in [(1.3, 1.15, 0.5), (0.990, 0.980, 0.75), (1.01, 0.97, 0.74), (1.0, 0.99, 0.6)]:
        R = O.local_rotation(np.diag(d), jacobi=False, fp64=True)
        assert np.array_equal(R, np.eye(3)), (d, R)
        R = O.local_rotation(np.diag(d), jacobi=False, fp64=False)
        assert np.abs(R - np.eye(3)).max() <= 4 * 2.0 ** -23, (d, R)


def test_first_iterate_and_fixed_points():
    X0, ex = run(0)
    ms = C.modes(ex["c"], ex["diagL"], weight())
    for v, ax, g, xstar, s0 in ms:
        assert X0[v, ax] == pytest.approx(s0, abs=1e-15)
    # the chosen masses and stiffness give the intended factors
    assert ms[0][2] == pytest.approx(3.6e9 / (3.636e9 + 3.6e9), rel=1e-12)
    assert ms[2][3] == pytest.approx(0.75, rel=1e-12)
    # the fixed points are where the iteration goes: a long run reaches them
    Xl, _ = run(3000)
    for v, ax, g, xstar, s0 in ms:
        assert Xl[v, ax] == pytest.approx(xstar, abs=1e-13)
    # coordinates that the iteration must not move stay put
    assert np.array_equal(Xl[0], np.zeros(3))
    assert Xl[1, 1] == 0 and Xl[1, 2] == 0 and Xl[2, 0] == 0 and Xl[2, 2] == 0
    assert Xl[3, 0] == 0 and Xl[3, 1] == 0


@pytest.mark.parametrize("n", ITERS)
def test_error_follows_the_closed_form_fp64(n):
    X, ex = run(n)
    for v, ax, g, xstar, s0 in C.modes(ex["c"], ex["diagL"], weight()):
        e0, en = s0 - xstar, X[v, ax] - xstar
        want = C.poly(n, g) * e0
        assert abs(en - want) <= 1e-14 + 1e-9 * abs(want), (n, v, en, want)


@pytest.mark.parametrize("n", [5, 12, 20, 50, 100])
def test_error_follows_the_closed_form_fp32(n):
    X, ex = run(n, fp64=False)
    for v, ax, g, xstar, s0 in C.modes(ex["c"], ex["diagL"], weight()):
        e0, en = s0 - xstar, X[v, ax] - xstar
        want = C.poly(n, g) * e0
        # float roundings of coordinates of size 1, and R = I only to a few
        # ulps (the force w (R - F) then moves the fixed point by ~1e-7)
        assert abs(en - want) <= 1e-6 * (1 + n / 10), (n, v, en, want)


@pytest.mark.parametrize("rho", [0.9, 0.99])
def test_rho_enters_as_written(rho):
    X, ex = run(40, rho=rho)
    for v, ax, g, xstar, s0 in C.modes(ex["c"], ex["diagL"], weight()):
        want = C.poly(40, g, rho=rho) * (s0 - xstar)
        assert abs(X[v, ax] - xstar - want) <= 1e-14 + 1e-9 * abs(want)


def test_plausible_mistakes_are_detected():
    """Each mistake moves the predicted iterate by much more than the tolerance
    used above, so the closed-form tests would fail on it."""
    X, ex = run(20)
    v, ax, g, xstar, s0 = C.modes(ex["c"], ex["diagL"], weight())[0]
    got = X[v, ax] - xstar
    e0 = s0 - xstar
    wrong = {
        "switch at 9": C.poly(20, g, switch=9) * e0,
        "switch at 11": C.poly(20, g, switch=11) * e0,
        "damping 0.8": C.poly(20, g, damping=0.8) * e0,
        "no damping": C.poly(20, g, damping=1.0) * e0,
        "rho 0.999": C.poly(20, g, rho=0.999) * e0,
    }
    for name, w in wrong.items():
        assert abs(got - w) > 1e3 * (1e-14 + 1e-9 * abs(w)), name
