"""Pins of the oracle's Neo-Hookean elasticity (NeoHookean08Energy,
neohookean08.cu:99-206), all on the CPU.

* rest state: zero energy and stress at F = I; frame indifference;
* uniform dilation F = s I in closed form;
* the first Piola stress equals central finite differences of the density;
* the 9x9 Hessian equals the exact second derivative (finite differences of
  P) minus the one term the reference's kernel forms but never adds,
  (lambda ln J - mu)/J dC/dF with C = dJ/dF = J F^-T (itself checked against
  finite differences of det);
* the total IPC energy with Neo-Hookean elasticity: its gradient matches
  finite differences, and a converged Newton step is a stationary point.
"""
import numpy as np
import pytest

import oracle as O
from sbs_amd import mesh

MU, LAM = 35714.29, 142857.1   # E = 1e5, nu = 0.4


def rand_F(rng, amp=0.3):
    while True:
        F = np.eye(3) + rng.uniform(-amp, amp, (3, 3))
        if np.linalg.det(F) > 0.3:
            return F


def vec(M):
    return M.T.reshape(-1)        # index 3c + r


def test_rest_state_and_frame_indifference():
    e, P, D = O.nh_density(np.eye(3), MU, LAM)
    assert e == 0.0 and np.abs(P).max() == 0.0
    rng = np.random.default_rng(1)
    for _ in range(5):
        F = rand_F(rng)
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] = -Q[:, 0]
        e1, P1, _ = O.nh_density(F, MU, LAM)
        e2, P2, _ = O.nh_density(Q @ F, MU, LAM)
        assert np.isclose(e1, e2, rtol=1e-12)
        np.testing.assert_allclose(P2, Q @ P1, rtol=1e-10, atol=1e-9 * np.abs(P1).max())


@pytest.mark.parametrize("s", [0.8, 0.95, 1.1, 1.5])
def test_uniform_dilation_closed_form(s):
    e, P, _ = O.nh_density(s * np.eye(3), MU, LAM)
    ls = np.log(s)
    assert np.isclose(e, MU / 2 * (3 * s * s - 3) - 3 * MU * ls + LAM / 2 * (3 * ls) ** 2, rtol=1e-12)
    np.testing.assert_allclose(P, (MU * (s - 1 / s) + 3 * LAM * ls / s) * np.eye(3), rtol=1e-12, atol=1e-6)


def test_stress_is_the_density_gradient():
    rng = np.random.default_rng(2)
    for _ in range(6):
        F = rand_F(rng)
        _, P, _ = O.nh_density(F, MU, LAM)
        h = 1e-6
        num = np.zeros((3, 3))
        for r in range(3):
            for c in range(3):
                E = np.zeros((3, 3))
                E[r, c] = h
                num[r, c] = (O.nh_density(F + E, MU, LAM)[0] - O.nh_density(F - E, MU, LAM)[0]) / (2 * h)
        np.testing.assert_allclose(P, num, rtol=1e-6, atol=1e-6 * np.abs(P).max())


def test_hessian_is_the_exact_one_without_the_cofactor_derivative():
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(6):
        F = rand_F(rng)
        J = np.linalg.det(F)
        C = J * np.linalg.inv(F).T
        # C = dJ/dF, checked against finite differences of det
        dJ = np.zeros((3, 3))
        for r in range(3):
            for c in range(3):
                E = np.zeros((3, 3))
                E[r, c] = h
                dJ[r, c] = (np.linalg.det(F + E) - np.linalg.det(F - E)) / (2 * h)
        np.testing.assert_allclose(C, dJ, rtol=1e-7, atol=1e-9)
        exact = np.zeros((9, 9))   # d vec P / d vec F, by finite differences of P
        dC = np.zeros((9, 9))      # d vec C / d vec F
        for b in range(9):
            E = np.zeros((3, 3))
            E[b % 3, b // 3] = h
            exact[:, b] = (vec(O.nh_density(F + E, MU, LAM)[1]) - vec(O.nh_density(F - E, MU, LAM)[1])) / (2 * h)
            Fp, Fm = F + E, F - E
            dC[:, b] = (vec(np.linalg.det(Fp) * np.linalg.inv(Fp).T) - vec(np.linalg.det(Fm) * np.linalg.inv(Fm).T)) / (2 * h)
        _, _, D = O.nh_density(F, MU, LAM)
        beta = (LAM * np.log(J) - MU) / J
        np.testing.assert_allclose(D, exact - beta * dC, rtol=1e-6, atol=1e-6 * np.abs(exact).max())
        # and the omitted term is not negligible: D is not the exact Hessian
        assert np.abs(D - exact).max() > 1e-3 * np.abs(exact).max()
        # D is symmetric positive definite here (mu I + alpha C C^T, alpha > 0)
        assert np.allclose(D, D.T) and np.linalg.eigvalsh(D).min() > 0


def nh_oracle(s, **kw):
    return O.IpcOracle(s.X, np.zeros_like(s.X), s.tet, s.tri, s.tri_fathers, s.mass, s.mu, s.lam, s.dbc > 0,
                       elastic="neohookean", **kw)


def test_total_energy_gradient_with_neo_hookean():
    s = mesh.stack_scene(2, gap=4e-4, nu=0.4)
    o = nh_oracle(s, dt=0.01, dhat=1e-3)
    rng = np.random.default_rng(4)
    X = s.X + rng.normal(scale=2e-3, size=s.X.shape)
    X[s.body_vert_start[1]:, 1] -= 1e-4
    o.set_state(s.X + 1e-3, 350.0)
    e, g = o.energy(X, grad=True)
    # the elastic part differs from the corotated energy of the same state
    oc = O.IpcOracle(s.X, np.zeros_like(s.X), s.tet, s.tri, s.tri_fathers, s.mass, s.mu, s.lam, s.dbc > 0,
                     dt=0.01, dhat=1e-3)
    oc.set_state(s.X + 1e-3, 350.0)
    assert abs(oc.energy(X) - e) > 1e-9 * abs(e)
    x = X.ravel()
    for i in rng.choice(x.size, 24, replace=False):
        h = 1e-9
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        num = (o.energy(xp) - o.energy(xm)) / (2 * h)
        assert np.isclose(g.ravel()[i], num, rtol=2e-5, atol=1e-7 * np.abs(g).max()), i


def test_neo_hookean_newton_step_reaches_a_stationary_point():
    s = mesh.cube_scene(2, pin_top=True, nu=0.4)
    o = nh_oracle(s, dt=0.01, dhat=1e-3, tol=1e-9)
    assert o.step()
    X, _ = o.get()
    _, g = o.energy(X, grad=True)
    free = s.dbc == 0
    assert np.abs(g[free]).max() < 1e-8 * np.abs(s.mass).max()
    assert o.stats()["newton_iterations"] >= 1
