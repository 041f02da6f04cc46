"""Pins of the IPC oracle (oracle/ipc_oracle.cpp), all on the CPU.

* squared distances and distance types: the reference's own fixture
  (tests/test_distance.cpp:30-51, tests/golden/ipc_distance_fixture.json); its
  points are printed to 6 decimals, so the recorded distances are matched to
  1e-4 relative;
* gradients and Hessians of every distance type and of the barrier: central
  finite differences (the reference's tests make the same comparison);
* makePD: numpy's eigendecomposition with the eigenvalues clamped at 1e-6;
* elastic and barrier gradients: finite differences of the total energy;
* the Newton step: the energy decreases, the step converges, the optimality
  condition holds at the solution, and two cubes pressed together keep a
  positive distance.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from sbs_amd import mesh

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "ipc_distance_fixture.json")))
T = {n: i for i, n in enumerate(O.DISTANCE_TYPES)}


def fd_grad(f, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e[i] = h
        g[i] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def fd_hess(grad, x, h=1e-6):
    H = np.zeros((x.size, x.size))
    for i in range(x.size):
        e = np.zeros_like(x)
        e[i] = h
        H[:, i] = (grad(x + e) - grad(x - e)) / (2 * h)
    return H


def test_reference_distance_fixture():
    pts = np.array(FIX["points"])
    for q in FIX["queries"]:
        x = pts[q["v"]].ravel()
        qt = O.QUERY_TYPES[q["type"]]
        d, _, _ = O.ipc_distance(qt, T[q["dtype"]], x)
        assert abs(d / q["d"] - 1) < 1e-4, (q, d)
        assert O.DISTANCE_TYPES[O.ipc_distance_type(qt, x)] == q["dtype"]


@pytest.mark.parametrize("k", range(8))
def test_fixture_derivatives_match_finite_differences(k):
    pts = np.array(FIX["points"])
    q = FIX["queries"][k]
    x = pts[q["v"]].ravel()
    qt, dt = O.QUERY_TYPES[q["type"]], T[q["dtype"]]
    d, g, H = O.ipc_distance(qt, dt, x)
    gf = fd_grad(lambda y: O.ipc_distance(qt, dt, y)[0], x, 1e-5)
    Hf = fd_hess(lambda y: O.ipc_distance(qt, dt, y)[1], x, 1e-5)
    np.testing.assert_allclose(g, gf, rtol=1e-5, atol=1e-7 * np.abs(g).max())
    np.testing.assert_allclose(H, Hf, rtol=1e-5, atol=1e-6 * np.abs(H).max())
    np.testing.assert_allclose(H, H.T, atol=1e-12 * np.abs(H).max())


@pytest.mark.parametrize("dtype", O.DISTANCE_TYPES)
def test_every_type_derivatives(dtype):
    rng = np.random.default_rng(T[dtype])
    qt = 1 if T[dtype] <= T["P_T"] else 2
    for _ in range(3):
        x = rng.normal(size=12)
        d, g, H = O.ipc_distance(qt, T[dtype], x)
        assert d >= 0
        gf = fd_grad(lambda y: O.ipc_distance(qt, T[dtype], y)[0], x)
        Hf = fd_hess(lambda y: O.ipc_distance(qt, T[dtype], y)[1], x)
        np.testing.assert_allclose(g, gf, rtol=1e-5, atol=1e-8 * max(1, np.abs(g).max()))
        np.testing.assert_allclose(H, Hf, rtol=1e-4, atol=1e-6 * max(1, np.abs(H).max()))


def test_distances_in_closed_form():
    # point above the middle of a unit right triangle in z = 0: plane distance
    x = np.array([0.2, 0.2, 0.3, 0, 0, 0, 1, 0, 0, 0, 1, 0], float)
    assert O.DISTANCE_TYPES[O.ipc_distance_type(1, x)] == "P_T"
    assert np.isclose(O.ipc_distance(1, T["P_T"], x)[0], 0.09)
    # beyond vertex 0: point-point
    x[:3] = [-1.0, -2.0, 0.5]
    assert O.DISTANCE_TYPES[O.ipc_distance_type(1, x)] == "P_T0"
    assert np.isclose(O.ipc_distance(1, T["P_T0"], x)[0], 1 + 4 + 0.25)
    # crossing skew edges: line-line distance = the offset along the common normal
    e = np.array([-1, 0, 0, 1, 0, 0, 0, -1, 0.7, 0, 1, 0.7], float)
    assert O.DISTANCE_TYPES[O.ipc_distance_type(2, e)] == "EA_EB"
    assert np.isclose(O.ipc_distance(2, T["EA_EB"], e)[0], 0.49)


def test_barrier_derivatives():
    dhat, kappa = 1e-3, 250.0
    for d in (1e-9, 3e-8, 5e-7, 9.9e-7):
        b, d1, d2 = O.ipc_barrier(d, dhat, kappa)
        h = d * 1e-5
        f = lambda y: O.ipc_barrier(y, dhat, kappa)[0]
        assert np.isclose(d1, (f(d + h) - f(d - h)) / (2 * h), rtol=1e-6)
        g = lambda y: O.ipc_barrier(y, dhat, kappa)[1]
        assert np.isclose(d2, (g(d + h) - g(d - h)) / (2 * h), rtol=1e-6)
    assert O.ipc_barrier(dhat * dhat, dhat, kappa)[0] == 0.0


def test_make_pd_matches_clamped_eigendecomposition():
    rng = np.random.default_rng(7)
    for n, shift in ((12, 0.0), (12, 5.0), (6, -1.0)):
        A = rng.normal(size=(n, n))
        A = A + A.T + shift * np.eye(n)
        w, Q = np.linalg.eigh(A)
        ref = (Q * np.maximum(w, 1e-6)) @ Q.T
        np.testing.assert_allclose(O.ipc_make_pd(A), ref, atol=1e-7 * np.abs(A).max())
    # a positive definite matrix is returned unchanged (to the rotation threshold)
    B = rng.normal(size=(12, 12))
    B = B @ B.T + np.eye(12)
    np.testing.assert_allclose(O.ipc_make_pd(B), B, atol=1e-8 * np.abs(B).max())


def scene_oracle(s, **kw):
    V = np.zeros_like(s.X)
    return O.IpcOracle(s.X, V, s.tet, s.tri, s.tri_fathers, s.mass, s.mu, s.lam, s.dbc > 0, **kw)


def test_energy_gradient_matches_finite_differences():
    s = mesh.stack_scene(2, gap=4e-4)
    o = scene_oracle(s, dt=0.01, dhat=1e-3)
    rng = np.random.default_rng(3)
    X = s.X + rng.normal(scale=2e-3, size=s.X.shape)
    X[s.body_vert_start[1]:, 1] -= 1e-4      # bring the faces within dhat
    o.set_state(s.X + 1e-3, 350.0)
    e, g = o.energy(X, grad=True)
    q, dt, d = o.queries()
    assert (d < 1e-6).any()                  # some barrier terms are active
    x = X.ravel()
    for i in rng.choice(x.size, 24, replace=False):
        h = 1e-9  # the closest active pair is 4e-6 m apart: the barrier is steep there
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        num = (o.energy(xp) - o.energy(xm)) / (2 * h)
        assert np.isclose(g.ravel()[i], num, rtol=2e-5, atol=1e-7 * np.abs(g).max()), i


def test_newton_step_reaches_a_stationary_point():
    s = mesh.cube_scene(2, pin_top=True)
    o = scene_oracle(s, dt=0.01, dhat=1e-3, tol=1e-9)
    assert o.step()
    X, V = o.get()
    e, g = o.energy(X, grad=True)
    free = s.dbc == 0
    assert np.abs(g[free]).max() < 1e-8 * np.abs(s.mass).max()
    assert o.stats()["newton_iterations"] >= 1


def test_contact_keeps_bodies_apart():
    s = mesh.stack_scene(2, gap=1.5e-3)
    o = scene_oracle(s, dt=0.01, dhat=1e-3, tol=1e-3)
    for _ in range(10):
        assert o.step()
        _, _, d = o.queries()
        assert d.min() > 0.0
    X, V = o.get()
    up = s.body_vert_start[1]
    assert (s.X[up:, 1] - X[up:, 1]).min() > 1.5e-3      # the upper cube came down onto the lower one
    assert o.stats()["kappa"] > 100.0                      # and the barrier stiffness adapted to the contact


# ------------------------------------------------------------ fixed bodies --
def fixed_setup(s, bodies, **kw):
    from sbs_amd import fixed as F
    o = scene_oracle(s, **kw)
    types = [b.type for b in bodies]
    models = np.stack([np.asarray(b.model, np.float32).T.reshape(-1) for b in bodies])
    o.set_fixed(types, O.fixed_params(types, models, [b.radius for b in bodies]))
    return o


def test_contact_area_closed_form():
    # one unit right tet: vertex 0 touches three faces of area 1/2, vertex 1
    # two of them and the slanted face (sqrt(3)/2)
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    o = O.IpcOracle(X, np.zeros_like(X), [[0, 1, 2, 3]], np.zeros((0, 3)), [], [1.0] * 4, [1.0], [1.0],
                    np.zeros(4))
    a = o.contact_area()
    np.testing.assert_allclose(a, [0.5] + [(1 + np.sqrt(3) / 2) / 3] * 3, rtol=1e-15)


def test_fixed_barrier_derivatives():
    from sbs_amd import fixed as F
    s = mesh.cube_scene(2, pin_top=False)
    # every vertex stays outside the bodies (jitter 1e-3) and some are within dhat
    bodies = [F.Plane((0, -2e-3, 0)), F.Sphere((0.1, 0.1, -0.053), radius=0.05),
              F.Cylinder((0.25, 0.0, 0.1), (90, 0, 0), (0.045, 1, 0.045))]
    o = fixed_setup(s, bodies, dt=0.01, dhat=1e-2)
    X = s.X.copy()
    hit = 0
    for v in range(s.nv):
        e, g, H = o.fixed_vertex(X, v)
        if e == 0.0:
            continue
        hit += 1
        x = X.ravel().copy()
        h = 1e-8
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[3 * v + k] += h
            xm[3 * v + k] -= h
            num = (o.fixed_vertex(xp, v)[0] - o.fixed_vertex(xm, v)[0]) / (2 * h)
            assert np.isclose(g[k], num, rtol=1e-5, atol=1e-9 * np.abs(g).max()), (v, k)
        np.testing.assert_allclose(H, H.T, atol=1e-12 * np.abs(H).max())
        assert np.linalg.eigvalsh(H).min() >= 1e-6 - 1e-12 * np.abs(H).max()   # made PD: clamped at 1e-6
    assert hit >= 4
    # a bottom vertex sees only the floor (normal +y): its Hessian is the exact
    # second derivative along y, and 1e-6 (the clamp) across
    v = int(np.argmin(X[:, 1]))
    _, g0, H = o.fixed_vertex(X, v)
    xp = X.ravel().copy()
    xp[3 * v + 1] += 1e-10
    _, g1, _ = o.fixed_vertex(xp, v)
    assert np.isclose(H[1, 1], (g1[1] - g0[1]) / 1e-10, rtol=1e-4)
    np.testing.assert_allclose([H[0, 0], H[2, 2]], [1e-6, 1e-6], rtol=1e-6)


def test_fixed_step_closed_form():
    from sbs_amd import fixed as F
    X = np.array([[0.0, 1.0, 0.0], [3.0, 5.0, 0.0], [10.0, 7.0, 0.0]])
    o = O.IpcOracle(X, np.zeros_like(X), np.zeros((0, 4)), np.zeros((0, 3)), [], [1.0] * 3, [], [], np.zeros(3))
    bodies = [F.Plane((0, 0, 0)), F.Sphere((0, 5, 0), radius=1.0), F.Cylinder((10, 0, 3), (0, 0, 0), (1, 1, 1))]
    types = [b.type for b in bodies]
    models = np.stack([np.asarray(b.model, np.float32).T.reshape(-1) for b in bodies])
    o.set_fixed(types, O.fixed_params(types, models, [b.radius for b in bodies]))
    # vertex 0 falls onto the plane: x - alpha p reaches y = 0 at alpha = 1/2
    p = np.array([[0, 2.0, 0], [0, 0, 0], [0, 0, 0]])
    assert np.isclose(o.fixed_step(X, p), 0.45)
    # vertex 1 moves towards the sphere: reaches its surface after 2 of 4 units
    p = np.array([[0, 0, 0], [4.0, 0, 0], [0, 0, 0]])
    assert np.isclose(o.fixed_step(X, p), 0.9 * 0.5)
    # vertex 2 moves towards the y-axis cylinder (radial distance 3) with an
    # axial part that does not change the hit: 2 of 3 radial units
    p = np.array([[0, 0, 0], [0, 0, 0], [0, -5.0, -3.0]])
    assert np.isclose(o.fixed_step(X, p), 0.9 * 2 / 3)
    # moving away from everything, or on a ray that misses the sphere: 1
    p = np.array([[0, -1.0, 0], [-1.0, 0, 5.0], [0, 0, -1.0]])
    assert o.fixed_step(X, p) == 1.0


def test_cube_rests_on_a_floor():
    from sbs_amd import fixed as F
    s = mesh.cube_scene(2, pin_top=False)
    X0 = s.X.copy()
    X0[:, 1] += 3e-3 - X0[:, 1].min()                 # 3 mm above the floor
    s.X[:] = X0
    o = fixed_setup(s, [F.Plane((0, 0, 0))], dt=0.01, dhat=1e-2, tol=1e-3)
    for _ in range(15):
        assert o.step()
        X, _ = o.get()
        assert X[:, 1].min() > 0.0                     # never through the floor
    assert X[:, 1].min() < 3e-3                        # it came down onto it
    e, g = o.energy(X, grad=True)
    assert np.isfinite(e)
