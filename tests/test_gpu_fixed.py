"""GPU parity of the fixed bodies (FixedBodyData::HandleCollisions,
src/simulation/fixedBodyData.cu:68-148) through the C ABI against the oracle
(oracle/fixed_oracle.cpp).

Bars: body parameters within 2e-6 of the oracle's (two float routes to the
plane normal: adjugate in the library, pivoted elimination in the oracle);
given one set of parameters, positions and velocities after the response
bit-equal in FP32 and FP64 (both sides evaluate the same IEEE sequence without
contraction); in a PD update the response is applied to XTilde and V after X
is taken (pdSolver.cu:226-231), checked against the oracle applied to the same
update without bodies (Jacobi global step: deterministic).
"""
import numpy as np
import pytest

import oracle as O
from sbs_amd import fixed as F
from sbs_amd import mesh
from sbs_amd.capi import DeviceArray
from sbs_amd.pd import PdSolver, SolverData, SolverParams, SolverType

pytestmark = pytest.mark.gpu


def scene_bodies():
    return [F.Plane((0, 0.05, 0), (5, 0, -3), (2, 2, 2)),
            F.Sphere((0.4, 0.3, 0.5), radius=0.25),
            F.Cylinder((0.1, 0.2, 0.1), (90, 0, 20), (0.15, 1.0, 0.15)),
            F.Plane((0, 1.0, 0), (0, 0, 180), (3, 3, 3)),
            F.Sphere((0.0, 0.7, 0.0), radius=0.2),
            F.Plane((-0.05, 0, 0), (0, 0, -90), (1, 1, 1)),
            F.Cylinder((0.8, 0.5, 0.8), (0, 0, 0), (0.3, 2.0, 0.3))]


def test_params_match_oracle(gpu):
    bodies = scene_bodies() + [F.Plane(p, r, s) for p, r, s in
                               [((0, 450, 0), (0, 0, 180), (450, 450, 450)), ((0, 0, -300), (90, 0, 0), (450, 450, 450))]]
    fb = F.FixedBodyData(bodies)
    try:
        got = fb.params()
    finally:
        fb.close()
    types = [b.type for b in bodies]
    models = np.stack([np.asarray(b.model, np.float32).T.reshape(-1) for b in bodies])
    ref = O.fixed_params(types, models, [b.radius for b in bodies])
    np.testing.assert_array_equal(got[:, [0, 1, 2, 6]], ref[:, [0, 1, 2, 6]])
    np.testing.assert_allclose(got[:, 3:6], ref[:, 3:6], atol=2e-6)


@pytest.mark.parametrize("precision", [32, 64])
def test_handle_collisions_bit_exact(gpu, precision):
    fp64 = precision == 64
    dt = np.float64 if fp64 else np.float32
    rng = np.random.default_rng(11)
    n = 20000
    X = rng.uniform(-0.2, 1.1, size=(n, 3)).astype(dt)
    V = rng.normal(scale=2.0, size=(n, 3)).astype(dt)
    V[:50] = 0.0                                     # |v_T| = 0 cases
    bodies = scene_bodies()
    fb = F.FixedBodyData(bodies)
    dX, dV = DeviceArray.from_numpy(X.reshape(-1)), DeviceArray.from_numpy(V.reshape(-1))
    try:
        prm = fb.params()
        fb.HandleCollisions(dX.ptr, dV.ptr, n, muT=0.4, muN=0.6, precision=precision)
        Xg = dX.download(dt, (n, 3))
        Vg = dV.download(dt, (n, 3))
    finally:
        dX.free(), dV.free(), fb.close()
    Xo, Vo = O.fixed_handle(X, V, [b.type for b in bodies], prm, 0.4, 0.6, fp64=fp64)
    moved = np.any(Xo != X, axis=1)
    assert moved.sum() > 1000                        # the bodies were engaged
    np.testing.assert_array_equal(Xg, Xo)
    np.testing.assert_array_equal(Vg, Vo)


def test_empty_inputs(gpu):
    fb = F.FixedBodyData([])
    X = np.ones((4, 3), np.float32)
    dX, dV = DeviceArray.from_numpy(X.reshape(-1)), DeviceArray.from_numpy(-X.reshape(-1))
    try:
        fb.HandleCollisions(dX.ptr, dV.ptr, 4, 0.5, 0.5)        # no bodies: nothing changes
        np.testing.assert_array_equal(dX.download(np.float32, (4, 3)), X)
        fb2 = F.FixedBodyData([F.Plane((0, 2, 0))])
        fb2.HandleCollisions(dX.ptr, dV.ptr, 0, 0.5, 0.5)       # no vertices: nothing to do
        fb2.HandleCollisions(dX.ptr, dV.ptr, 4, 0.5, 0.5)
        np.testing.assert_allclose(dX.download(np.float32, (4, 3))[:, 1], 2.0)
        fb2.close()
    finally:
        dX.free(), dV.free(), fb.close()


def test_pd_update_applies_fixed_bodies_after_the_step(gpu):
    s = mesh.cube_scene(4, pin_top=False)
    V0 = np.zeros_like(s.X)
    V0[:, 1] = -2.0                                  # the bottom layer goes through the floor in one step
    floor = F.Plane((0, -0.01, 0))
    wall = F.Sphere((0.2, 0.0, 0.2), radius=0.06)
    p = SolverParams(numIterations=10, dt=1.0 / 60.0, muN=0.5, muT=0.3)
    runs = {}
    for name, bodies in (("bare", None), ("fixed", [floor, wall])):
        data = SolverData(s, 32, V0=V0)
        fb = F.FixedBodyData(bodies) if bodies else None
        data.set_fixed_bodies(fb)
        solver = PdSolver(128, data)
        solver.SetGlobalSolver(SolverType.Jacobi)
        solver.SetPerf(True)
        solver.Update(data, p)
        runs[name] = {k: data.get(k) for k in ("X", "XTilde", "V")}
        runs[name]["perf"] = dict(solver.GetPerformanceData())
        runs[name]["prm"] = fb.params() if fb else None
        solver.close()
        data.free()
        if fb:
            fb.close()
    bare, fx = runs["bare"], runs["fixed"]
    np.testing.assert_array_equal(fx["X"], bare["XTilde"])   # X taken before the bodies act
    Xo, Vo = O.fixed_handle(bare["XTilde"], bare["V"], [floor.type, wall.type], fx["prm"], 0.3, 0.5)
    assert (Xo != bare["XTilde"]).any(axis=1).sum() >= 25   # the bottom layer hit the floor
    np.testing.assert_array_equal(fx["XTilde"], Xo)
    np.testing.assert_array_equal(fx["V"], Vo)
    assert fx["perf"]["collision handling(fixed)"] > 0.0
