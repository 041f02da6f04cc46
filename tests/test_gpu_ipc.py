"""GPU parity of the FP64 IPC step (row a16) through the C ABI against the IPC
oracle (oracle/ipc_oracle.cpp), whose derivatives come from a different
derivation (forward jets there, closed forms here).

Bars (FP64): energies within 1e-12 relative, gradients within 1e-10 of their
largest entry, positions after each step within 1e-10 relative (max norm
over the displacement scale), the same Newton iteration and line-search
counts, and the barrier stiffness within 1e-8 (it is a ratio of gradient
maxima dominated by the closest pair, where the barrier gradient grows like
1/d and amplifies rounding-level position differences).
"""
import numpy as np
import pytest

import oracle as O
from sbs_amd import capi, mesh
from sbs_amd.ipc import IpcData, IpcParams, IpcSolver, LinearSolverType

pytestmark = pytest.mark.gpu


def oracle_for(s, p: IpcParams):
    return O.IpcOracle(s.X, np.zeros_like(s.X), s.tet, s.tri, s.tri_fathers, s.mass, s.mu, s.lam, s.dbc > 0,
                       dt=p.dt, damp=p.damp, dhat=p.dhat, tol=p.tol, lin_tol=p.lin_tol, max_iter=p.maxIterations)


def test_energy_and_gradient_match_oracle(gpu):
    s = mesh.stack_scene(2, gap=4e-4)
    p = IpcParams(dt=0.01, dhat=1e-3)
    data = IpcData(s)
    solver = IpcSolver(data)
    o = oracle_for(s, p)
    rng = np.random.default_rng(3)
    X = s.X + rng.normal(scale=2e-3, size=s.X.shape)
    X[s.body_vert_start[1]:, 1] -= 1e-4
    XT = s.X + 1e-3
    solver.set_state(XT, 350.0)
    o.set_state(XT, 350.0)
    e_gpu, g_gpu = solver.energy(X, grad=True)
    e_orc, g_orc = o.energy(X, grad=True)
    _, _, d = o.queries()
    assert (d < p.dhat ** 2).sum() > 0
    assert abs(e_gpu - e_orc) <= 1e-12 * abs(e_orc)
    assert np.abs(g_gpu - g_orc).max() <= 1e-10 * np.abs(g_orc).max()
    solver.close()
    data.free()


def run_pair(s, p, nsteps):
    data = IpcData(s)
    solver = IpcSolver(data)
    o = oracle_for(s, p)
    out = []
    for k in range(nsteps):
        ok_g = solver.Update(data, p)
        ok_o = o.step()
        assert ok_g and ok_o, (k, solver.stats(), o.stats())
        Xg, Xo = data.get("X"), o.get()[0]
        out.append((Xg, Xo, solver.stats(), o.stats()))
    Vg, Vo = data.get("V"), o.get()[1]
    solver.close()
    data.free()
    return out, Vg, Vo


def test_free_fall_steps_match_oracle(gpu):
    s = mesh.cube_scene(3, pin_top=False)
    p = IpcParams(dt=0.01, dhat=1e-3, tol=1e-6)
    out, Vg, Vo = run_pair(s, p, 3)
    for Xg, Xo, sg, so in out:
        scale = max(np.abs(Xo - s.X).max(), 1e-6)
        assert np.abs(Xg - Xo).max() <= 1e-10 * scale
        assert sg["newton_iterations"] == so["newton_iterations"]
    assert np.abs(Vg - Vo).max() <= 1e-8 * np.abs(Vo).max()


def test_contact_steps_match_oracle(gpu):
    s = mesh.stack_scene(2, gap=1.5e-3)
    p = IpcParams(dt=0.01, dhat=1e-3, tol=1e-3)
    out, Vg, Vo = run_pair(s, p, 10)
    active = 0
    for Xg, Xo, sg, so in out:
        scale = max(np.abs(Xo - s.X).max(), 1e-6)
        assert np.abs(Xg - Xo).max() <= 1e-10 * scale, (np.abs(Xg - Xo).max(), scale, sg, so)
        assert sg["newton_iterations"] == so["newton_iterations"]
        assert sg["line_search_halvings"] == so["line_search_halvings"]
        # kappa is a ratio of gradient maxima whose barrier part grows like 1/d
        # at the closest pair: position rounding differences reach it amplified
        assert np.isclose(sg["kappa"], so["kappa"], rtol=1e-8)
        active += sg["kappa"] > 100.0
    assert active > 0           # the barrier was engaged


def test_failure_latch_and_reset(gpu):
    s = mesh.stack_scene(2, gap=1.5e-3)
    data = IpcData(s)
    solver = IpcSolver(data)
    strict = IpcParams(dt=0.01, dhat=1e-3, tol=1e-12, maxIterations=1)
    assert not solver.Update(data, strict)
    assert solver.stats()["failed"]
    X_before = data.get("X")
    assert not solver.Update(data, IpcParams(dt=0.01, dhat=1e-3, tol=1e-3))   # latched: no step taken
    assert np.array_equal(data.get("X"), X_before)
    solver.Reset()
    assert solver.Update(data, IpcParams(dt=0.01, dhat=1e-3, tol=1e-3))
    solver.SetLinearSolver(LinearSolverType.PCGJacobi)
    with pytest.raises(capi.SbsError):
        solver.SetLinearSolver(LinearSolverType.Jacobi)
    solver.close()
    data.free()


# ------------------------------------------------------------ fixed bodies --
def fixed_pair(s, bodies, p):
    from sbs_amd import fixed as F
    fb = F.FixedBodyData(bodies)
    data = IpcData(s, fixed=fb)
    solver = IpcSolver(data)
    o = oracle_for(s, p)
    o.set_fixed([b.type for b in bodies], fb.params())          # one set of float parameters
    return fb, data, solver, o


def test_fixed_barrier_energy_and_gradient_match_oracle(gpu):
    from sbs_amd import fixed as F
    # the energy entry point evaluates with the default dhat (1e-3) until a step
    # sets the parameters: bodies 0.5 mm from the vertices (jitter 0.1 mm)
    s = mesh.cube_scene(3, pin_top=False, jitter=1e-4)
    bodies = [F.Plane((0, -5e-4, 0)), F.Sphere((0.1, 0.1, -0.0505), radius=0.05),
              F.Cylinder((0.35, 0.0, 0.1), (90, 0, 0), (0.0495, 1, 0.0495))]
    p = IpcParams(dt=0.01, dhat=1e-3)
    fb, data, solver, o = fixed_pair(s, bodies, p)
    try:
        X = s.X.copy()
        XT = s.X + 1e-3
        solver.set_state(XT, 250.0)
        o.set_state(XT, 250.0)
        e_gpu, g_gpu = solver.energy(X, grad=True)
        e_orc, g_orc = o.energy(X, grad=True)
        active = sum(o.fixed_vertex(X, v)[0] != 0.0 for v in range(s.nv))
        assert active >= 10
        assert abs(e_gpu - e_orc) <= 1e-12 * abs(e_orc)
        # at the rest shape F = I up to rounding, so the elastic term leaves
        # ~1e-14 of rounding noise per vertex (2 mu V0 |Dm^-1| h^2 eps); the
        # barrier terms themselves are ~4e-5 here
        assert np.abs(g_gpu - g_orc).max() <= 1e-10 * np.abs(g_orc).max() + 2e-13
    finally:
        solver.close(), data.free(), fb.close()


def test_cube_falling_on_a_floor_matches_oracle(gpu):
    from sbs_amd import fixed as F
    s = mesh.cube_scene(3, pin_top=False)
    s.X[:, 1] += 4e-3 - s.X[:, 1].min()
    bodies = [F.Plane((0, 0, 0)), F.Plane((-0.01, 0, 0), (0, 0, -90))]   # a floor and a side wall
    p = IpcParams(dt=0.01, dhat=1e-2, tol=1e-3)
    fb, data, solver, o = fixed_pair(s, bodies, p)
    try:
        for k in range(8):
            assert solver.Update(data, p) and o.step(), k
            Xg, Xo = data.get("X"), o.get()[0]
            sg, so = solver.stats(), o.stats()
            scale = max(np.abs(Xo - s.X).max(), 1e-6)
            assert np.abs(Xg - Xo).max() <= 1e-10 * scale, (k, np.abs(Xg - Xo).max(), scale)
            assert sg["newton_iterations"] == so["newton_iterations"]
            assert sg["line_search_halvings"] == so["line_search_halvings"]
            assert np.isclose(sg["alpha0"], so["alpha0"], rtol=1e-12)
            assert Xg[:, 1].min() > 0.0
        assert Xg[:, 1].min() < 4e-3                               # it came down onto the floor
    finally:
        solver.close(), data.free(), fb.close()


def test_pcg_stops_at_the_iteration_limit(gpu):
    """The PCG's stop logic lives on the device and the host reads it only at
    check points (the first after 9/10 of the previous solve's count, then
    every 16 or more): every solve of a step limited to lin_max_iter
    iterations stops at exactly that many, inside a chunk, at its end and
    across a boundary (linear_iterations sums the step's newton_iterations + 1
    solves), and the unlimited solves need more."""
    s = mesh.stack_scene(4, gap=1.5e-3)
    counts = {}
    for cap in (7, 16, 17, 20000):
        data = IpcData(s)
        solver = IpcSolver(data)
        try:
            solver.Update(data, IpcParams(dt=0.01, dhat=1e-3, tol=1e-3, lin_tol=1e-13, lin_max_iter=cap))
            st = solver.stats()
            counts[cap] = (st["linear_iterations"], st["newton_iterations"] + 1)
        finally:
            solver.close()
            data.free()
    lin, solves = counts[20000]
    assert lin > 17 * solves, counts
    for cap in (7, 16, 17):
        assert counts[cap][0] == cap * counts[cap][1], counts
