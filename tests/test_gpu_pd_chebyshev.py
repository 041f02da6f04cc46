"""GPU parity of the reference's default global step (Jacobi + Chebyshev,
pdSolver.cu:193-199, pdUtil.cu:195-226) past the switch to omega != 1, and of
the step's less common inputs: dragged vertices (moreDBC), a vertex in no tet
under the folded PCG + Chebyshev finish, and Reset with a captured step graph.

Bars: FP32 positions within 1e-5 relative of the FP32 oracle, FP64 within
1e-10 of the FP64 oracle (relative L2 and max norm); on the closed-form scene
the error after n iterations follows the Chebyshev polynomial of cheb_case.
"""
import numpy as np
import pytest

import cheb_case as C
import oracle as O
from conftest import max_rel, rel_err
from sbs_amd import mesh
from sbs_amd.pd import PdSolver, SolverData, SolverParams, SolverType

pytestmark = pytest.mark.gpu

TOL = {32: 1e-5, 64: 1e-10}


def jacobi_params(n_iter, gravity=9.8, dt=1.0 / 60.0):
    return SolverParams(numIterations=n_iter, dt=dt, gravity=gravity, tol=1e-6, rho=C.RHO)


def oracle_jacobi(n_iter, gravity=9.8, dt=1.0 / 60.0):
    return O.make_params(solver=O.SOLVER_JACOBI, num_iterations=n_iter, dt=dt, gravity=gravity, rho=C.RHO)


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("n_iter", [12, 20, 100])
def test_closed_form_scene(gpu, precision, n_iter):
    """One step of the linear scene: the GPU's error after n iterations equals
    the closed-form Chebyshev polynomial and the oracle."""
    s = C.scene()
    data = SolverData(s, precision, V0=C.velocity())
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.Jacobi)
    pd.Update(data, jacobi_params(n_iter, gravity=0.0, dt=C.DT))
    pd.synchronize()
    X = data.get("X").astype(np.float64)
    ex = pd.export_system()
    pd.close()
    data.free()
    o = O.PdOracle(s, oracle_jacobi(n_iter, gravity=0.0, dt=C.DT), fp64=(precision == 64), V0=C.velocity())
    o.step(1)
    Xo = o.get("X")
    w = 1e6 * (1.0 / C.DT) ** 2
    for v, ax, g, xstar, s0 in C.modes(np.asarray(ex["c"], np.float64), np.asarray(ex["diagL"], np.float64), w):
        want = C.poly(n_iter, g) * (s0 - xstar)
        tol = 1e-13 if precision == 64 else 1e-6 * (1 + n_iter / 10)
        assert abs(X[v, ax] - xstar - want) <= tol, (v, X[v, ax] - xstar, want)
        # FP32: the two implementations round differently (a few ulps of 0.75
        # per iteration, carried by the Chebyshev extrapolation)
        assert abs(X[v, ax] - Xo[v, ax]) <= (1e-14 if precision == 64 else tol), (v, X[v, ax], Xo[v, ax])


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("n_iter,nsteps", [(20, 50), (100, 20)])
def test_cube_jacobi_chebyshev(gpu, precision, n_iter, nsteps):
    """The 8^3 cube on the reference's default global step with the Chebyshev
    phase active: 20 iterations (context.json-like short runs) and 100 (the
    reference context's numIterations)."""
    s = mesh.cube_scene(8)
    data = SolverData(s, precision)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.Jacobi)
    p = jacobi_params(n_iter)
    for _ in range(nsteps):
        pd.Update(data, p)
    pd.synchronize()
    X = data.get("X")
    pd.close()
    data.free()
    o = O.PdOracle(s, oracle_jacobi(n_iter), fp64=(precision == 64))
    o.step(nsteps)
    Xo = o.get("X")
    assert rel_err(X, Xo) < TOL[precision], rel_err(X, Xo)
    assert max_rel(X, Xo) < TOL[precision], max_rel(X, Xo)
    # the cube moved: the comparison is not of two resting states
    assert np.abs(Xo - s.X).max() > 1e-3


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("solver", [SolverType.Jacobi, SolverType.PCG])
def test_dragged_vertices(gpu, precision, solver):
    """moreDBC (setMDt_2MoreDBC, computeSn, computeDBCLocal, getErrorKern,
    updateVelPos): dragged vertices sit at their targets with zero velocity and
    pull their neighbours, as in the oracle."""
    s = mesh.cube_scene(6)
    rng = np.random.default_rng(11)
    more = np.zeros(s.nv)
    sel = rng.choice(np.flatnonzero(s.dbc == 0), 12, replace=False)
    more[sel] = 5e3
    targets = s.X + np.array([0.05, -0.03, 0.02])
    kw = dict(numIterations=20, dt=1.0 / 60.0, gravity=9.8, tol=1e-6, rho=C.RHO)
    if solver == SolverType.PCG:
        kw.update(pcg_chebyshev=True, pcg_tol=1e-6 if precision == 32 else 1e-13,
                  pcg_floor=1e-8 if precision == 32 else 0.0, pcg_max_iter=400, pcg_adaptive=False)
    p = SolverParams(**kw)
    data = SolverData(s, precision)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(solver)
    pd.Update(data, p)                     # prepares (DBCX <- X0), one undisturbed step
    data.set_more_dbc(more, targets)
    for _ in range(5):
        pd.Update(data, p)
    pd.synchronize()
    X, V = data.get("X"), data.get("V")
    pd.close()
    data.free()
    op = O.make_params(solver=O.SOLVER_JACOBI if solver == SolverType.Jacobi else O.SOLVER_PCG,
                       num_iterations=20, rho=C.RHO, pcg_chebyshev=int(solver == SolverType.PCG),
                       pcg_tol=kw.get("pcg_tol", 1e-6), pcg_floor=kw.get("pcg_floor", 1e-7), pcg_max_iter=400)
    o = O.PdOracle(s, op, fp64=(precision == 64))
    o.step(1)
    o.set_more_dbc(more, targets)
    o.step(5)
    Xo = o.get("X")
    assert rel_err(X, Xo) < TOL[precision], rel_err(X, Xo)
    assert max_rel(X, Xo) < TOL[precision], max_rel(X, Xo)
    if solver == SolverType.Jacobi:
        # getErrorKern keeps a dragged vertex at its prediction: the target
        np.testing.assert_allclose(X[sel], targets[sel], rtol=0, atol=1e-6 if precision == 32 else 1e-12)
    assert np.abs(V[sel]).max() == 0


@pytest.mark.parametrize("precision", [32, 64])
def test_vertex_in_no_tet_with_folded_finish(gpu, precision):
    """A vertex that no tet touches falls freely; the folded PCG + Chebyshev
    finish (second sn/prev pair) keeps it well defined."""
    base = mesh.cube_scene(8)
    X = np.vstack([base.X, [[2.0, 0.4, 0.3]]])
    s = mesh.Scene(X=X, tet=base.tet, tri=base.tri, tri_fathers=base.tri_fathers,
                   mass=np.append(base.mass, base.mass[0]), mu=base.mu, lam=base.lam,
                   dbc=np.append(base.dbc, 0.0), body_vert_start=np.array([0, X.shape[0]]),
                   body_tet_start=base.body_tet_start)
    kw = dict(numIterations=20, dt=1.0 / 60.0, gravity=9.8, tol=1e-6, rho=C.RHO, pcg_chebyshev=True,
              pcg_halo=True, pcg_adaptive=False, pcg_max_iter=400)
    if precision == 64:
        kw.update(pcg_tol=1e-13, pcg_floor=0.0)
    p = SolverParams(**kw)
    data = SolverData(s, precision)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.PCG)
    nsteps = 10
    for _ in range(nsteps):
        pd.Update(data, p)
    pd.synchronize()
    st = pd.stats()
    Xg = data.get("X")
    pd.close()
    data.free()
    assert st.halo_groups > 0 and st.pcg_cap_hits == 0
    assert np.isfinite(Xg).all()
    o = O.PdOracle(s, O.make_params(solver=O.SOLVER_PCG, num_iterations=20, rho=C.RHO, pcg_chebyshev=1,
                                    pcg_tol=p.pcg_tol, pcg_floor=p.pcg_floor, pcg_max_iter=400),
                   fp64=(precision == 64))
    o.step(nsteps)
    Xo = o.get("X")
    assert rel_err(Xg, Xo) < TOL[precision]
    assert max_rel(Xg, Xo) < TOL[precision]
    # implicit-Euler free fall of the lone vertex: y drops by g h^2 n(n+1)/2
    drop = 9.8 * p.dt ** 2 * nsteps * (nsteps + 1) / 2
    assert abs(Xg[-1, 1] - (0.4 - drop)) < (1e-6 if precision == 32 else 1e-12)


def test_reset_drops_the_captured_graph(gpu):
    """Reset with use_graph: the next update prepares again (new buffers) and
    must capture a new graph instead of replaying one that refers to freed
    memory; positions equal an eager run that is reset the same way."""
    from sbs_amd.capi import Stream
    s = mesh.cube_scene(8)
    runs = []
    for use_graph in (True, False):
        data = SolverData(s, 32)
        pd = PdSolver(128, data)
        pd.SetGlobalSolver(SolverType.Jacobi)
        st = Stream()
        pd.set_stream(st)
        runs.append((data, pd, st, SolverParams(numIterations=10, dt=1.0 / 60.0, gravity=9.8, rho=C.RHO,
                                                use_graph=use_graph)))
    try:
        def advance(n):
            for _ in range(n):
                for data, pd, _, pp in runs:
                    pd.Update(data, pp)
                    pd.synchronize()
                np.testing.assert_array_equal(runs[0][0].get("X"), runs[1][0].get("X"))
        advance(4)
        before = runs[0][1].stats().graph_captures
        for _, pd, _, _ in runs:
            pd.Reset()
        advance(4)
        after = runs[0][1].stats()
        assert after.graph_captures == before + 1, (before, after.graph_captures)
        assert after.graph_replays == 8
    finally:
        for data, pd, st, _ in runs:
            pd.close()
            data.free()
            st.close()
