"""GPU parity of the dense small-body solver (SolverType.DenseCholesky):
per-body inverses applied with MFMA (FP32 or BF16 fragments, FP32
accumulation) and FP64 iterative refinement, against the oracle's direct
(Cholesky) mode.

Bars: FP32 positions within 1e-5 relative of the FP32 oracle; FP64 within
1e-10 of the FP64 oracle; every solve reaches the refinement tolerance.
"""
import numpy as np
import pytest

import oracle as O
from conftest import max_rel, rel_err
from sbs_amd import capi, mesh
from sbs_amd.pd import PdSolver, SolverData, SolverParams, SolverType

pytestmark = pytest.mark.gpu

TOL = {32: 1e-5, 64: 1e-10}


def params(n_iter=10, tol=1e-6, dense_precision=0, dense_tol=1e-10, dense_max_iter=8):
    return SolverParams(numIterations=n_iter, dt=1.0 / 60.0, gravity=9.8, tol=tol, rho=0.9992,
                        dense_precision=dense_precision, dense_tol=dense_tol, dense_max_iter=dense_max_iter)


def run_dense(scene, p, nsteps, precision):
    data = SolverData(scene, precision)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.DenseCholesky)
    worst = 0
    for _ in range(nsteps):
        pd.Update(data, p)
        info = pd.dense_info()
        assert info["iterations"] >= 0, info      # converged within dense_max_iter
        worst = max(worst, info["iterations"])
    pd.synchronize()
    X = data.get("X")
    pd.close()
    data.free()
    return X, info, worst


def run_oracle(scene, p, nsteps, fp64):
    o = O.PdOracle(scene, O.make_params(solver=O.SOLVER_CHOLESKY, num_iterations=p.numIterations, dt=p.dt,
                                        gravity=p.gravity, tol=p.tol, rho=p.rho), fp64=fp64)
    o.step(nsteps)
    return o.get("X")


@pytest.mark.parametrize("dense_precision", [0, 1])
def test_dense_cube_100_steps_fp32(gpu, dense_precision):
    s = mesh.cube_scene(8)
    p = params(dense_precision=dense_precision, dense_max_iter=16)
    X, info, worst = run_dense(s, p, 100, 32)
    r = run_oracle(s, p, 100, False)
    assert rel_err(X, r) < TOL[32] and max_rel(X, r) < TOL[32]
    assert info["bodies"] == 1 and info["max_body"] == s.nv


def test_dense_cube_100_steps_fp64(gpu):
    s = mesh.cube_scene(8)
    p = params(tol=0.0, dense_tol=1e-14, dense_max_iter=12)
    X, info, worst = run_dense(s, p, 100, 64)
    r = run_oracle(s, p, 100, True)
    assert rel_err(X, r) < TOL[64] and max_rel(X, r) < TOL[64]


@pytest.mark.parametrize("dense_precision", [0, 1])
def test_dense_many_small_bodies(gpu, dense_precision):
    """Eight 4^3 bodies of different stiffness, the mixed-precision workload's shape."""
    rng = np.random.default_rng(5)
    bodies = [dict(nx=4, ny=4, nz=4, origin=(0.6 * i, 0.0, 0.0), seed=10 + i, pin_top=True,
                   E=float(rng.uniform(5e4, 5e5))) for i in range(8)]
    s = mesh.make_scene(bodies)
    p = params(n_iter=10, dense_precision=dense_precision, dense_max_iter=16)
    X, info, worst = run_dense(s, p, 20, 32)
    r = run_oracle(s, p, 20, False)
    assert rel_err(X, r) < TOL[32] and max_rel(X, r) < TOL[32]
    assert info["bodies"] == 8 and info["max_body"] == 125
    assert info["ratio"] <= 1e-10


@pytest.mark.parametrize("precision", [32, 64])
def test_dense_linear_solve(gpu, precision):
    s = mesh.make_scene([dict(nx=3, ny=3, nz=3, seed=1, pin_top=True),
                         dict(nx=4, ny=3, nz=2, origin=(1.0, 0, 0), seed=2, pin_top=True)])
    dt = np.float32 if precision == 32 else np.float64
    data = SolverData(s, precision)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.DenseCholesky)
    p = params(dense_tol=1e-12, dense_max_iter=12)
    pd.Update(data, p)
    ex = pd.export_system()
    A = np.zeros((s.nv, s.nv))
    for row in range(s.nv):
        sl = slice(ex["rowptr"][row], ex["rowptr"][row + 1])
        A[row, ex["col"][sl]] = ex["val"][sl]
    xs = np.random.default_rng(3).normal(size=(s.nv, 3))
    b = (A @ xs).astype(dt)
    db = capi.DeviceArray.from_numpy(b.ravel())
    dx = capi.DeviceArray(b.size * np.dtype(dt).itemsize)
    pd.solve(db.ptr, dx.ptr, p)
    x = dx.download(dt, b.shape)
    # the refinement solves the system of the given (rounded) b to 1e-12
    xe = np.linalg.solve(A, b.astype(np.float64))
    assert rel_err(x, xe) < (1e-6 if precision == 32 else 1e-10)
    pd.close()
    data.free()


def test_dense_256_bodies_of_6cube_vs_fp64_oracle(gpu):
    """The dense small-body workload at its size: 256 bodies of 6^3 hexes
    (87,808 vertices, 331,776 tets, 343 vertices per body, coordinates up to
    ~16 m), 10 steps of 10 iterations, no early exit, against the FP64 oracle.

    The error is FP32's, not the MFMA precision's: positions at 16 m round at
    ~1e-6 m, and the FP32 oracle (the reference's arithmetic) is itself
    4.5e-6 (l2) / 2.7e-4 m (max) from FP64 here.  Bar: both MFMA modes within
    1.5x the FP32 oracle's distance from FP64, and within 1e-5 (l2) / 5e-4 m
    (max) absolutely; every solve refined to 1e-10."""
    s = mesh.small_bodies_scene(256, 6)
    assert (s.nv, s.nt) == (87808, 331776)
    o64 = O.PdOracle(s, O.make_params(solver=O.SOLVER_CHOLESKY, num_iterations=10, dt=1.0 / 60.0, gravity=9.8,
                                      tol=0.0, rho=0.9992), fp64=True)
    o64.step(10)
    r64 = o64.get("X")
    o64.close()
    r32 = run_oracle(s, params(tol=0.0), 10, False)
    e_orc = rel_err(r32, r64)
    m_orc = np.abs(r32 - r64).max()
    for prec in (0, 1):
        X, info, worst = run_dense(s, params(tol=0.0, dense_precision=prec, dense_tol=1e-10, dense_max_iter=16),
                                   10, 32)
        assert info["bodies"] == 256 and info["max_body"] == 343
        assert info["ratio"] <= 1e-10
        e, m = rel_err(X, r64), np.abs(X - r64).max()
        assert e < min(1e-5, 1.5 * e_orc), (prec, e, e_orc)
        assert m < min(5e-4, 1.5 * m_orc), (prec, m, m_orc)


@pytest.mark.parametrize("dense_precision", [0, 1])
def test_dense_body_of_1000_vertices(gpu, dense_precision):
    """A 9^3-hex body (1000 vertices, padded to 1024 rows): with FP32
    fragments its x and residual fragments exceed 64 KB of LDS, so each
    correction and residual runs as a launch over all row blocks; with BF16
    fragments the body is solved in one workgroup.  Both against the oracle's
    direct mode over 20 steps."""
    s = mesh.make_scene([dict(nx=9, ny=9, nz=9, seed=4, pin_top=True, E=2e5)])
    assert s.nv == 1000
    p = params(n_iter=10, dense_precision=dense_precision, dense_max_iter=16)
    X, info, worst = run_dense(s, p, 20, 32)
    r = run_oracle(s, p, 20, False)
    assert rel_err(X, r) < TOL[32] and max_rel(X, r) < TOL[32]
    assert info["bodies"] == 1 and info["max_body"] == 1000
    assert info["ratio"] <= 1e-10
