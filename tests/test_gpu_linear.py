"""GPU parity of the general linear solve, LinearSolver<T>::Solve with the
PCGJacobiSolver algorithm (linear.h:62, pcgJacobi.cu:89-173), through the C
ABI against the oracle (oracle/linear_oracle.py) on the same triplets.

Bars: FP64 solutions within 1e-12 of the oracle's (relative, max norm) with
the same iteration count (+-1 where the absolute stop test sits at a
rounding-level boundary); FP32 within the accuracy its stopping rule gives
(||r|| < tol) of the exact solution; the summed entry count equal to
scipy's.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import linear_oracle as LO
from sbs_amd import capi, mesh
from sbs_amd.linear import PCGJacobiSolver
from sbs_amd.pd import PdSolver, SolverData, SolverParams, SolverType
from test_oracle_linear import random_spd_coo

pytestmark = pytest.mark.gpu


def upload(*arrays):
    return [capi.DeviceArray.from_numpy(np.ascontiguousarray(a)) for a in arrays]


def solve_gpu(n, b, v, r, c, precision, tol, max_iter=2000, guess=None):
    dt = np.float64 if precision == 64 else np.float32
    db, dA, dr, dc = upload(b.astype(dt), v.astype(dt), r.astype(np.int32), c.astype(np.int32))
    dx = capi.DeviceArray(n * np.dtype(dt).itemsize)
    dg = upload(guess.astype(dt))[0] if guess is not None else None
    s = PCGJacobiSolver(n, max_iter=max_iter, tolerance=tol, precision=precision)
    try:
        k = s.Solve(n, db, dx, dA, v.size, dr, dc, dg)
        return dx.download(dt, (n,)).astype(np.float64), k, s.nnz()
    finally:
        s.close()
        for a in (db, dA, dr, dc, dx) + ((dg,) if dg is not None else ()):
            a.free()


@pytest.mark.parametrize("n,seed", [(50, 1), (1000, 2), (20000, 3)])
def test_fp64_matches_oracle(gpu, n, seed):
    r, c, v = random_spd_coo(n, seed)
    b = np.random.default_rng(seed + 10).normal(size=n)
    x, k, nnz = solve_gpu(n, b, v, r, c, 64, tol=1e-8)
    xo, ko = LO.pcg_jacobi(n, b, v, r, c, tol=1e-8)
    assert abs(k - ko) <= 1, (k, ko)
    if k == ko:
        assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    assert nnz == A.nnz
    assert np.linalg.norm(A @ x - b) < 1e-7


def test_fp64_with_guess_and_limits(gpu):
    n = 3000
    r, c, v = random_spd_coo(n, 4)
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    xs = np.random.default_rng(5).normal(size=n)
    b = A @ xs
    guess = xs + np.random.default_rng(6).normal(scale=1e-3, size=n)
    x, k, _ = solve_gpu(n, b, v, r, c, 64, tol=1e-9, guess=guess)
    xo, ko = LO.pcg_jacobi(n, b, v, r, c, tol=1e-9, guess=guess)
    assert abs(k - ko) <= 1 and np.abs(x - xs).max() < 1e-8
    x, k, _ = solve_gpu(n, b, v, r, c, 64, tol=1e-9, guess=xs)       # exact guess: no iteration
    assert k == 0 and np.array_equal(x, xs)
    _, k, _ = solve_gpu(n, b, v, r, c, 64, tol=0.0, max_iter=5)
    assert k == 5


def test_fp32_solution(gpu):
    n = 5000
    r, c, v = random_spd_coo(n, 7)
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    b = np.random.default_rng(8).normal(size=n)
    x, k, _ = solve_gpu(n, b, v, r, c, 32, tol=1e-3)
    xe = spla.spsolve(A.tocsc(), b)
    assert 0 < k < 2000
    # the added diagonal is at least 0.5, so ||A^-1|| <= 2: ||r|| < 1e-3 bounds
    # the error by 2e-3, plus FP32 rounding of the recursive residual
    assert np.abs(x - xe).max() < 2e-3 + 1e-4


def test_pd_system_as_ipc_calls_it(gpu):
    """The prepared PD matrix of an 8^3 cube expanded to 3N x 3N (A_s (x) I3, the
    reference's layout), its triplets shuffled and split: the solve equals the
    PD solver's own answer to the tolerance."""
    s = mesh.cube_scene(8)
    data = SolverData(s, 64)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.CuSolverCholesky)
    pd.Update(data, SolverParams(numIterations=1, dt=1.0 / 60.0))
    ex = pd.export_system()
    pd.close()
    data.free()
    rows = np.repeat(np.arange(s.nv), np.diff(ex["rowptr"]))
    cols, vals = np.asarray(ex["col"]), np.asarray(ex["val"])
    R = np.concatenate([3 * rows + k for k in range(3)])
    C = np.concatenate([3 * cols + k for k in range(3)])
    V = np.concatenate([vals] * 3)
    rng = np.random.default_rng(9)
    f = rng.uniform(0.3, 0.7, V.size)
    R, C, V = np.concatenate([R, R]), np.concatenate([C, C]), np.concatenate([V * f, V * (1 - f)])
    perm = rng.permutation(V.size)
    R, C, V = R[perm], C[perm], V[perm]
    n = 3 * s.nv
    xs = s.X.ravel().copy()
    A = sp.coo_matrix((V, (R, C)), shape=(n, n)).tocsr()
    b = A @ xs
    x, k, nnz = solve_gpu(n, b, V, R, C, 64, tol=1e-3)
    xo, ko = LO.pcg_jacobi(n, b, V, R, C, tol=1e-3)
    assert nnz == 3 * len(vals)
    assert abs(k - ko) <= 1, (k, ko)
    if k == ko:
        assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()
    # the pinned rows (weight 1e6/h^2) dominate b: the absolute stop test
    # leaves the free rows' error at the scale of ||A^-1|| 1e-3
    assert np.abs(x - xs).max() < 1e-3


def test_errors_are_reported(gpu):
    n = 4
    db, dA, dr, dc = upload(np.ones(n), np.ones(2), np.array([0, 5], np.int32), np.array([0, 1], np.int32))
    dx = capi.DeviceArray(n * 8)
    s = PCGJacobiSolver(n, precision=64)
    try:
        with pytest.raises(capi.SbsError):
            s.Solve(n, db, dx, dA, 2, dr, dc)
        assert s.Solve(0, db, dx) == 0
    finally:
        s.close()
        for a in (db, dA, dr, dc, dx):
            a.free()
