"""The allocation-free linear solve (sbs_linear_solve_noalloc) through the C
ABI against the allocating one (sbs_linear_solve) on the same triplets: two
calls in a row, the second smaller (the shared workspace keeps the first
call's contents), both bit-identical to the allocating solve (same kernels,
fixed-order reductions); a system larger than the workspace is rejected."""
import numpy as np
import pytest

from sbs_amd import capi
from sbs_amd.linear import PCGJacobiSolver
from test_oracle_linear import random_spd_coo

pytestmark = pytest.mark.gpu


def run(solver, n, r, c, v, b, dt, no_alloc, guess=None):
    arrs = [capi.DeviceArray.from_numpy(np.ascontiguousarray(a)) for a in
            (b.astype(dt), v.astype(dt), r.astype(np.int32), c.astype(np.int32))]
    db, dA, dr, dc = arrs
    dg = capi.DeviceArray.from_numpy(guess.astype(dt)) if guess is not None else None
    dx = capi.DeviceArray(n * np.dtype(dt).itemsize)
    try:
        f = solver.SolveNoAlloc if no_alloc else solver.Solve
        k = f(n, db, dx, dA, v.size, dr, dc, dg)
        return dx.download(dt, (n,)), k
    finally:
        for a in arrs + [dx] + ([dg] if dg is not None else []):
            a.free()


@pytest.mark.parametrize("precision", [64, 32])
def test_no_alloc_matches_allocating_solve(gpu, precision):
    dt = np.float64 if precision == 64 else np.float32
    s = PCGJacobiSolver(max_iter=2000, tolerance=1e-8 if precision == 64 else 1e-3, precision=precision)
    try:
        for n, seed, guess in [(8000, 1, False), (700, 2, True)]:
            r, c, v = random_spd_coo(n, seed)
            b = np.random.default_rng(seed + 5).normal(size=n)
            g = np.random.default_rng(seed + 6).normal(scale=0.1, size=n) if guess else None
            xa, ka = run(s, n, r, c, v, b, dt, False, g)
            xn, kn = run(s, n, r, c, v, b, dt, True, g)
            assert kn == ka and ka > 0
            np.testing.assert_array_equal(xn, xa)
    finally:
        s.close()


def test_no_alloc_rejects_a_system_beyond_its_workspace(gpu):
    s = PCGJacobiSolver(precision=64)
    n = 2_000_000
    db = capi.DeviceArray(n * 8)
    dx = capi.DeviceArray(n * 8)
    try:
        with pytest.raises(capi.SbsError, match="workspace"):
            s.SolveNoAlloc(n, db, dx, None, 0, None, None)
    finally:
        db.free()
        dx.free()
        s.close()
