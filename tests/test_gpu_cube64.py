"""The 64^3-hex cube (274,625 vertices, 1,572,864 tets) on its measured step:
20 local/global iterations, Jacobi-PCG to 1e-6 (floor 1e-8 in FP32) +
Chebyshev, h = 1/60, top face pinned.

* parity: the FP32 GPU run against the FP32 oracle (the reference's step in
  the reference's precision) over 10 steps, within 1e-5 relative (L2 and max
  norm);
* accuracy: over 50 steps the FP32 run's distance to a tightly solved FP64
  run stays what FP32 arithmetic gives this algorithm -- within 1.5 times the
  FP32 oracle's own distance to that FP64 run after 10 steps, and about 1.4e-4
  of the displacement at every step count (the relative position error grows
  with the displacement, 6e-5 of |X| after 50 steps; the PCG tolerance is not
  the cause: an FP64 run with the FP32 tolerances is within 1.3e-6).

Both FP32 rotations of the local step are held to these bars: the TR1690 SVD
(as the reference) and Newton's polar iteration (the measured step), which
agree to rounding on this workload's elements (singular values within
[0.95, 1.92] on a 16^3 cube of the same setup; the SVD's sweeps have
converged there).
"""
import numpy as np
import pytest

import oracle as O
from conftest import max_rel, rel_err
from sbs_amd import mesh
from sbs_amd.pd import PdSolver, SolverData, SolverParams, SolverType

pytestmark = pytest.mark.gpu


def params(fp64, tight=False, polar=False):
    return SolverParams(numIterations=20, dt=1.0 / 60.0, gravity=9.8, tol=1e-6, rho=0.9992,
                        pcg_tol=1e-12 if tight else 1e-6, pcg_floor=0.0 if (fp64 or tight) else 1e-8,
                        pcg_max_iter=400, pcg_chebyshev=True, pcg_adaptive=True, polar_rotation=polar)


def run(scene, precision, p, marks):
    data = SolverData(scene, precision)
    pd = PdSolver(128, data)
    pd.SetGlobalSolver(SolverType.PCG)
    out = {}
    try:
        for k in range(1, max(marks) + 1):
            pd.Update(data, p)
            if k in marks:
                pd.synchronize()
                out[k] = data.get("X").astype(np.float64)
        assert pd.stats().pcg_cap_hits == 0
    finally:
        pd.close()
        data.free()
    return out


@pytest.fixture(scope="module")
def cube():
    return mesh.cube_scene(64)


@pytest.mark.parametrize("polar", [False, True])
def test_fp32_matches_fp32_oracle_10_steps(gpu, cube, polar):
    g = run(cube, 32, params(False, polar=polar), [10])[10]
    o = O.PdOracle(cube, O.make_params(solver=O.SOLVER_PCG, num_iterations=20, rho=0.9992, pcg_tol=1e-6,
                                       pcg_floor=1e-8, pcg_max_iter=400, pcg_chebyshev=1), fp64=False)
    o.step(10)
    r = o.get("X")
    o.close()
    assert rel_err(g, r) < 1e-5, rel_err(g, r)
    assert max_rel(g, r) < 1e-5, max_rel(g, r)


@pytest.mark.parametrize("polar", [False, True])
def test_fp32_accuracy_over_50_steps(gpu, cube, polar):
    marks = [10, 25, 50]
    a = run(cube, 32, params(False, polar=polar), marks)
    c = run(cube, 64, params(True, tight=True), marks)
    o = O.PdOracle(cube, O.make_params(solver=O.SOLVER_PCG, num_iterations=20, rho=0.9992, pcg_tol=1e-6,
                                       pcg_floor=1e-8, pcg_max_iter=400, pcg_chebyshev=1), fp64=False)
    o.step(10)
    own = rel_err(o.get("X"), c[10])
    o.close()
    assert rel_err(a[10], c[10]) <= 1.5 * own, (rel_err(a[10], c[10]), own)
    for k in marks:
        disp = np.linalg.norm(c[k] - cube.X)
        assert np.linalg.norm(a[k] - c[k]) <= 2e-4 * disp, (k, np.linalg.norm(a[k] - c[k]) / disp)
    assert rel_err(a[50], c[50]) < 1e-4
