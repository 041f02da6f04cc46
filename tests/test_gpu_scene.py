"""Scenes from a scene file (sbs_amd/scene.py) on the GPU: the reference's own
assets (tests/golden/tetgen/) loaded through the library's TetGen loader and
stepped by the library's PD and IPC solvers.

Bars: a float context without fixed bodies or collision handling matches the
FP32 oracle's Jacobi steps on the same flat scene to 1e-5 (relative L2 and
max norm); with a floor plane every vertex of the projected end-of-step
positions stays on or above the floor (to float rounding) and the
"collision handling(fixed)" timer runs; a double context takes IPC steps
that converge.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import max_rel, rel_err
from sbs_amd.pd import SolverType
from sbs_amd.scene import SceneFile, SimulationContext

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tetgen")

SCENE = {
    "num of iterations": 10,
    "softBodies": [
        {"name": "house1", "nodeFile": "house2.node", "eleFile": "house2.ele", "centralize": True,
         "start index": 1, "scale": [1, 1, 1], "mass": 1.0, "mu": 20000.0, "lambda": 5000.0, "DBC": [0, 7, 30]},
        {"name": "cube", "nodeFile": "cube.1.node", "eleFile": "cube.1.ele", "faceFile": "cube.1.face",
         "start index": 1, "scale": [0.5, 0.5, 0.5], "mass": 0.5, "mu": 20000.0, "lambda": 5000.0},
    ],
    "fixedBodies": [{"type": "plane", "name": "floor", "scale": [10, 10, 10]}],
    "contexts": [
        {"name": "bodies", "precision": "float", "dt": 0.01, "gravity": 9.8,
         "softBodies": [{"name": "house1", "pos": [0, 20, 0]}, {"name": "cube", "pos": [20, 5, 0]}]},
        {"name": "floor", "precision": "float", "dt": 0.01, "gravity": 98,
         "softBodies": [{"name": "cube", "pos": [0, 0.6, 0]}], "fixedBodies": [{"name": "floor"}]},
        {"name": "ipc", "precision": "double", "dt": 0.01, "dhat": 1e-3, "tolerance": 1e-2,
         "softBodies": [{"name": "cube", "pos": [0, 0, 0]}, {"name": "cube", "pos": [0, 1.02, 0]}]},
    ],
}


@pytest.fixture(scope="module")
def scene_file(tmp_path_factory):
    p = tmp_path_factory.mktemp("scene") / "scene.json"
    p.write_text(json.dumps(SCENE))
    return SceneFile(str(p), search_dirs=(G,))


def test_float_context_matches_oracle(gpu, scene_file):
    sim = SimulationContext(scene_file, "bodies", SolverType.Jacobi, handle_collision=False)
    try:
        for _ in range(3):
            sim.Update()
        X = sim.positions()
        scene, p = sim.scene, sim.params
    finally:
        sim.close()
    o = O.PdOracle(scene, O.make_params(solver=O.SOLVER_JACOBI, num_iterations=p.numIterations, dt=p.dt,
                                        gravity=p.gravity, tol=p.tol, rho=p.rho), fp64=False)
    o.step(3)
    Xo = o.get("X")
    assert np.abs(Xo - scene.X).max() > 1e-4                     # the bodies moved
    assert rel_err(X, Xo) <= 1e-5 and max_rel(X, Xo) <= 1e-5


def test_floor_context_keeps_the_cube_above_the_plane(gpu, scene_file):
    sim = SimulationContext(scene_file, "floor", SolverType.Jacobi)
    try:
        lowest = []
        for _ in range(15):
            sim.Update()
            lowest.append(sim.data.get("XTilde")[:, 1].min())
        perf = dict(sim.GetPerformanceData())
    finally:
        sim.close()
    assert min(lowest) >= -1e-6                                   # projected onto the floor, never below
    assert lowest[-1] < 0.05                                      # the cube reached the floor
    assert perf["collision handling(fixed)"] > 0.0 and perf["local step"] > 0.0


def test_double_context_runs_ipc(gpu, scene_file):
    sim = SimulationContext(scene_file, "ipc")
    try:
        for _ in range(2):
            assert sim.Update()
        X = sim.positions()
    finally:
        sim.close()
    assert np.isfinite(X).all() and X.shape == (16, 3)
