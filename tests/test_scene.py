"""Scene-file parsing (sbs_amd/scene.py) on the CPU: the reference's context.json
structure (definitions by name, per-instance overrides, defaults of
getJsonVec3 / value(), "load": false contexts skipped, precision selection),
loaded through the library's TetGen loader from tests/golden/tetgen/."""
import json
import os

import numpy as np
import pytest

from sbs_amd import fixed as F
from sbs_amd.scene import SceneFile

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tetgen")

SCENE = {
    "num of iterations": 12,
    "threads per block": 64,
    "softBodies": [
        {"name": "house1", "nodeFile": "../assets/house2/house2.node", "eleFile": "../assets/house2/house2.ele",
         "centralize": True, "start index": 1, "pos": [0.0, 40.0, 0.0], "scale": [4, 4, 4], "mass": 10.0,
         "mu": 200000.0, "lambda": 5000.0, "DBC": [0, 1]},
        {"name": "cube", "nodeFile": "../assets/cube/cube.1.node", "eleFile": "../assets/cube/cube.1.ele",
         "scale": [5, 5, 5], "start index": 1},
    ],
    "fixedBodies": [
        {"type": "sphere", "name": "sphere1", "pos": [0.0, -50.0, 0.0], "radius": 10.0},
        {"type": "cylinder", "name": "cylinder1", "pos": [0.0, 10.0, 0.0], "rot": [90.0, 0.0, 0.0],
         "scale": [10.0, 75.0, 10.0]},
        {"type": "plane", "name": "bottom plane", "scale": [450.0, 450.0, 450.0]},
        {"type": "plane", "name": "right plane", "pos": [450.0, 0.0, 0.0], "rot": [0.0, 0.0, 90.0]},
    ],
    "contexts": [
        {"name": "skipped", "load": False, "softBodies": [{"name": "cube"}]},
        {"name": "pd", "precision": "float", "dt": 0.01, "gravity": 98, "tolerance": 1e-9,
         "softBodies": [{"name": "house1", "DBC": [5]}, {"name": "cube", "pos": [60, 300, 0], "mass": 2.0}],
         "fixedBodies": [{"name": "cylinder1", "pos": [-25.0, 80.0, 0.0]}, {"name": "bottom plane"},
                         {"name": "right plane"}, {"name": "sphere1", "radius": 3.0}]},
        {"name": "ipc", "precision": "double", "dt": 0.05, "softBodies": [{"name": "cube"}]},
    ],
}


@pytest.fixture
def scene_file(tmp_path):
    p = tmp_path / "context.json"
    p.write_text(json.dumps(SCENE))
    return SceneFile(str(p), search_dirs=(G,))


def test_contexts_and_defaults(scene_file):
    sf = scene_file
    assert [c["name"] for c in sf.contexts] == ["pd", "ipc"]        # "load": false skipped
    assert sf.num_iterations == 12 and sf.threads_per_block == 64
    s, dl = sf.soft_scene(sf.context("pd"), 32)
    assert s.nv == 400 + 8 and s.nt == 1389 + 6 and dl.names == ["house1", "cube"]
    assert (s.mass[:400] == 10.0).all() and (s.mass[400:] == 2.0).all()     # instance override
    assert (s.mu[:1389] == 2e5).all() and (s.mu[1389:] == 1000.0).all()      # default mu
    assert (s.lam[1389:] == 1000.0).all()
    np.testing.assert_array_equal(np.nonzero(s.dbc)[0], [5])                 # the instance's DBC wins
    cube = s.X[400:]
    np.testing.assert_allclose(cube.mean(axis=0), [60, 300, 0], atol=1e-4)   # cube.1 is centred at 0, scaled 5
    np.testing.assert_allclose(np.ptp(cube, axis=0), [10, 10, 10], atol=1e-4)
    house = s.X[:400]
    np.testing.assert_allclose(house.mean(axis=0), [0, 40, 0], atol=1e-3)    # centralize, then pos


def test_fixed_bodies_from_definitions(scene_file):
    sf = scene_file
    bodies = sf.fixed_bodies(sf.context("pd"))
    assert [b.type for b in bodies] == [F.CYLINDER, F.PLANE, F.PLANE, F.SPHERE]
    cyl, bottom, right, sph = bodies
    np.testing.assert_allclose(cyl.model[:3, 3], [-25, 80, 0])               # instance pos
    assert cyl.radius == 10.0                                                 # scale[0] of the definition
    np.testing.assert_allclose(cyl.model[:3, 1] / 75.0, [0, 0, 1], atol=1e-6)  # rot 90 about x: y axis -> z
    np.testing.assert_allclose(bottom.model[:3, :3], 450 * np.eye(3))
    np.testing.assert_allclose(right.model[:3, 3], [450, 0, 0])
    assert sph.radius == 3.0 and np.allclose(sph.model[:3, 3], [0, -50, 0])  # radius override, def pos


def test_missing_file_is_reported(tmp_path):
    bad = dict(SCENE)
    bad["softBodies"] = [{"name": "cube", "nodeFile": "nowhere/x.node", "eleFile": "nowhere/x.ele"}]
    bad["contexts"] = [{"name": "pd", "precision": "float", "softBodies": [{"name": "cube"}]}]
    p = tmp_path / "bad.json"
    p.write_text(json.dumps(bad))
    sf = SceneFile(str(p))
    with pytest.raises(FileNotFoundError):
        sf.soft_scene(sf.context("pd"), 32)
