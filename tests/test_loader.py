"""TetGen loading (sbs_mesh_load_tetgen, host code of the library) against the
loader oracle (oracle/tetgen_oracle.py), on the CPU.

Fixtures: tests/golden/tetgen/ holds three of the reference's own assets
(assets/cube/cube.1.*, assets/tet/tet.*, assets/house2/house2.*), loaded with
the options the reference's context.json gives them.

Pins beyond the oracle: the boundary extracted from the cube's tets is the
face set TetGen itself wrote (cube.1.face); extracted surfaces are closed
2-manifolds with every directed edge used once, and the volume they enclose
(divergence theorem) equals the sum of the tet volumes; a synthetic hex-grid
body written as TetGen files comes back with its own vertices, tets and the
surface the scene builder derives independently; the model transform in
closed form.
"""
import os

import numpy as np
import pytest

import tetgen_oracle as TO
from sbs_amd import capi, mesh
from sbs_amd.loader import DataLoader, SoftBodyAttribute, load_tetgen

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tetgen")
CASES = {  # name: (node, ele, face, start index, centralize, pos, rot, scale) as context.json gives them
    "cube": ("cube.1.node", "cube.1.ele", None, 1, False, (0, 10, 25), (0, 0, 0), (5, 5, 5)),
    "cube_faces": ("cube.1.node", "cube.1.ele", "cube.1.face", 1, False, (0, 10, 25), (0, 0, 0), (5, 5, 5)),
    "tet": ("tet.node", "tet.ele", None, 1, False, (0, 10, 25), (0, 0, 0), (4, 4, 4)),
    "house1": ("house2.node", "house2.ele", None, 1, True, (0, 40, 0), (0, 0, 0), (4, 4, 4)),
    "house_rotated": ("house2.node", "house2.ele", None, 1, True, (0, 180, 0), (0, 0, 75), (4, 4, 4)),
}


def both(name, precision=32):
    node, ele, face, start, cen, pos, rot, scale = CASES[name]
    args = (os.path.join(G, node), os.path.join(G, ele), os.path.join(G, face) if face else None)
    kw = dict(pos=pos, scale=scale, rot=rot, centralize=cen, start_index=start, precision=precision)
    return load_tetgen(*args, **kw), TO.load(*args, **kw)


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("precision", [32, 64])
def test_loader_matches_oracle(name, precision):
    got, ref = both(name, precision)
    for k in ("tet", "tri", "edges"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    assert got["tris_from_file"] == ref["tris_from_file"]
    if CASES[name][6] == (0, 0, 0):
        np.testing.assert_array_equal(got["X"], ref["X"])          # same IEEE sequence
    else:                                                            # cos/sin of two libraries
        np.testing.assert_allclose(got["X"], ref["X"], rtol=0, atol=1e-5 * np.abs(ref["X"]).max())


def test_extracted_cube_surface_is_tetgens_own():
    got, _ = both("cube")
    faces, _ = both("cube_faces")
    assert faces["tris_from_file"] and not got["tris_from_file"]
    key = lambda t: sorted(map(tuple, np.sort(t, axis=1).tolist()))
    assert key(got["tri"]) == key(faces["tri"])
    assert len(faces["edges"]) == 18                      # 12 cube edges + 6 face diagonals


def signed_volume(X, tri):
    a, b, c = X[tri[:, 0]], X[tri[:, 1]], X[tri[:, 2]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0


@pytest.mark.parametrize("name", ["house1", "cube", "tet"])
def test_extracted_surface_is_closed_and_encloses_the_tets(name):
    got, _ = both(name, 64)
    X, tri, tet = got["X"], got["tri"].astype(np.int64), got["tet"].astype(np.int64)
    directed = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    assert len({tuple(e) for e in directed.tolist()}) == len(directed)          # each directed edge once
    und = {tuple(sorted(e)) for e in directed.tolist()}
    assert len(und) * 2 == len(directed) == 2 * len(got["edges"])                # each edge in two faces
    d = X[tet[:, 1:]] - X[tet[:, :1]]
    vol = np.abs(np.linalg.det(d)).sum() / 6.0
    assert abs(abs(signed_volume(X, tri)) - vol) <= 1e-9 * vol


def test_centring_and_transform_closed_form():
    node = os.path.join(G, "house2.node")
    raw = np.array([[float(v) for v in l.split()[1:4]] for l in open(node).read().splitlines()[1:401]])
    raw = raw.astype(np.float32).astype(np.float64)
    m = load_tetgen(node, os.path.join(G, "house2.ele"), None, pos=(1, 2, 3), scale=(2, 2, 2),
                    centralize=True, start_index=1, precision=64)
    c = raw - raw.mean(axis=0)
    expect = np.stack([c[:, 0], c[:, 2], c[:, 1]], axis=1) * 2 + [1, 2, 3]
    np.testing.assert_allclose(m["X"], expect, atol=1e-5)
    # Rz(90) takes +x to +y
    t = load_tetgen(os.path.join(G, "tet.node"), os.path.join(G, "tet.ele"), None, rot=(0, 0, 90),
                    start_index=1, precision=64)
    raw4 = np.array([[1, 1, -1], [-1, -1, -1], [1, -1, -1], [1, -1, 1.0]])
    np.testing.assert_allclose(t["X"], np.stack([-raw4[:, 1], raw4[:, 0], raw4[:, 2]], 1), atol=1e-6)


def test_synthetic_body_round_trip(tmp_path):
    s = mesh.cube_scene(3, jitter=1e-3, seed=4)
    node, ele = tmp_path / "b.node", tmp_path / "b.ele"
    with open(node, "w") as f:
        f.write(f"# written by the test\n{s.nv} 3 0 0\n")
        for i, p in enumerate(s.X):
            f.write(f"{i + 1} {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
    with open(ele, "w") as f:
        f.write(f"{s.nt} 4 0\n\n")
        for i, t in enumerate(s.tet):
            f.write(f"{i + 1} " + " ".join(str(int(v) + 1) for v in t) + "  # a comment\n")
    m = load_tetgen(str(node), str(ele), str(tmp_path / "missing.face"), start_index=1, precision=64)
    np.testing.assert_allclose(m["X"], s.X, atol=1e-7)
    np.testing.assert_array_equal(m["tet"], s.tet)
    np.testing.assert_array_equal(m["tri"], s.tri)      # the scene builder's own face-parity surface
    assert not m["tris_from_file"]


def test_errors_are_reported():
    with pytest.raises(capi.SbsError, match="unable to open"):
        load_tetgen(os.path.join(G, "nope.node"), os.path.join(G, "tet.ele"))
    with pytest.raises(capi.SbsError, match="out of range"):      # 1-based file read as 0-based
        load_tetgen(os.path.join(G, "tet.node"), os.path.join(G, "tet.ele"), start_index=0)


def test_data_loader_concatenates_bodies():
    dl = DataLoader(precision=32)
    dl.CollectData(os.path.join(G, "cube.1.node"), os.path.join(G, "cube.1.ele"), None, (0, 10, 25), (5, 5, 5),
                   (0, 0, 0), False, 1, SoftBodyAttribute(mass=10.0, mu=2e5, lam=5e3, DBC=[0, 3]))
    dl.CollectData(os.path.join(G, "house2.node"), os.path.join(G, "house2.ele"), None, (0, 40, 0), (4, 4, 4),
                   (0, 0, 0), True, 1, SoftBodyAttribute(mass=1.0, mu=2e4, lam=5e3))
    s = dl.AllocData()
    assert s.nv == 408 and s.nt == 1395 and s.ntri == 12 + 650
    assert s.tet[6:].min() == 8 and s.tri[12:].min() >= 8
    np.testing.assert_array_equal(s.tri_fathers, [0] * 12 + [1] * 650)
    np.testing.assert_array_equal(np.nonzero(s.dbc)[0], [0, 3])
    assert (s.mass[:8] == 10.0).all() and (s.mass[8:] == 1.0).all()
    assert (s.mu[:6] == 2e5).all() and (s.mu[6:] == 2e4).all()
    assert len(dl.edges) == 18 + 975 and dl.edges[18:].min() >= 8
