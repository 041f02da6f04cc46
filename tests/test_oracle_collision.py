"""Pins the CPU oracle of the collision path against facts it does not share.

  * CCD           -> the known answer held by the reference's tests
                     (tests/test_solveCubic.cpp:126-172: times of impact and the
                     normal it records), textbook cubics with known roots, and
                     closed-form plane crossings of a moving vertex / edge;
  * broad phase   -> closed-form query counts (two triangles: 6 VF + 18 EE),
                     the adjacency, same-triangle and same-body filters, and
                     monotonicity in the separation of two bodies;
  * hit handling  -> the vertex set and normal signs of a single crossing;
  * CCD response  -> V = -(dx.n) n on hit vertices, X = XTilde elsewhere.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from sbs_amd import mesh

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_ccd_reference_fixture():
    fx = json.load(open(os.path.join(GOLDEN, "ccd_reference_fixture.json")))
    X, XT = np.array(fx["X"]), np.array(fx["XTilde"])
    tois = []
    for q in fx["queries"]:
        t, n = O.ccd_test(q[0], q[1:], X, XT, fp64=True)
        tois.append(t)
        if q[1:] == [3, 4, 7, 10]:
            # the fixture records toi 0.994723 and normal (0.000010, 1, 0.000010)
            assert abs(t - 0.994723) < 1e-6
            np.testing.assert_allclose(n, [0.000010, 1.0, 0.000010], atol=1e-6)
    assert min(tois) != 1.0
    assert sum(t < 1 for t in tois) == 1


@pytest.mark.parametrize("roots", [(0.1, 0.5, 0.9), (-1.5, 0.25, 2.0), (0.3, 0.31, 5.0)])
def test_solve_cubic_textbook(roots):
    a, b, c = roots
    coef = np.poly([a, b, c])  # (t-a)(t-b)(t-c)
    x = np.sort(O.solve_cubic(*coef))
    np.testing.assert_allclose(x, np.sort(roots), atol=2e-6)


def test_solve_cubic_single_real_root():
    # t^3 + t + 1: one real root, no real critical point
    x = O.solve_cubic(1.0, 0.0, 1.0, 1.0)
    assert len(x) == 1
    assert abs(x[0] ** 3 + x[0] + 1.0) < 1e-9


def _tri_scene(h0, h1, offset=(0.0, 0.0)):
    # static triangle in the plane y = 0 and one vertex moving from y = h0 to y = h1
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.2 + offset[0], h0, 0.2 + offset[1]]])
    XT = X.copy()
    XT[3, 1] = h1
    return X, XT


@pytest.mark.parametrize("fp64", [True, False])
def test_ccd_vertex_through_face_closed_form(fp64):
    X, XT = _tri_scene(0.3, -0.1)
    t, n = O.ccd_test(O.QUERY_VF, [3, 0, 1, 2], X, XT, fp64=fp64)
    assert abs(t - 0.75) < (1e-9 if fp64 else 1e-6)      # 0.3 / (0.3 + 0.1)
    np.testing.assert_allclose(np.abs(n), [0, 1, 0], atol=1e-6)
    # the same motion outside the triangle, or stopping short of the plane: no hit
    X2, XT2 = _tri_scene(0.3, -0.1, offset=(1.0, 1.0))
    assert O.ccd_test(O.QUERY_VF, [3, 0, 1, 2], X2, XT2, fp64=fp64)[0] == 1.0
    X3, XT3 = _tri_scene(0.3, 0.05)
    assert O.ccd_test(O.QUERY_VF, [3, 0, 1, 2], X3, XT3, fp64=fp64)[0] == 1.0


def test_ccd_edge_edge_crossing_closed_form():
    # edge a along x at height y = 0.4 moving down to -0.4; edge b along z at y = 0
    X = np.array([[-1.0, 0.4, 0.3], [1.0, 0.4, 0.3], [0.2, 0.0, -1.0], [0.2, 0.0, 1.0]])
    XT = X.copy()
    XT[0:2, 1] = -0.4
    t, n = O.ccd_test(O.QUERY_EE, [0, 1, 2, 3], X, XT, fp64=True)
    assert abs(t - 0.5) < 1e-9
    np.testing.assert_allclose(np.abs(n), [0, 1, 0], atol=1e-9)


def test_broad_phase_two_triangles_count():
    # two overlapping triangles with distinct vertices: each ordered pair gives
    # 3 VF + 9 EE; EE (eA, eB) and (eB, eA) stay distinct -> 6 VF + 18 EE
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [0.2, 0.01, 0.2], [1.2, 0.01, 0.2], [0.2, 0.01, 1.2]], float)
    tri = np.array([[0, 1, 2], [3, 4, 5]])
    out = O.collision_detect(X, X, tri, [0, 1], ignore_self=True, resolve=False)
    q = out["queries"]
    assert (q[:, 0] == O.QUERY_VF).sum() == 6 and (q[:, 0] == O.QUERY_EE).sum() == 18
    # same body with ignore_self: nothing
    assert len(O.collision_detect(X, X, tri, [0, 0], ignore_self=True, resolve=False)["queries"]) == 0
    assert len(O.collision_detect(X, X, tri, [0, 0], ignore_self=False, resolve=False)["queries"]) == 24
    # sharing a vertex: excluded
    tri2 = np.array([[0, 1, 2], [2, 4, 5]])
    assert len(O.collision_detect(X, X, tri2, [0, 1], ignore_self=False, resolve=False)["queries"]) == 0
    # 0.02 plus a margin apart along y: boxes (padded by 0.01 each) do not touch
    X3 = X.copy()
    X3[3:, 1] = 0.0205
    assert len(O.collision_detect(X3, X3, tri, [0, 1], resolve=False)["queries"]) == 0


def two_cube_scene(gap, n=3, drop=0.0, fp64=False):
    s = mesh.make_scene([dict(nx=n, ny=n, nz=n, origin=(0.0, 0.0, 0.0), seed=1),
                         dict(nx=n, ny=n, nz=n, origin=(0.05, n * 0.1 + gap, 0.03), seed=2)])
    X = s.X.copy()
    XT = X.copy()
    upper = np.arange(s.nv) >= s.body_vert_start[1]
    XT[upper, 1] -= drop
    return s, X, XT


def test_broad_phase_filters_and_monotonicity():
    counts = []
    for gap in (0.5, 0.05, 0.015, 0.0):
        s, X, XT = two_cube_scene(gap)
        out = O.collision_detect(X, XT, s.tri, s.tri_fathers, ignore_self=True, resolve=False)
        q = out["queries"]
        counts.append(len(q))
        body = (q[:, 1:] >= s.body_vert_start[1])
        # every query pairs the two bodies
        if len(q):
            vf = q[:, 0] == O.QUERY_VF
            assert np.all(body[vf, 0] != body[vf, 1]) and np.all(body[vf, 1] == body[vf, 2])
            assert np.all(body[vf, 2] == body[vf, 3])
            ee = ~vf
            assert np.all(body[ee, 0] == body[ee, 1]) and np.all(body[ee, 2] == body[ee, 3])
            assert np.all(body[ee, 0] != body[ee, 2])
            # canonical and unique
            assert np.all(q[vf, 2] < q[vf, 3]) and np.all(q[vf, 3] < q[vf, 4])
            assert np.all(q[ee, 1] < q[ee, 2]) and np.all(q[ee, 3] < q[ee, 4])
            assert len(np.unique(q, axis=0)) == len(q)
    assert counts[0] == 0
    assert counts == sorted(counts)
    assert counts[-1] > 0


@pytest.mark.parametrize("fp64", [True, False])
def test_hit_resolution_and_response(fp64):
    # a single vertex crossing a static triangle; the vertex belongs to a
    # second triangle whose other corners are far away and do not move
    X, XT = _tri_scene(0.3, -0.1)
    X = np.vstack([X, [[5.0, 5.0, 5.0], [5.0, 6.0, 5.0]]])
    XT = np.vstack([XT, [[5.0, 5.0, 5.0], [5.0, 6.0, 5.0]]])
    tri = np.array([[0, 1, 2], [3, 4, 5]])
    out = O.collision_detect(X, XT, tri, [0, 1], ignore_self=True, fp64=fp64)
    assert out["hits"] == 1
    assert np.array_equal(out["tI"] < 1, np.array([1, 1, 1, 1, 0, 0], bool))
    nv = out["vertex_normals"]
    assert abs(abs(nv[3, 1]) - 1) < 1e-6
    np.testing.assert_array_equal(nv[0], -nv[3])
    V = np.zeros_like(X)
    Xn, Vn = O.ccd_response(X, XT, V, out["tI"], nv, fp64=fp64)
    dx = XT[3] - X[3]
    np.testing.assert_allclose(Vn[3], -np.dot(dx, nv[3].astype(float)) * nv[3], atol=1e-6)
    np.testing.assert_array_equal(Xn[3], X[3].astype(Xn.dtype))      # hit: X kept
    np.testing.assert_array_equal(Xn[4:], XT[4:].astype(Xn.dtype))   # no hit: X = XTilde
