"""Pins of the fixed-body oracle (oracle/fixed_oracle.cpp) and of the scene-side
model matrices (sbs_amd/fixed.py), on the CPU.

* body parameters: the six walls of the reference's context.json
  (context.json:109-117) have inward unit normals known in closed form; a
  sheared, scaled, rotated plane matches numpy's inverse transpose; cylinder
  axes are the rotated y axis;
* the response in closed form: penetration removed exactly along the normal,
  v_N reflected times muN, v_T shortened by muT (1 + muN) |v_N| and clamped at
  0 (fixedBodyData.cu:80-85); the plane acts only on vertices moving into it
  (:76); the cylinder is infinite and keeps the axial coordinate (:118-133);
* the visiting order: spheres, then planes, then cylinders, whatever the
  creation order (fixedBodyData.cu:139-147).
"""
import numpy as np
import pytest

import oracle as O
from sbs_amd import fixed as F

WALLS = [  # (pos, rot, scale) of context.json's fixed planes, and the inward normal
    ((0, 0, 0), (0, 0, 0), (450, 450, 450), (0, 1, 0)),
    ((0, 450, 0), (0, 0, 180), (450, 450, 450), (0, -1, 0)),
    ((450, 0, 0), (0, 0, 90), (450, 450, 450), (-1, 0, 0)),
    ((-450, 0, 0), (0, 0, -90), (450, 450, 450), (1, 0, 0)),
    ((0, 0, -300), (90, 0, 0), (450, 450, 450), (0, 0, 1)),
    ((0, 0, 450), (-90, 0, 0), (450, 450, 450), (0, 0, -1)),
]


def params_of(bodies):
    types = [b.type for b in bodies]
    models = np.stack([np.asarray(b.model, np.float32).T.reshape(-1) for b in bodies])
    radii = [b.radius for b in bodies]
    return types, O.fixed_params(types, models, radii)


def test_model_matrix_composition():
    M = F.model_matrix((1, 2, 3), (0, 0, 90), (2, 3, 4))
    # T R S applied to the y axis: S stretches it by 3, Rz(90) turns +y into -x
    p = M @ np.array([0, 1, 0, 1.0])
    np.testing.assert_allclose(p[:3], [1 - 3, 2, 3], atol=1e-5)
    np.testing.assert_allclose(M[:3, 3], [1, 2, 3])


def test_wall_normals_point_inward():
    bodies = [F.Plane(p, r, s) for p, r, s, _ in WALLS]
    _, prm = params_of(bodies)
    for (p, _, _, up), q in zip(WALLS, prm):
        np.testing.assert_allclose(q[3:6], up, atol=1e-6)
        np.testing.assert_allclose(q[:3], p, atol=1e-3)


def test_plane_normal_of_a_general_affine_model():
    rng = np.random.default_rng(5)
    for _ in range(20):
        A = rng.normal(size=(3, 3)) + 2 * np.eye(3)
        M = np.eye(4, dtype=np.float32)
        M[:3, :3] = A
        M[:3, 3] = rng.normal(size=3)
        b = F.FixedBody(F.PLANE, M)
        _, prm = params_of([b])
        n = np.linalg.inv(M.astype(np.float64)).T @ np.array([0, 1.0, 0, 0])
        n = n[:3] / np.linalg.norm(n[:3])
        np.testing.assert_allclose(prm[0, 3:6], n, atol=2e-6)


def test_cylinder_axis_and_radius():
    b = F.Cylinder((5, 80, 0), (90, 0, 0), (10, 75, 10))
    _, prm = params_of([b])
    np.testing.assert_allclose(prm[0, 3:6], [0, 0, 1], atol=1e-6)
    assert prm[0, 6] == 10.0


@pytest.mark.parametrize("fp64", [False, True])
def test_plane_response_closed_form(fp64):
    types, prm = params_of([F.Plane()])
    X = np.array([[0.2, -0.1, 0.3], [0.2, -0.1, 0.3], [0.0, -0.1, 0.0], [0.0, 0.2, 0.0], [1.0, -0.05, 2.0]])
    V = np.array([[3.0, -4.0, 0.0], [10.0, -4.0, 0.0], [1.0, 2.0, 0.0], [0.0, -3.0, 0.0], [0.0, -2.0, 0.0]])
    Xo, Vo = O.fixed_handle(X, V, types, prm, mu_t=0.5, mu_n=0.5, fp64=fp64)
    tol = 1e-12 if fp64 else 1e-6
    # 0: a = 1 - 0.5 * 1.5 * 4 / 3 = 0: the tangential velocity stops
    np.testing.assert_allclose(Xo[0], [0.2, 0.0, 0.3], atol=tol)
    np.testing.assert_allclose(Vo[0], [0.0, 2.0, 0.0], atol=tol)
    # 1: a = 1 - 0.75 * 4 / 10 = 0.7
    np.testing.assert_allclose(Vo[1], [7.0, 2.0, 0.0], atol=10 * tol)
    # 2: below the plane but moving out of it: untouched
    np.testing.assert_array_equal(Xo[2], X[2].astype(Xo.dtype))
    np.testing.assert_array_equal(Vo[2], V[2].astype(Vo.dtype))
    # 3: above the plane: untouched
    np.testing.assert_array_equal(Vo[3], V[3].astype(Vo.dtype))
    # 4: no tangential velocity (|v_T| = 0): a = 0, v = -muN v_N
    np.testing.assert_allclose(Vo[4], [0.0, 1.0, 0.0], atol=tol)


@pytest.mark.parametrize("fp64", [False, True])
def test_sphere_and_cylinder_projection(fp64):
    types, prm = params_of([F.Sphere((1, 1, 1), radius=2.0), F.Cylinder((0, 0, 10), (0, 0, 0), (1, 5, 1))])
    X = np.array([[1.6, 1.8, 1.0], [0.3, 5.0, 10.4], [4.0, 4.0, 4.0]])
    V = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 1.0]])
    Xo, Vo = O.fixed_handle(X, V, types, prm, mu_t=0.0, mu_n=1.0, fp64=fp64)
    tol = 1e-12 if fp64 else 1e-6
    # inside the sphere (distance 1): pushed out along the ray to radius 2
    np.testing.assert_allclose(Xo[0], [1 + 1.2, 1 + 1.6, 1.0], atol=tol)
    # muN = 1, muT = 0: elastic reflection of the normal part, tangent kept
    n = np.array([0.6, 0.8, 0.0])
    v = np.array([1.0, 0.0, 0.0])
    np.testing.assert_allclose(Vo[0], v - 2 * (v @ n) * n, atol=tol)
    # inside the y-axis cylinder (radial distance 0.5): radial move to 1, y kept
    np.testing.assert_allclose(Xo[1], [0.6, 5.0, 10.8], atol=tol)
    np.testing.assert_allclose(Vo[1], [0.0, 1.0, 0.0], atol=tol)
    np.testing.assert_array_equal(Xo[2], X[2].astype(Xo.dtype))


def test_spheres_before_planes_whatever_the_creation_order():
    s, p = F.Sphere((0, 0, 0), radius=1.0), F.Plane((0, 0.5, 0))
    X = np.array([[0.0, 0.2, 0.0]])   # inside the sphere and below the plane
    V = np.array([[0.0, -1.0, 0.0]])
    t1, p1 = params_of([s, p])
    t2, p2 = params_of([p, s])
    a = O.fixed_handle(X, V, t1, p1, 0.5, 0.5)
    b = O.fixed_handle(X, V, t2, p2, 0.5, 0.5)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    # by hand: the sphere pushes it to (0, 1, 0) and reflects v to (0, 0.5, 0);
    # it is then above the plane, which leaves it
    np.testing.assert_allclose(a[0][0], [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(a[1][0], [0, 0.5, 0], atol=1e-6)


def test_unknown_type_is_refused():
    with pytest.raises(ValueError):
        O.fixed_params([7], np.eye(4, dtype=np.float32).reshape(1, 16), [1.0])
