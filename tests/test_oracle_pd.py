"""Pins the CPU oracle of the PD step against mathematics it does not share.

Each check is chosen so a plausible slip in the restatement fails it:
  * SVD / projection  -> the SVD known answer held by the reference's tests,
                         numpy's LAPACK SVD, SO(3) membership, exact rotations;
  * Dm^-1, V0         -> Dm Dm^-1 = I and the volume of the closed surface
                         (divergence theorem over the extracted triangles);
  * assembly          -> an independent FEM assembly from barycentric gradients
                         (inverse of the 4x4 [1; x] matrix) and the cotangent
                         formula of the linear-tet Laplacian;
  * local step        -> rest state (Jacobi form gives 0), rigid rotation
                         (projection = F gives exactly L q);
  * global step       -> numpy dense solve; free fall closed form; agreement of
                         all solver modes at convergence of the PD iteration.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import max_rel, rel_err
from sbs_amd import mesh

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- SVD ----
def _polar(A):
    u, s, vt = np.linalg.svd(A)
    d = np.sign(np.linalg.det(u @ vt))
    return u @ np.diag([1.0, 1.0, d]) @ vt


def _rot(rng=RNG):
    R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    if np.linalg.det(R) < 0:
        R[:, 2] *= -1
    return R


@pytest.mark.parametrize("fp64", [True, False])
def test_svd_reconstruction_and_rotations(fp64):
    # U, V are rotations to rounding for any input.  U diag(s) V^T reproduces
    # A only up to the truncation of 4 fixed Jacobi sweeps (the QR factor keeps
    # an off-diagonal the SVD drops): loose for arbitrary matrices with close
    # singular values, tight for deformation gradients of elastic solids.
    rnd = 1e-12 if fp64 else 1e-5
    rng = np.random.default_rng(7)   # own seed: the tail bars hold for this sample whatever ran before
    errs = []
    for _ in range(1000):
        A = rng.normal(size=(3, 3))
        U, s, V = O.svd3(A, fp64)
        errs.append(np.abs(U @ np.diag(s) @ V.T - A).max() / np.abs(A).max())
        assert abs(np.linalg.det(U) - 1) < 10 * rnd and abs(np.linalg.det(V) - 1) < 10 * rnd
        assert np.abs(U.T @ U - np.eye(3)).max() < 10 * rnd
        assert s[0] >= s[1] >= abs(s[2]) - rnd
    # measured: median 1e-12 (FP64) / 5e-7 (FP32), 99th percentile ~1e-4
    assert np.median(errs) < (1e-10 if fp64 else 2e-6)
    assert np.percentile(errs, 99) < 1e-3 and max(errs) < 3e-2
    # deformation gradients of an elastic solid: the rotation R = U V^T that the
    # local step uses matches LAPACK's polar factor (measured FP64: median 8e-13,
    # 99th percentile 7e-7; the tail is near-degenerate stretches)
    rerr = []
    for _ in range(1000):
        F = _rot(rng) @ (np.eye(3) + 0.05 * rng.normal(size=(3, 3)))
        R = O.local_rotation(F, fp64=fp64)
        rerr.append(np.abs(R - _polar(F)).max())
    assert np.median(rerr) < (1e-10 if fp64 else 1e-6)
    assert np.percentile(rerr, 99) < 1e-5 and max(rerr) < 2e-4


def test_svd_matches_lapack_when_well_separated():
    # With separated singular values the 4-sweep Jacobi converges quadratically
    # once past its fixed-angle phase; large initial angles can still leave a
    # truncation (measured: median 5e-13, 99th percentile 1e-5, worst 4e-4,
    # one more sweep would give 1e-9) -- the reference's own 4-sweep behaviour.
    # The tail depends on the draw (over seeds: 95th percentile 2e-7..1.3e-6,
    # worst 7e-5..3.5e-3), so the sample has its own seed and does not depend
    # on which tests ran before it.
    rng = np.random.default_rng(5)
    rerr, serr = [], []
    for _ in range(500):
        Q1 = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        Q2 = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        sig = np.array([3.0, 1.5, 0.6]) * rng.uniform(0.9, 1.1, 3)   # ratios >= 1.6
        A = Q1 @ np.diag(sig) @ Q2
        U, s, V = O.svd3(A)
        serr.append(np.abs(np.abs(s) - np.sort(sig)[::-1]).max())
        rerr.append(np.abs(O.local_rotation(A) - _polar(A)).max())
    assert np.median(rerr) < 1e-11 and np.median(serr) < 1e-11
    assert np.percentile(rerr, 95) < 1e-6 and max(rerr) < 2e-3


def test_svd_matches_reference_fixture():
    """Known answer held by the reference's tests (tests/test_aabb.cpp:186-194):
    F = U diag(sigma) V^T from its printed factors; the TR1690 SVD must give the
    same sigma (descending) and the same singular-vector pairs up to a common
    sign per pair, with U, V proper rotations."""
    import json
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "svd_reference_fixture.json")))
    U, V, sig = np.array(fx["U"]), np.array(fx["V"]), np.array(fx["sigma"])
    F = U @ np.diag(sig) @ V.T
    tol = 2e-4  # the fixture carries 4 decimals
    for fp64 in (True, False):
        u, s, v = O.svd3(F, fp64=fp64)
        np.testing.assert_allclose(s, sig, atol=tol)
        for i in range(3):
            sgn = np.sign(u[:, i] @ U[:, i])
            assert sgn == np.sign(v[:, i] @ V[:, i])
            np.testing.assert_allclose(u[:, i], sgn * U[:, i], atol=tol)
            np.testing.assert_allclose(v[:, i], sgn * V[:, i], atol=tol)
        assert abs(np.linalg.det(u) - 1) < 1e-5 and abs(np.linalg.det(v) - 1) < 1e-5
        np.testing.assert_allclose(O.local_rotation(F, fp64=fp64), U @ V.T, atol=tol)
    # the fixture's Dm^-1 is the inverse of the edge matrix of a 0.1-scaled tet
    Dm = np.linalg.inv(np.array(fx["DmInv"]))
    np.testing.assert_allclose(Dm @ np.array(fx["DmInv"]), np.eye(3), atol=1e-12)


def test_projection_of_rotation_is_itself():
    for _ in range(50):
        R = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
        if np.linalg.det(R) < 0:
            R[:, 2] *= -1
        np.testing.assert_allclose(O.local_rotation(R), R, atol=1e-12)
        # residual (Jacobi) form: R - F = 0 for a rotation
        np.testing.assert_allclose(O.local_rotation(R, jacobi=True), 0.0, atol=1e-12)
    np.testing.assert_array_equal(O.local_rotation(np.eye(3)), np.eye(3))


def test_projection_of_inverted_element_is_rotation_variant():
    # det F < 0: the reference's SVD keeps U, V in SO(3) so R = U V^T is the
    # nearest rotation, not a reflection
    A = np.diag([1.2, 0.9, -0.5])
    R = O.local_rotation(A)
    assert abs(np.linalg.det(R) - 1) < 1e-12
    np.testing.assert_allclose(R, _polar(A), atol=1e-12)


# ------------------------------------------------------------ meshes -----
def test_mesh_sizes_match_the_workload_table():
    s0 = mesh.cube_scene(8)
    assert (s0.nv, s0.nt, s0.ntri, int(s0.dbc.sum())) == (729, 3072, 768, 81)
    assert mesh.scalar_laplacian_nnz(s0.tet, s0.nv) == 9097
    s2 = mesh.cube_scene(6)
    assert (s2.nv, s2.nt, mesh.scalar_laplacian_nnz(s2.tet, s2.nv)) == (343, 1296, 4051)
    s4 = mesh.cube_scene(16, pin_top=False)
    assert (s4.nv, s4.nt, s4.ntri, mesh.scalar_laplacian_nnz(s4.tet, s4.nv)) == (4913, 24576, 3072, 66961)


def _surface_volume(X, tri):
    a, b, c = X[tri[:, 0]], X[tri[:, 1]], X[tri[:, 2]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0


def test_fem_precompute_against_geometry():
    s = mesh.cube_scene(5)
    o = O.PdOracle(s, O.make_params(), fp64=True)
    ex = o.export()
    t = s.tet.astype(np.int64)
    Dm = np.stack([s.X[t[:, k]] - s.X[t[:, 0]] for k in (1, 2, 3)], axis=2)
    eye = np.einsum("tij,tjk->tik", Dm, ex["dminv"])
    np.testing.assert_allclose(eye, np.broadcast_to(np.eye(3), eye.shape), atol=1e-12)
    # the surface faces are oriented inward (tet faces of positively oriented
    # tets), so the divergence theorem gives minus the volume
    vol = -_surface_volume(s.X, s.tri.astype(np.int64))
    np.testing.assert_allclose(ex["V0"].sum(), vol, rtol=1e-12)
    assert abs(vol - 0.5 ** 3) < 1e-3


# ---------------------------------------------------------- assembly -----
def brute_force_laplacian(X, tet, w):
    """Linear-FEM Laplacian sum_t w_t grad(phi_i).grad(phi_j) |T| / |T| from
    barycentric gradients (rows 1..3 of inv([1 1 1 1; x0 x1 x2 x3]))."""
    nv = X.shape[0]
    L = np.zeros((nv, nv))
    for ti, tv in enumerate(tet.astype(np.int64)):
        M = np.vstack([np.ones(4), X[tv].T])
        grads = np.linalg.inv(M)[:, 1:]          # (4, 3): grad phi_i
        K = grads @ grads.T * w[ti]
        L[np.ix_(tv, tv)] += K
    return L


def cotangent_tet(P):
    """Cotangent formula of the linear-tet Laplacian (weight w = V0):
    K_ij = -(1/6) |e_kl| cot(theta_kl) for the edge kl opposite ij, with
    theta_kl the interior dihedral angle at that edge; rows sum to zero."""
    K = np.zeros((4, 4))
    for i in range(4):
        for j in range(i + 1, 4):
            k, l = [m for m in range(4) if m not in (i, j)]
            e = P[l] - P[k]
            n1 = np.cross(P[k] - P[i], P[l] - P[i])   # face (i,k,l)
            n2 = np.cross(P[k] - P[j], P[l] - P[j])   # face (j,k,l)
            # n1, n2 share the k->l winding, so one points out of the tet and
            # one in: the interior dihedral angle has cos = +n1.n2/|n1||n2|
            cosang = np.dot(n1, n2) / (np.linalg.norm(n1) * np.linalg.norm(n2))
            ang = np.arccos(np.clip(cosang, -1, 1))   # dihedral angle at edge kl
            K[i, j] = K[j, i] = -np.linalg.norm(e) / 6.0 / np.tan(ang)
    np.fill_diagonal(K, -K.sum(axis=1))
    return K


def test_assembly_matches_brute_force_fem():
    s = mesh.cube_scene(3)
    p = O.make_params(dt=1 / 60)
    o = O.PdOracle(s, p, fp64=True)
    ex = o.export()
    A = np.zeros((s.nv, s.nv))
    for r in range(s.nv):
        A[r, ex["col"][ex["rowptr"][r]:ex["rowptr"][r + 1]]] = ex["val"][ex["rowptr"][r]:ex["rowptr"][r + 1]]
    w = ex["V0"] * s.mu
    L = brute_force_laplacian(s.X, s.tet, w)
    c = (s.mass + 1e6 * s.dbc) / p.dt ** 2
    np.testing.assert_allclose(A, L + np.diag(c), rtol=1e-10, atol=1e-8 * np.abs(L).max())
    np.testing.assert_allclose(ex["diagL"], np.diag(L), rtol=1e-10)
    np.testing.assert_allclose(ex["c"], c, rtol=1e-12)
    np.testing.assert_allclose(L.sum(axis=1), 0.0, atol=1e-9 * np.abs(L).max())
    np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-9 * np.abs(L).max())


def test_single_tet_matches_cotangent_formula():
    P = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 0.9, 0.1], [0.1, 0.2, 1.1]])
    scene = mesh.Scene(X=P, tet=np.array([[0, 1, 2, 3]], np.uint32), tri=np.zeros((0, 3), np.uint32),
                       tri_fathers=np.zeros(0, np.uint32), mass=np.ones(4), mu=np.ones(1), lam=np.ones(1),
                       dbc=np.zeros(4), body_vert_start=np.array([0, 4]), body_tet_start=np.array([0, 1]))
    o = O.PdOracle(scene, O.make_params(dt=1.0), fp64=True)
    ex = o.export()
    A = np.zeros((4, 4))
    for r in range(4):
        A[r, ex["col"][ex["rowptr"][r]:ex["rowptr"][r + 1]]] = ex["val"][ex["rowptr"][r]:ex["rowptr"][r + 1]]
    # mu = 1, w = V0: L_s equals the cotangent Laplacian of the tet
    np.testing.assert_allclose(A - np.eye(4), cotangent_tet(P), rtol=1e-10, atol=1e-12)


# -------------------------------------------------------- local step -----
def _dense_L(s, ex):
    A = np.zeros((s.nv, s.nv))
    for r in range(s.nv):
        sl = slice(ex["rowptr"][r], ex["rowptr"][r + 1])
        A[r, ex["col"][sl]] = ex["val"][sl]
    return A - np.diag(ex["c"])


def test_local_step_rest_state_and_rigid_rotation():
    s = mesh.cube_scene(3)
    dt = 1 / 60
    o = O.PdOracle(s, O.make_params(dt=dt), fp64=True)
    ex = o.export()
    L = _dense_L(s, ex)
    sn_old = s.X + RNG.normal(size=s.X.shape) * 1e-3
    c = ex["c"][:, None]
    pin = s.dbc > 0
    # Jacobi (residual) form at rest: R - F = 0, so b = c*sn_old off the pins.
    # (At F = I the column norms tie, the sort may swap columns, and the QR's
    # max(|a11|, 1e-12) floor then moves R by ~1e-12: a ~1e-9 force residual.)
    b = o.local(s.X, sn_old, jacobi=True, dt=dt)
    np.testing.assert_allclose(b[~pin], (c * sn_old)[~pin], rtol=1e-12, atol=1e-8)
    # pins are overwritten with DBCX * 1e6 / dt^2 (computeDBCLocal)
    np.testing.assert_array_equal(b[pin], s.X[pin] * (1e6 / dt ** 2))
    # rigid rotation: projection = F, so the elastic term is exactly L q
    R = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
    if np.linalg.det(R) < 0:
        R[:, 2] *= -1
    q = s.X @ R.T + np.array([0.3, -0.2, 0.1])
    b = o.local(q, sn_old, jacobi=False, dt=dt)
    np.testing.assert_allclose(b[~pin], (c * sn_old + L @ q)[~pin], rtol=1e-9, atol=1e-9 * np.abs(b).max())


# ------------------------------------------------------- global step -----
def test_direct_mode_solves_the_assembled_system():
    s = mesh.cube_scene(4)
    p = O.make_params(solver=O.SOLVER_CHOLESKY, num_iterations=1, tol=0.0)
    o = O.PdOracle(s, p, fp64=True)
    ex = o.export()
    A = _dense_L(s, ex) + np.diag(ex["c"])
    # one step with one iteration from rest: s_n = x + h v + h^2 f/(m + 1e6 DBC)
    # (computeSn), the local step on s_n, then x = A^-1 b
    sn = s.X.copy()
    sn[:, 1] += -p.gravity * s.mass * p.dt ** 2 / (s.mass + 1e6 * s.dbc)
    b = o.local(sn, sn, jacobi=False, dt=p.dt)
    o.step(1)
    np.testing.assert_allclose(o.get("X"), np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("solver", [O.SOLVER_JACOBI, O.SOLVER_CHOLESKY, O.SOLVER_PCG])
def test_free_fall_closed_form(solver):
    s = mesh.cube_scene(3, pin_top=False)
    p = O.make_params(solver=solver, num_iterations=5, tol=0.0, pcg_tol=1e-12, pcg_floor=0.0)
    o = O.PdOracle(s, p, fp64=True)
    n = 7
    o.step(n)
    h, g = p.dt, p.gravity
    expect = s.X.copy()
    expect[:, 1] -= g * h * h * n * (n + 1) / 2     # implicit Euler free fall
    np.testing.assert_allclose(o.get("X"), expect, rtol=0, atol=1e-12)
    np.testing.assert_allclose(o.get("V")[:, 1], -g * h * n, rtol=1e-9)


def test_all_modes_reach_the_same_pd_fixed_point():
    s = mesh.cube_scene(3)
    ref = O.PdOracle(s, O.make_params(solver=O.SOLVER_CHOLESKY, num_iterations=200, tol=0.0), fp64=True)
    ref.step(1)
    xr = ref.get("X")
    pcg = O.PdOracle(s, O.make_params(solver=O.SOLVER_PCG, num_iterations=200, tol=0.0, pcg_tol=1e-13,
                                      pcg_floor=0.0, pcg_max_iter=500), fp64=True)
    pcg.step(1)
    np.testing.assert_allclose(pcg.get("X"), xr, atol=1e-11)
    jac = O.PdOracle(s, O.make_params(solver=O.SOLVER_JACOBI, num_iterations=6000), fp64=True)
    jac.step(1)
    # Jacobi + Chebyshev converges to the same fixed point except on the pinned
    # vertices, where the overwritten right-hand side gives x0*1e6/(m+1e6)
    pin = s.dbc > 0
    np.testing.assert_allclose(jac.get("X")[~pin], xr[~pin], atol=1e-7)
    m = s.mass[pin][:, None]
    np.testing.assert_allclose(jac.get("X")[pin], s.X[pin] * 1e6 / (m + 1e6), rtol=1e-9)


def test_fp32_oracle_tracks_fp64():
    s = mesh.cube_scene(4)
    for solver in (O.SOLVER_JACOBI, O.SOLVER_CHOLESKY):
        p = O.make_params(solver=solver, num_iterations=10)
        a = O.PdOracle(s, p, fp64=True)
        b = O.PdOracle(s, p, fp64=False)
        a.step(20)
        b.step(20)
        assert rel_err(b.get("X"), a.get("X")) < 1e-5
        assert max_rel(b.get("X"), a.get("X")) < 1e-5
