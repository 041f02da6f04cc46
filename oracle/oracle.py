"""ctypes bindings for the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this module, and only as the checker or the timed CPU baseline.  The product
library never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

SOLVER_JACOBI = 0
SOLVER_CHOLESKY = 1
SOLVER_EIGEN_CHOLESKY = 2
SOLVER_PCG = 3


class OrcParams(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int), ("num_iterations", ctypes.c_int), ("dt", ctypes.c_double),
                ("gravity", ctypes.c_double), ("tol", ctypes.c_double), ("rho", ctypes.c_double),
                ("positional_weight", ctypes.c_double), ("pcg_tol", ctypes.c_double),
                ("pcg_floor", ctypes.c_double), ("pcg_max_iter", ctypes.c_int), ("pcg_chebyshev", ctypes.c_int)]


def make_params(solver=SOLVER_JACOBI, num_iterations=10, dt=1.0 / 60.0, gravity=9.8, tol=1e-6, rho=0.9992,
                positional_weight=1e6, pcg_tol=1e-6, pcg_floor=1e-7, pcg_max_iter=200,
                pcg_chebyshev=0) -> OrcParams:
    return OrcParams(solver, num_iterations, dt, gravity, tol, rho, positional_weight, pcg_tol, pcg_floor,
                     pcg_max_iter, pcg_chebyshev)


_LIBS: dict[str, ctypes.CDLL] = {}


def build(bench: bool = False) -> str:
    name = "liboracle_bench.so" if bench else "liboracle.so"
    subprocess.run(["make", "-s", "-C", HERE, name], check=True)
    return os.path.join(HERE, name)


def lib(bench: bool = False) -> ctypes.CDLL:
    name = "liboracle_bench.so" if bench else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        path = build(bench)
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    up = ctypes.POINTER(ctypes.c_uint32)
    L.orc_svd3_f32.argtypes = [fp, fp, fp, fp]
    L.orc_svd3_f64.argtypes = [dp, dp, dp, dp]
    L.orc_local_rotation_f64.argtypes = [dp, ctypes.c_int, dp]
    L.orc_local_rotation_f32.argtypes = [fp, ctypes.c_int, fp]
    L.orc_pd_create.restype = ctypes.c_void_p
    L.orc_pd_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, up, dp, dp, dp,
                                ctypes.POINTER(OrcParams)]
    L.orc_pd_destroy.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.orc_pd_step.restype = ctypes.c_int
    L.orc_pd_step.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(OrcParams), ctypes.c_int, ip]
    L.orc_pd_get.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, dp]
    L.orc_pd_nnz.restype = ctypes.c_int
    L.orc_pd_nnz.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.orc_pd_export.argtypes = [ctypes.c_int, ctypes.c_void_p, dp, dp, dp, dp, ip, ip, dp]
    L.orc_pd_set_more_dbc.argtypes = [ctypes.c_int, ctypes.c_void_p, dp, dp]
    L.orc_pd_local.argtypes = [ctypes.c_int, ctypes.c_void_p, dp, dp, ctypes.c_int, ctypes.c_double,
                               ctypes.c_double, dp]
    L.orc_num_threads.restype = ctypes.c_int
    vp = ctypes.c_void_p
    L.orc_col_detect.restype = vp
    L.orc_col_detect.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, up, up, vp, vp, ctypes.c_int, vp, fp]
    L.orc_col_destroy.argtypes = [ctypes.c_int, vp]
    L.orc_col_num_queries.restype = ctypes.c_longlong
    L.orc_col_num_queries.argtypes = [ctypes.c_int, vp]
    L.orc_col_num_hits.restype = ctypes.c_int
    L.orc_col_num_hits.argtypes = [ctypes.c_int, vp]
    L.orc_col_get.argtypes = [ctypes.c_int, vp, up, dp, dp]
    L.orc_ccd_test.restype = ctypes.c_double
    L.orc_ccd_test.argtypes = [ctypes.c_int, ctypes.c_int, up, vp, vp, dp]
    L.orc_solve_cubic.restype = ctypes.c_int
    L.orc_solve_cubic.argtypes = [ctypes.c_double] * 4 + [dp]
    L.orc_ccd_response.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, fp]
    L.orc_col_min_step.restype = ctypes.c_double
    L.orc_col_min_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, up, up, vp, vp, ctypes.c_int]
    L.orc_col_proximity.restype = ctypes.c_longlong
    L.orc_col_proximity.argtypes = [ctypes.c_int, up, up, dp, ctypes.c_double, ctypes.c_int, up]
    if hasattr(L, "orc_fixed_params"):
        L.orc_fixed_params.restype = ctypes.c_int
        L.orc_fixed_params.argtypes = [ctypes.c_int, ip, fp, fp, fp]
        L.orc_fixed_handle.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ip, fp, ctypes.c_double, ctypes.c_double]
    if hasattr(L, "orc_ipc_create"):
        ucp = ctypes.POINTER(ctypes.c_ubyte)
        L.orc_ipc_distance_type.restype = ctypes.c_int
        L.orc_ipc_distance_type.argtypes = [ctypes.c_int, dp]
        L.orc_ipc_distance.restype = ctypes.c_double
        L.orc_ipc_distance.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.orc_ipc_barrier.restype = ctypes.c_double
        L.orc_ipc_barrier.argtypes = [ctypes.c_double] * 3 + [dp, dp]
        L.orc_ipc_make_pd.argtypes = [ctypes.c_int, dp]
        L.orc_ipc_create.restype = vp
        L.orc_ipc_create.argtypes = [ctypes.c_int] * 3 + [dp, dp, up, up, up, dp, dp, dp, ucp] + \
            [ctypes.c_double] * 5 + [ctypes.c_int]
        L.orc_ipc_destroy.argtypes = [vp]
        L.orc_ipc_set_elastic.argtypes = [vp, ctypes.c_int]
        L.orc_nh_density.restype = ctypes.c_double
        L.orc_nh_density.argtypes = [dp, ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_ipc_step.restype = ctypes.c_int
        L.orc_ipc_step.argtypes = [vp]
        L.orc_ipc_get.argtypes = [vp, dp, dp]
        L.orc_ipc_set_x.argtypes = [vp, dp]
        L.orc_ipc_set_xv.argtypes = [vp, dp, dp, ctypes.c_int]
        L.orc_ipc_stats.argtypes = [vp, ip, dp]
        L.orc_ipc_energy.restype = ctypes.c_double
        L.orc_ipc_energy.argtypes = [vp, dp, dp]
        L.orc_ipc_set_state.argtypes = [vp, dp, ctypes.c_double]
        L.orc_ipc_queries.restype = ctypes.c_longlong
        L.orc_ipc_queries.argtypes = [vp, up, ip, dp]
        L.orc_ipc_set_fixed.argtypes = [vp, ctypes.c_int, ip, fp]
        L.orc_ipc_contact_area.argtypes = [vp, dp]
        L.orc_ipc_fixed_vertex.restype = ctypes.c_double
        L.orc_ipc_fixed_vertex.argtypes = [vp, dp, ctypes.c_int, dp, dp]
        L.orc_ipc_fixed_step.restype = ctypes.c_double
        L.orc_ipc_fixed_step.argtypes = [vp, dp, dp]
    _LIBS[name] = L
    return L


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def svd3(A: np.ndarray, fp64: bool = True):
    """TR1690 SVD of one 3x3 matrix: returns U, sigma, V with A = U diag(sigma) V^T."""
    L = lib()
    dt, ct = (np.float64, ctypes.c_double) if fp64 else (np.float32, ctypes.c_float)
    A = np.ascontiguousarray(A, dtype=dt)
    U = np.zeros((3, 3), dt); s = np.zeros(3, dt); V = np.zeros((3, 3), dt)
    (L.orc_svd3_f64 if fp64 else L.orc_svd3_f32)(_p(A, ct), _p(U, ct), _p(s, ct), _p(V, ct))
    return U, s, V


def local_rotation(F: np.ndarray, jacobi: bool = False, fp64: bool = True) -> np.ndarray:
    L = lib()
    dt, ct = (np.float64, ctypes.c_double) if fp64 else (np.float32, ctypes.c_float)
    F = np.ascontiguousarray(F, dtype=dt)
    R = np.zeros((3, 3), dt)
    (L.orc_local_rotation_f64 if fp64 else L.orc_local_rotation_f32)(_p(F, ct), int(jacobi), _p(R, ct))
    return R


class PdOracle:
    """The reference PD step on the CPU for one flat scene (see pd_oracle.cpp)."""

    def __init__(self, scene, params: OrcParams, fp64: bool = True, bench: bool = False, V0=None):
        self.L = lib(bench)
        self.fp64 = int(fp64)
        self.params = params
        self.nv, self.nt = scene.nv, scene.nt
        X = np.ascontiguousarray(scene.X, dtype=np.float64)
        tet = np.ascontiguousarray(scene.tet, dtype=np.uint32)
        mass = np.ascontiguousarray(scene.mass, dtype=np.float64)
        mu = np.ascontiguousarray(scene.mu, dtype=np.float64)
        dbc = np.ascontiguousarray(scene.dbc, dtype=np.float64)
        v0 = None if V0 is None else np.ascontiguousarray(V0, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        self.h = self.L.orc_pd_create(self.fp64, self.nv, self.nt, _p(X, ctypes.c_double),
                                      None if v0 is None else _p(v0, ctypes.c_double),
                                      _p(tet, ctypes.c_uint32), _p(mass, ctypes.c_double),
                                      _p(mu, ctypes.c_double), _p(dbc, ctypes.c_double), ctypes.byref(params))
        if not self.h:
            raise RuntimeError("oracle: factorisation failed")
        del dp

    def close(self):
        if self.h:
            self.L.orc_pd_destroy(self.fp64, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1) -> tuple[int, int]:
        pcg = ctypes.c_int(0)
        it = self.L.orc_pd_step(self.fp64, self.h, ctypes.byref(self.params), n, ctypes.byref(pcg))
        return it, pcg.value

    def set_more_dbc(self, weights, targets):
        """moreDBC weights (nv) and the targets of the weighted vertices (nv, 3)."""
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1)
        self.L.orc_pd_set_more_dbc(self.fp64, self.h, _p(w, ctypes.c_double), _p(t, ctypes.c_double))

    def get(self, which: str = "X") -> np.ndarray:
        out = np.zeros(3 * self.nv)
        self.L.orc_pd_get(self.fp64, self.h, {"X": 0, "V": 1, "XTilde": 2}[which], _p(out, ctypes.c_double))
        return out.reshape(-1, 3)

    def export(self) -> dict:
        nnz = self.L.orc_pd_nnz(self.fp64, self.h)
        d = dict(dminv=np.zeros(9 * self.nt), V0=np.zeros(self.nt), diagL=np.zeros(self.nv), c=np.zeros(self.nv),
                 rowptr=np.zeros(self.nv + 1, np.int32), col=np.zeros(nnz, np.int32), val=np.zeros(nnz))
        cd, ci = ctypes.c_double, ctypes.c_int
        self.L.orc_pd_export(self.fp64, self.h, _p(d["dminv"], cd), _p(d["V0"], cd), _p(d["diagL"], cd),
                             _p(d["c"], cd), _p(d["rowptr"], ci), _p(d["col"], ci), _p(d["val"], cd))
        d["dminv"] = d["dminv"].reshape(-1, 3, 3)
        return d

    def local(self, q: np.ndarray, sn_old: np.ndarray, jacobi: bool, dt: float, positional_weight: float = 1e6):
        q = np.ascontiguousarray(q, dtype=np.float64).ravel()
        s = np.ascontiguousarray(sn_old, dtype=np.float64).ravel()
        out = np.zeros(3 * self.nv)
        cd = ctypes.c_double
        self.L.orc_pd_local(self.fp64, self.h, _p(q, cd), _p(s, cd), int(jacobi), dt, positional_weight,
                            _p(out, cd))
        return out.reshape(-1, 3)


def num_threads(bench: bool = False) -> int:
    return int(lib(bench).orc_num_threads())


# ------------------------------------------------------------- collision --
QUERY_VF = 1
QUERY_EE = 2


def _vec(a, fp64):
    return np.ascontiguousarray(a, dtype=np.float64 if fp64 else np.float32).reshape(-1)


def collision_detect(X, XT, tri, fathers, ignore_self=True, fp64=False, resolve=True):
    """DetectCollision on the CPU: sorted unique queries (n, 5) [type, v0..v3],
    their toi and normals, and (resolve=True) tI per vertex and per-vertex
    normals (float32) after storeTi."""
    L = lib()
    prec = 64 if fp64 else 32
    X = _vec(X, fp64)
    XT = _vec(XT, fp64)
    nv = X.size // 3
    tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1)
    fathers = np.ascontiguousarray(fathers, dtype=np.uint32).reshape(-1)
    tI = np.ones(nv, dtype=X.dtype) if resolve else None
    nrm_v = np.zeros(3 * nv, dtype=np.float32) if resolve else None
    h = L.orc_col_detect(prec, nv, tri.size // 3, _p(tri, ctypes.c_uint32), _p(fathers, ctypes.c_uint32),
                         X.ctypes.data, XT.ctypes.data, int(ignore_self),
                         tI.ctypes.data if resolve else None, _p(nrm_v, ctypes.c_float) if resolve else None)
    n = L.orc_col_num_queries(prec, h)
    q = np.zeros((n, 5), dtype=np.uint32)
    toi = np.zeros(n)
    nrm = np.zeros((n, 3))
    if n:
        L.orc_col_get(prec, h, _p(q, ctypes.c_uint32), _p(toi, ctypes.c_double), _p(nrm, ctypes.c_double))
    hits = L.orc_col_num_hits(prec, h)
    L.orc_col_destroy(prec, h)
    out = dict(queries=q, toi=toi, normals=nrm, hits=hits)
    if resolve:
        out["tI"] = tI
        out["vertex_normals"] = nrm_v.reshape(-1, 3)
    return out


def ccd_test(qtype, v, X, XT, fp64=True):
    """ccdCollisionTest of one query: (toi, normal)."""
    L = lib()
    X = _vec(X, fp64)
    XT = _vec(XT, fp64)
    v = np.ascontiguousarray(v, dtype=np.uint32)
    n = np.zeros(3)
    t = L.orc_ccd_test(64 if fp64 else 32, int(qtype), _p(v, ctypes.c_uint32), X.ctypes.data, XT.ctypes.data,
                       _p(n, ctypes.c_double))
    return t, n


def solve_cubic(a, b, c, d):
    x = np.zeros(3)
    k = lib().orc_solve_cubic(a, b, c, d, _p(x, ctypes.c_double))
    return x[:k]


def ccd_response(X, XT, V, tI, vertex_normals, fp64=False):
    """CCDKernel on the CPU; returns the new X and V."""
    X = _vec(X, fp64).copy()
    XT = _vec(XT, fp64)
    V = _vec(V, fp64).copy()
    tI = np.ascontiguousarray(tI, dtype=X.dtype)
    nrm = np.ascontiguousarray(vertex_normals, dtype=np.float32).reshape(-1)
    lib().orc_ccd_response(64 if fp64 else 32, X.size // 3, X.ctypes.data, XT.ctypes.data, V.ctypes.data,
                           tI.ctypes.data, _p(nrm, ctypes.c_float))
    return X.reshape(-1, 3), V.reshape(-1, 3)


# -------------------------------------------------------- fixed bodies --
def fixed_params(types, models_colmajor, radii):
    """Body parameters (8 floats each: translation, plane normal / cylinder
    axis, radius, 0) from column-major float model matrices."""
    t = np.ascontiguousarray(types, dtype=np.int32)
    m = np.ascontiguousarray(models_colmajor, dtype=np.float32).reshape(-1)
    r = np.ascontiguousarray(radii, dtype=np.float32)
    out = np.zeros((len(t), 8), np.float32)
    rc = lib().orc_fixed_params(len(t), _p(t, ctypes.c_int), _p(m, ctypes.c_float), _p(r, ctypes.c_float),
                                _p(out, ctypes.c_float))
    if rc != 0:
        raise ValueError("fixed body: unknown type or singular plane matrix")
    return out


def fixed_handle(X, V, types, params, mu_t, mu_n, fp64=False):
    """HandleCollisions on the CPU; returns the new X and V."""
    X = _vec(X, fp64).copy()
    V = _vec(V, fp64).copy()
    t = np.ascontiguousarray(types, dtype=np.int32)
    prm = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
    lib().orc_fixed_handle(64 if fp64 else 32, X.size // 3, X.ctypes.data, V.ctypes.data, len(t),
                           _p(t, ctypes.c_int), _p(prm, ctypes.c_float), float(mu_t), float(mu_n))
    return X.reshape(-1, 3), V.reshape(-1, 3)


# ----------------------------------------------------------------- IPC --
QUERY_TYPES = {"VF": 1, "EE": 2}
DISTANCE_TYPES = ["P_T0", "P_T1", "P_T2", "P_E0", "P_E1", "P_E2", "P_T", "EA0_EB0", "EA0_EB1", "EA1_EB0",
                  "EA1_EB1", "EA_EB0", "EA_EB1", "EA0_EB", "EA1_EB", "EA_EB"]


def update_queries(X, tri, fathers, dhat, ignore_self=False):
    """UpdateQueries (broadphase.cu:574-587) on the CPU: all-pairs candidates
    whose boxes grown by 2 dhat overlap, their distance types and squared
    distances, kept while d <= dhat; (queries (n, 5), types (n,), d (n,))
    sorted by (d, type, vertices)."""
    L = lib()
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
    tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1)
    fathers = np.ascontiguousarray(fathers, dtype=np.uint32).reshape(-1)
    n = L.orc_col_proximity(tri.size // 3, _p(tri, ctypes.c_uint32), _p(fathers, ctypes.c_uint32),
                            _p(X, ctypes.c_double), 2 * dhat, int(ignore_self), None)
    raw = np.zeros((n, 5), np.uint32)
    if n:
        L.orc_col_proximity(tri.size // 3, _p(tri, ctypes.c_uint32), _p(fathers, ctypes.c_uint32),
                            _p(X, ctypes.c_double), 2 * dhat, int(ignore_self), _p(raw, ctypes.c_uint32))
    P = X.reshape(-1, 3)
    rows = []
    for q in raw:
        x12 = P[q[1:5]].reshape(12)
        t = ipc_distance_type(int(q[0]), x12)
        d = ipc_distance(int(q[0]), t, x12)[0]
        if d <= dhat:
            rows.append((d, t, *[int(v) for v in q]))
    rows.sort()
    q = np.array([r[2:] for r in rows], dtype=np.uint32).reshape(-1, 5)
    return q, np.array([r[1] for r in rows], dtype=np.int32), np.array([r[0] for r in rows])


def ipc_distance_type(qtype, x12):
    x = np.ascontiguousarray(x12, dtype=np.float64).reshape(12)
    return lib().orc_ipc_distance_type(int(qtype), _p(x, ctypes.c_double))


def ipc_distance(qtype, dtype, x12):
    """Squared distance of the given type, its gradient (12) and Hessian (12x12)."""
    x = np.ascontiguousarray(x12, dtype=np.float64).reshape(12)
    g = np.zeros(12)
    H = np.zeros((12, 12))
    d = lib().orc_ipc_distance(int(qtype), int(dtype), _p(x, ctypes.c_double), _p(g, ctypes.c_double),
                               _p(H, ctypes.c_double))
    return d, g, H


def ipc_barrier(d, dhat, kappa):
    d1, d2 = ctypes.c_double(), ctypes.c_double()
    b = lib().orc_ipc_barrier(d, dhat, kappa, ctypes.byref(d1), ctypes.byref(d2))
    return b, d1.value, d2.value


def ipc_make_pd(M):
    A = np.ascontiguousarray(M, dtype=np.float64).copy()
    lib().orc_ipc_make_pd(A.shape[0], _p(A, ctypes.c_double))
    return A


def nh_density(F, mu, lam):
    """NeoHookean08 energy density, first Piola stress P (3x3) and the 9x9
    d2Psi/dF2 that GradHessianKern assembles (vec index 3c + r)."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    P = np.zeros((3, 3))
    D = np.zeros((9, 9))
    e = lib().orc_nh_density(_p(F, ctypes.c_double), float(mu), float(lam), _p(P, ctypes.c_double),
                             _p(D, ctypes.c_double))
    return e, P, D


class IpcOracle:
    """The reference's IPCSolver step (ipc.cu:142-258) on the CPU, FP64."""

    def __init__(self, X, V, tet, tri, fathers, mass, mu, lam, fixed, dt=0.01, damp=0.999, dhat=1e-3, tol=5e-2,
                 lin_tol=1e-13, max_iter=100, elastic="corotated"):
        self.L = lib()
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        self.nv = X.size // 3
        tet = np.ascontiguousarray(tet, dtype=np.uint32).reshape(-1)
        tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1)
        fathers = np.ascontiguousarray(fathers, dtype=np.uint32).reshape(-1)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (mass, mu, lam)]
        fx = np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        self._keep += [X, V, tet, tri, fathers, fx]
        cd, cu = ctypes.c_double, ctypes.c_uint32
        self.h = self.L.orc_ipc_create(self.nv, tet.size // 4, tri.size // 3, _p(X, cd), _p(V, cd), _p(tet, cu),
                                       _p(tri, cu), _p(fathers, cu), _p(self._keep[0], cd), _p(self._keep[1], cd),
                                       _p(self._keep[2], cd), _p(fx, ctypes.c_ubyte), dt, damp, dhat, tol, lin_tol,
                                       int(max_iter))
        self.L.orc_ipc_set_elastic(self.h, {"corotated": 0, "neohookean": 1}[elastic])

    def step(self, n=1):
        ok = True
        for _ in range(n):
            ok = bool(self.L.orc_ipc_step(self.h)) and ok
        return ok

    def get(self):
        X = np.zeros(3 * self.nv)
        V = np.zeros(3 * self.nv)
        self.L.orc_ipc_get(self.h, _p(X, ctypes.c_double), _p(V, ctypes.c_double))
        return X.reshape(-1, 3), V.reshape(-1, 3)

    def stats(self):
        ints = np.zeros(4, np.int32)
        reals = np.zeros(2)
        self.L.orc_ipc_stats(self.h, _p(ints, ctypes.c_int), _p(reals, ctypes.c_double))
        return dict(newton_iterations=int(ints[0]), line_search_halvings=int(ints[1]), linear_iterations=int(ints[2]),
                    queries=int(ints[3]), kappa=float(reals[0]), alpha0=float(reals[1]))

    def set_xv(self, X, V, ignore_self=True):
        """Continue from another run's positions and velocities (the detector's
        ignore-self flag as a line search leaves it)."""
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        self.L.orc_ipc_set_xv(self.h, _p(X, ctypes.c_double), _p(V, ctypes.c_double), int(ignore_self))

    def set_state(self, XT, kappa):
        XT = np.ascontiguousarray(XT, dtype=np.float64).reshape(-1)
        self.L.orc_ipc_set_state(self.h, _p(XT, ctypes.c_double), float(kappa))

    def energy(self, X, grad=False):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        g = np.zeros(X.size) if grad else None
        e = self.L.orc_ipc_energy(self.h, _p(X, ctypes.c_double), None if g is None else _p(g, ctypes.c_double))
        return (e, g.reshape(-1, 3)) if grad else e

    def set_fixed(self, types, params):
        """Fixed bodies of the implicit barrier: types (0 sphere, 1 plane, 2
        cylinder) and 8 floats each (the library's / fixed_params' layout)."""
        t = np.ascontiguousarray(types, dtype=np.int32)
        prm = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        self._keep += [t, prm]
        self.L.orc_ipc_set_fixed(self.h, len(t), _p(t, ctypes.c_int), _p(prm, ctypes.c_float))

    def contact_area(self):
        out = np.zeros(self.nv)
        self.L.orc_ipc_contact_area(self.h, _p(out, ctypes.c_double))
        return out

    def fixed_vertex(self, X, v):
        """Fixed-body barrier of vertex v (no h^2): energy, gradient (3), Hessian (3x3)."""
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        g, H = np.zeros(3), np.zeros(9)
        e = self.L.orc_ipc_fixed_vertex(self.h, _p(X, ctypes.c_double), int(v), _p(g, ctypes.c_double),
                                        _p(H, ctypes.c_double))
        return e, g, H.reshape(3, 3)

    def fixed_step(self, X, p):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        return self.L.orc_ipc_fixed_step(self.h, _p(X, ctypes.c_double), _p(p, ctypes.c_double))

    def queries(self):
        n = self.L.orc_ipc_queries(self.h, None, None, None)
        q = np.zeros((n, 5), np.uint32)
        dt = np.zeros(n, np.int32)
        d = np.zeros(n)
        if n:
            self.L.orc_ipc_queries(self.h, _p(q, ctypes.c_uint32), _p(dt, ctypes.c_int), _p(d, ctypes.c_double))
        return q, dt, d

    def close(self):
        if self.h:
            self.L.orc_ipc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
