"""Python mirror of the reference's PD solver interface over the C ABI.

Reference shapes kept (file:line in the reference):
  SolverData<Scalar>          src/def.h:22-56           -> SolverData
  SolverParams<Scalar>        src/def.h:83-100          -> SolverParams
  PdSolver                    src/simulation/solver/projective/pdSolver.h:12-44 -> PdSolver
    SetGlobalSolver(SolverType) / Update / Reset / SetPerf / GetPerformanceData
  LinearSolver<T>::Solve      src/simulation/solver/linear/linear.h:62 -> PdSolver.solve
"""
from __future__ import annotations

import ctypes
import dataclasses
import enum

import numpy as np

from . import capi


class SolverType(enum.IntEnum):
    """PdSolver::SolverType (pdSolver.h:14-17) plus the Jacobi-PCG global step."""
    Jacobi = capi.SOLVER_JACOBI
    CuSolverCholesky = capi.SOLVER_CUSOLVER_CHOLESKY
    EigenCholesky = capi.SOLVER_EIGEN_CHOLESKY
    PCG = capi.SOLVER_PCG
    DenseCholesky = capi.SOLVER_DENSE_CHOLESKY
    BatchedPCG = capi.SOLVER_BATCHED_PCG


@dataclasses.dataclass
class SolverParams:
    """SolverParams<Scalar> fields used by the PD step (def.h:83-100)."""
    numIterations: int = 1
    dt: float = 0.001
    gravity: float = 9.8
    tol: float = 1e-2
    rho: float = 0.9992
    handleCollision: bool = False
    pcg_tol: float = 1e-6
    pcg_floor: float = 1e-8
    pcg_max_iter: int = 100
    pcg_chebyshev: bool = False
    pcg_adaptive: bool = True
    dense_precision: int = 0      # 0: FP32 MFMA inverse, 1: BF16 MFMA inverse
    dense_tol: float = 1e-10
    dense_max_iter: int = 8
    pcg_halo: bool = True         # halo-staged PCG kernels (LDS) where the mesh can be tiled
    muN: float = 0.5              # fixed-body response: normal restitution
    muT: float = 0.5              # fixed-body response: friction
    use_graph: bool = False       # replay the step as a captured HIP graph (needs a non-null stream)
    batch_autotune: bool = True   # BatchedPCG: time each size class's workgroup sizes at the first update
    polar_rotation: bool = False  # FP32 local step: Newton polar iteration instead of the TR1690 SVD

    def to_c(self) -> capi.SbsSolverParams:
        return capi.SbsSolverParams(int(self.numIterations), float(self.dt), float(self.gravity), float(self.tol),
                                    float(self.rho), int(self.handleCollision), float(self.pcg_tol),
                                    float(self.pcg_floor), int(self.pcg_max_iter), int(self.pcg_chebyshev),
                                    int(self.pcg_adaptive), int(self.dense_precision), float(self.dense_tol),
                                    int(self.dense_max_iter), int(self.pcg_halo), float(self.muN),
                                    float(self.muT), int(self.use_graph), int(self.batch_autotune),
                                    int(self.polar_rotation))


class SolverData:
    """Device arrays of one flat scene, allocated through the library.

    Vectors are interleaved xyz (3*N_v scalars), tets u32[4*N_t]; scalars are
    float32 (precision 32) or float64 (precision 64).
    """

    def __init__(self, scene, precision: int = 32, V0=None):
        self.precision = precision
        self.dtype = np.float32 if precision == 32 else np.float64
        dt = self.dtype
        self.nv, self.nt = scene.nv, scene.nt
        X = np.ascontiguousarray(scene.X, dtype=dt).reshape(-1)
        V = np.zeros_like(X) if V0 is None else np.ascontiguousarray(V0, dtype=dt).reshape(-1)
        self._bufs = {}

        def put(name, arr):
            self._bufs[name] = capi.DeviceArray.from_numpy(arr)
            return self._bufs[name].ptr

        self.c = capi.SbsSolverData()
        self.c.num_verts, self.c.num_tets = self.nv, self.nt
        self.c.num_dbc = int(np.count_nonzero(scene.dbc))
        self.c.X = put("X", X)
        self.c.V = put("V", V)
        self.c.X0 = put("X0", X)
        self.c.XTilde = put("XTilde", X)
        self.c.DBCX = put("DBCX", X)
        self.c.Tet = put("Tet", np.ascontiguousarray(scene.tet, dtype=np.uint32).reshape(-1))
        self.c.mass = put("mass", np.ascontiguousarray(scene.mass, dtype=dt))
        self.c.mu = put("mu", np.ascontiguousarray(scene.mu, dtype=dt))
        self.c.DBC = put("DBC", np.ascontiguousarray(scene.dbc, dtype=dt))
        self.c.moreDBC = None
        self.c.V0 = put("V0", np.zeros(self.nt, dtype=dt))
        # surface for collision handling (dev_TriFathers, dev_tIs, dev_Normals)
        tri = getattr(scene, "tri", None)
        if tri is not None and len(tri):
            self.c.num_tris = int(np.asarray(tri).reshape(-1, 3).shape[0])
            self.c.Tri = put("Tri", np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1))
            self.c.TriFathers = put("TriFathers", np.ascontiguousarray(scene.tri_fathers, dtype=np.uint32))
        self.c.tIs = put("tIs", np.ones(self.nv, dtype=dt))
        self.c.Normals = put("Normals", np.zeros(3 * self.nv, dtype=np.float32))

    def set_more_dbc(self, weights, targets=None) -> None:
        """SolverData::moreDBC (def.h:32), the mouse-drag weights, with the
        dragged vertices' targets written into DBCX (computeSn, pdUtil.cu:80-87,
        writes the target there); None clears the weights.  The first update
        sets DBCX from X0, so targets are written after it."""
        if weights is None:
            self.c.moreDBC = None
            return
        w = np.ascontiguousarray(weights, dtype=self.dtype).reshape(-1)
        assert w.size == self.nv
        if "moreDBC" in self._bufs:
            self._bufs.pop("moreDBC").free()
        self._bufs["moreDBC"] = capi.DeviceArray.from_numpy(w)
        self.c.moreDBC = self._bufs["moreDBC"].ptr
        if targets is not None:
            d = self.get("DBCX")
            sel = w > 0
            d[sel] = np.asarray(targets, dtype=np.float64).reshape(-1, 3)[sel]
            self._bufs["DBCX"].upload(np.ascontiguousarray(d, dtype=self.dtype).reshape(-1))

    def set_fixed_bodies(self, fixed) -> None:
        """SolverData::pFixedBodies (def.h:43): a FixedBodyData, or None."""
        self.fixed_bodies = fixed
        self.c.fixed_bodies = fixed.h if fixed is not None else None

    def get(self, name: str) -> np.ndarray:
        if name in ("X", "V", "X0", "XTilde", "DBCX"):
            return self._bufs[name].download(self.dtype, (self.nv, 3))
        if name in ("mass", "DBC", "tIs"):
            return self._bufs[name].download(self.dtype, (self.nv,))
        if name == "Normals":
            return self._bufs[name].download(np.float32, (self.nv, 3))
        if name in ("mu", "V0"):
            return self._bufs[name].download(self.dtype, (self.nt,))
        raise KeyError(name)

    def ptr(self, name: str) -> int:
        return self._bufs[name].ptr

    def free(self):
        for b in self._bufs.values():
            b.free()
        self._bufs.clear()


class PdSolver:
    """PdSolver (pdSolver.h:12-44) over the gfx950 library."""

    def __init__(self, threadsPerBlock: int, solverData: SolverData):
        self._L = capi.lib()
        self.data = solverData
        self.precision = solverData.precision
        h = ctypes.c_void_p()
        capi.check(self._L.sbs_pd_create(self.precision, int(threadsPerBlock), ctypes.byref(solverData.c),
                                         ctypes.byref(h)))
        self.h = h.value

    def close(self):
        if getattr(self, "h", None):
            self._L.sbs_pd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def SetGlobalSolver(self, val: SolverType) -> None:
        capi.check(self._L.sbs_pd_set_global_solver(self.h, int(val)))

    def Update(self, solverData: SolverData, solverParams: SolverParams) -> None:
        p = solverParams.to_c()
        capi.check(self._L.sbs_pd_update(self.h, ctypes.byref(solverData.c), ctypes.byref(p)))

    def Reset(self) -> None:
        capi.check(self._L.sbs_pd_reset(self.h))

    def SetPerf(self, val: bool) -> None:
        capi.check(self._L.sbs_pd_set_perf(self.h, int(bool(val))))

    def GetPerformanceData(self) -> list[tuple[str, float]]:
        n = ctypes.c_int(0)
        capi.check(self._L.sbs_pd_perf_count(self.h, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            name = ctypes.c_char_p()
            ms = ctypes.c_double()
            capi.check(self._L.sbs_pd_perf_entry(self.h, i, ctypes.byref(name), ctypes.byref(ms)))
            out.append((name.value.decode(), ms.value))
        return out

    def stats(self) -> capi.SbsPdStats:
        s = capi.SbsPdStats()
        capi.check(self._L.sbs_pd_stats(self.h, ctypes.byref(s)))
        return s

    def solve(self, d_b: int, d_x: int, params: SolverParams, d_guess: int | None = None) -> int:
        it = ctypes.c_int(0)
        p = params.to_c()
        capi.check(self._L.sbs_pd_solve(self.h, d_b, d_x, d_guess, ctypes.byref(p), ctypes.byref(it)))
        return it.value

    def local_step(self, d_q: int, d_sn_old: int, residual_form: bool, dt: float, d_b: int) -> None:
        capi.check(self._L.sbs_pd_local_step(self.h, d_q, d_sn_old, int(residual_form), float(dt), d_b))

    def dense_info(self) -> dict:
        """DenseCholesky mode: bodies, largest body, inverse bytes, and the last
        solve's refinement corrections and residual ratio."""
        b, m, it = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nb, r = ctypes.c_longlong(), ctypes.c_double()
        capi.check(self._L.sbs_pd_dense_info(self.h, ctypes.byref(b), ctypes.byref(m), ctypes.byref(nb),
                                             ctypes.byref(it), ctypes.byref(r)))
        return dict(bodies=b.value, max_body=m.value, inverse_bytes=nb.value, iterations=it.value, ratio=r.value)

    def batch_info(self) -> dict:
        """BatchedPCG: bodies, size classes (tets range, bodies, tuned threads
        per workgroup, the tuner's (threads, ms) trials) and per-body PCG
        statistics (accumulated over the updates)."""
        nb, nc = ctypes.c_int(), ctypes.c_int()
        capi.check(self._L.sbs_pd_batch_info(self.h, ctypes.byref(nb), ctypes.byref(nc)))
        classes = []
        for c in range(nc.value):
            info = (ctypes.c_int * 4)()
            tried = (ctypes.c_double * 16)()
            nt = ctypes.c_int()
            capi.check(self._L.sbs_pd_batch_class(self.h, c, info, tried, 8, ctypes.byref(nt)))
            classes.append(dict(min_tets=info[0], max_tets=info[1], bodies=info[2], threads=info[3],
                                tried=[(int(tried[2 * k]), tried[2 * k + 1]) for k in range(min(nt.value, 8))]))
        st = np.zeros((nb.value, 6), np.int32)
        capi.check(self._L.sbs_pd_batch_bodies(self.h, st.ctypes.data))
        return dict(bodies=nb.value, classes=classes, body_stats=st)

    def export_system(self) -> dict:
        dt = self.data.dtype
        nv, nt = self.data.nv, self.data.nt
        nnz = ctypes.c_longlong(0)
        capi.check(self._L.sbs_pd_system_nnz(self.h, ctypes.byref(nnz)))
        d = dict(c=np.zeros(nv, dt), diagA=np.zeros(nv, dt), diagL=np.zeros(nv, dt), V0=np.zeros(nt, dt),
                 w=np.zeros(nt, dt), dminv=np.zeros(9 * nt, dt), rowptr=np.zeros(nv + 1, np.int32),
                 col=np.zeros(nnz.value, np.int32), val=np.zeros(nnz.value, dt))
        a = {k: v.ctypes.data for k, v in d.items()}
        capi.check(self._L.sbs_pd_export_system(self.h, a["c"], a["diagA"], a["diagL"], a["V0"], a["w"],
                                                a["dminv"], a["rowptr"], a["col"], a["val"]))
        d["dminv"] = d["dminv"].reshape(nt, 3, 3)
        return d

    def set_kernel_timing(self, kernels) -> None:
        """Time every launch of the named kernels (capi.KERNELS) with HIP events."""
        mask = 0
        for k in kernels:
            mask |= 1 << capi.KERNELS.index(k)
        capi.check(self._L.sbs_pd_set_kernel_timing(self.h, mask))

    def kernel_time(self, kernel: str) -> tuple[int, float]:
        """(launches, summed milliseconds) of a timed kernel."""
        n = ctypes.c_int(0)
        ms = ctypes.c_double(0)
        capi.check(self._L.sbs_pd_kernel_time(self.h, capi.KERNELS.index(kernel), None, ctypes.byref(n),
                                              ctypes.byref(ms)))
        return n.value, ms.value

    def set_stream(self, stream) -> None:
        """Queue the solver's work on `stream` (a capi.Stream, a raw hipStream_t or None)."""
        ptr = getattr(stream, "ptr", stream)
        self._stream = stream
        capi.check(self._L.sbs_pd_set_stream(self.h, ptr))

    def synchronize(self) -> None:
        capi.check(self._L.sbs_device_synchronize())


class CollisionDetection:
    """CollisionDetection<Scalar> (src/collision/bvh.h:44-89) over the gfx950
    library: DetectCollision, ComputeMinStepSize, GetNumQueries/GetQueries.

    The triangle arrays are uploaded once; positions are numpy arrays (n, 3)
    uploaded per call, or device pointers (ints) used as they are."""

    def __init__(self, num_verts: int, tri, tri_fathers, precision: int = 32):
        self._L = capi.lib()
        self.precision = precision
        self.dtype = np.float32 if precision == 32 else np.float64
        self.nv = int(num_verts)
        tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1)
        self.ntri = tri.size // 3
        self._tri = capi.DeviceArray.from_numpy(tri)
        self._fat = capi.DeviceArray.from_numpy(np.ascontiguousarray(tri_fathers, dtype=np.uint32))
        self._x = capi.DeviceArray(3 * self.nv * np.dtype(self.dtype).itemsize)
        self._xt = capi.DeviceArray(3 * self.nv * np.dtype(self.dtype).itemsize)
        self._ti = capi.DeviceArray(self.nv * np.dtype(self.dtype).itemsize)
        self._n = capi.DeviceArray(3 * self.nv * 4)
        self._n.upload(np.zeros(3 * self.nv, np.float32))
        h = ctypes.c_void_p()
        capi.check(self._L.sbs_collision_create(precision, self.nv, self.ntri, self._tri.ptr, self._fat.ptr,
                                                ctypes.byref(h)))
        self.h = h.value

    def _pos(self, X, XT):
        self._x.upload(np.ascontiguousarray(X, dtype=self.dtype).reshape(-1))
        self._xt.upload(np.ascontiguousarray(XT, dtype=self.dtype).reshape(-1))

    def DetectCollision(self, X, XTilde, ignoreSelfCollision: bool = True):
        """Returns (tI (nv,), normals (nv, 3) float32)."""
        self._pos(X, XTilde)
        capi.check(self._L.sbs_collision_detect(self.h, self._x.ptr, self._xt.ptr, int(ignoreSelfCollision),
                                                self._ti.ptr, self._n.ptr))
        return self._ti.download(self.dtype, (self.nv,)), self._n.download(np.float32, (self.nv, 3))

    def ComputeMinStepSize(self, X, XTilde, ignoreSelfCollision: bool = False) -> float:
        self._pos(X, XTilde)
        t = ctypes.c_double()
        capi.check(self._L.sbs_collision_min_step(self.h, self._x.ptr, self._xt.ptr, int(ignoreSelfCollision),
                                                  ctypes.byref(t)))
        return t.value

    def UpdateQueries(self, X, dhat: float):
        """UpdateQueries(numVerts, numTris, Tri, X, TriFathers, dhat) (bvh.h:52,
        broadphase.cu:574-587), precision 64: (queries (n, 5) u32, distance
        types (n,), squared distances (n,)) of the kept queries, sorted by
        (distance, type)."""
        self._x.upload(np.ascontiguousarray(X, dtype=self.dtype).reshape(-1))
        n = ctypes.c_longlong()
        capi.check(self._L.sbs_collision_update_queries(self.h, self._x.ptr, float(dhat), ctypes.byref(n)))
        q = np.zeros((n.value, 5), np.uint32)
        t = np.zeros(n.value, np.int32)
        d = np.zeros(n.value)
        if n.value:
            capi.check(self._L.sbs_collision_get_proximity(self.h, q.ctypes.data, t.ctypes.data, d.ctypes.data))
        return q, t, d

    def SetBuildType(self, build_type: int) -> None:
        """BVH<Scalar>::BuildType: 0 Serial, 1 Atomic (default), 2 Cooperative."""
        capi.check(self._L.sbs_collision_set_build_type(self.h, int(build_type)))

    def GetBuildType(self) -> int:
        t = ctypes.c_int()
        capi.check(self._L.sbs_collision_get_build_type(self.h, ctypes.byref(t)))
        return t.value

    def GetNumQueries(self) -> int:
        n = ctypes.c_longlong()
        capi.check(self._L.sbs_collision_num_queries(self.h, ctypes.byref(n)))
        return n.value

    def GetQueries(self):
        """(queries (n, 5) u32 [type, v0..v3], toi (n,), normals (n, 3))."""
        n = self.GetNumQueries()
        q = np.zeros((n, 5), np.uint32)
        toi = np.zeros(n)
        nrm = np.zeros((n, 3))
        if n:
            capi.check(self._L.sbs_collision_get_queries(self.h, q.ctypes.data, toi.ctypes.data, nrm.ctypes.data))
        return q, toi, nrm

    def stats(self):
        p, h = ctypes.c_longlong(), ctypes.c_int()
        capi.check(self._L.sbs_collision_stats(self.h, ctypes.byref(p), ctypes.byref(h)))
        return p.value, h.value

    def close(self):
        if self.h:
            capi.check(self._L.sbs_collision_destroy(self.h))
            self.h = None
        for b in (self._tri, self._fat, self._x, self._xt, self._ti, self._n):
            b.free()


def ccd_response(X, XTilde, V, tI, normals, precision: int = 32):
    """CCDKernel (collisionUtil.cu:50-70) on the GPU; returns new (X, V)."""
    dt = np.float32 if precision == 32 else np.float64
    L = capi.lib()
    bufs = [capi.DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=t).reshape(-1))
            for a, t in ((X, dt), (XTilde, dt), (V, dt), (tI, dt), (normals, np.float32))]
    nv = np.asarray(X).reshape(-1, 3).shape[0]
    capi.check(L.sbs_ccd_response(precision, nv, *(b.ptr for b in bufs)))
    Xn = bufs[0].download(dt, (nv, 3))
    Vn = bufs[2].download(dt, (nv, 3))
    for b in bufs:
        b.free()
    return Xn, Vn
