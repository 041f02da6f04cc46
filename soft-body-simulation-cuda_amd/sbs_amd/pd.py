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

    def to_c(self) -> capi.SbsSolverParams:
        return capi.SbsSolverParams(int(self.numIterations), float(self.dt), float(self.gravity), float(self.tol),
                                    float(self.rho), int(self.handleCollision), float(self.pcg_tol),
                                    float(self.pcg_floor), int(self.pcg_max_iter), int(self.pcg_chebyshev),
                                    int(self.pcg_adaptive))


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

    def get(self, name: str) -> np.ndarray:
        if name in ("X", "V", "X0", "XTilde", "DBCX"):
            return self._bufs[name].download(self.dtype, (self.nv, 3))
        if name in ("mass", "DBC"):
            return self._bufs[name].download(self.dtype, (self.nv,))
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

    def synchronize(self) -> None:
        capi.check(self._L.sbs_device_synchronize())
