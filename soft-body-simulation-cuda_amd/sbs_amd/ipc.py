"""Python mirror of the reference's IPC solver interface over the C ABI.

IPCSolver (src/simulation/solver/IPC/ipc.h:11-40): constructed from the FP64
SolverData of a scene, then Update(data, params) per time step, Reset() and
SetLinearSolver(type).  The arrays live on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import dataclasses
import enum

import numpy as np

from . import capi


class LinearSolverType(enum.IntEnum):
    """IPCSolver::SolverType (ipc.h:14-17)."""
    CuSolverCholesky = 0
    ICPCG = 1
    PCGJacobi = 2
    Jacobi = 3


@dataclasses.dataclass
class IpcParams:
    """The IPC fields of SolverParams<double> (def.h:83-100) plus the linear solve."""
    dt: float = 0.01
    damp: float = 0.999
    dhat: float = 1e-3
    tol: float = 5e-2
    maxIterations: int = 100
    lin_tol: float = 1e-13
    lin_max_iter: int = 20000

    def to_c(self) -> capi.SbsIpcParams:
        return capi.SbsIpcParams(float(self.dt), float(self.damp), float(self.dhat), float(self.tol),
                                 int(self.maxIterations), float(self.lin_tol), int(self.lin_max_iter))


class IpcData:
    """Device arrays (FP64) of a scene for the IPC solver."""

    def __init__(self, scene, V=None, fixed=None):
        self.nv, self.nt, self.ntri = scene.nv, scene.nt, scene.ntri
        self.fixed_bodies = fixed          # a fixed.FixedBodyData: the implicit barrier's bodies
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
        u32 = lambda a: np.ascontiguousarray(a, dtype=np.uint32).ravel()
        self.arrays = {
            "X": capi.DeviceArray.from_numpy(f64(scene.X)),
            "V": capi.DeviceArray.from_numpy(f64(np.zeros_like(scene.X) if V is None else V)),
            "XTilde": capi.DeviceArray.from_numpy(f64(scene.X)),
            "Tet": capi.DeviceArray.from_numpy(u32(scene.tet)),
            "Tri": capi.DeviceArray.from_numpy(u32(scene.tri)),
            "TriFathers": capi.DeviceArray.from_numpy(u32(scene.tri_fathers)),
            "mass": capi.DeviceArray.from_numpy(f64(scene.mass)),
            "mu": capi.DeviceArray.from_numpy(f64(scene.mu)),
            "lambda": capi.DeviceArray.from_numpy(f64(scene.lam)),
            "DBC": capi.DeviceArray.from_numpy(f64(scene.dbc)),
        }

    def c(self) -> capi.SbsIpcData:
        a = self.arrays
        return capi.SbsIpcData(self.nv, self.nt, self.ntri, a["X"].ptr, a["V"].ptr, a["XTilde"].ptr, a["Tet"].ptr,
                               a["Tri"].ptr, a["TriFathers"].ptr, a["mass"].ptr, a["mu"].ptr, a["lambda"].ptr,
                               a["DBC"].ptr, self.fixed_bodies.h if self.fixed_bodies is not None else None)

    def get(self, name: str) -> np.ndarray:
        return self.arrays[name].download(np.float64, (self.nv, 3))

    def free(self):
        for a in self.arrays.values():
            a.free()
        self.arrays = {}


class IpcSolver:
    """IPCSolver(threadsPerBlock, SolverData<double>) over the C ABI."""

    ELASTIC = {"corotated": 0, "neohookean": 1}

    def __init__(self, data: IpcData, elastic: str = "corotated"):
        """elastic: the IPEnergy elastic term, "corotated" (what ip.cu:22
        builds) or "neohookean" (NeoHookean08Energy, the swap there)."""
        self._L = capi.lib()
        self._data = data
        h = ctypes.c_void_p()
        cd = data.c()
        capi.check(self._L.sbs_ipc_create(ctypes.byref(cd), ctypes.byref(h)))
        self.h = h.value
        capi.check(self._L.sbs_ipc_set_elastic(self.h, self.ELASTIC[elastic]))

    def Update(self, data: IpcData, params: IpcParams) -> bool:
        cd, cp, ok = data.c(), params.to_c(), ctypes.c_int()
        capi.check(self._L.sbs_ipc_update(self.h, ctypes.byref(cd), ctypes.byref(cp), ctypes.byref(ok)))
        return bool(ok.value)

    def Reset(self):
        capi.check(self._L.sbs_ipc_reset(self.h))

    def SetLinearSolver(self, t: LinearSolverType):
        capi.check(self._L.sbs_ipc_set_linear_solver(self.h, int(t)))

    def stats(self) -> dict:
        s = capi.SbsIpcStats()
        capi.check(self._L.sbs_ipc_stats(self.h, ctypes.byref(s)))
        return dict(newton_iterations=s.newton_iterations, line_search_halvings=s.line_search_halvings,
                    linear_iterations=s.linear_iterations, failed=bool(s.failed), queries=s.queries,
                    kappa=s.kappa, alpha0=s.alpha0, energy=s.energy)

    def energy(self, X, grad=False):
        X = np.ascontiguousarray(X, dtype=np.float64).ravel()
        e = ctypes.c_double()
        g = np.zeros_like(X) if grad else None
        capi.check(self._L.sbs_ipc_energy(self.h, X.ctypes.data, ctypes.byref(e),
                                          None if g is None else g.ctypes.data))
        return (e.value, g.reshape(-1, 3)) if grad else e.value

    def set_state(self, XTilde, kappa: float):
        XT = np.ascontiguousarray(XTilde, dtype=np.float64).ravel()
        capi.check(self._L.sbs_ipc_set_state(self.h, XT.ctypes.data, float(kappa)))

    def close(self):
        if self.h:
            capi.check(self._L.sbs_ipc_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
