"""ctypes binding of the C ABI in include/sbs/sbs.h (lib/libsbs_amd.so).

The library is required: importing works without it, but every call that
needs it raises LibraryNotBuilt -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsbs_amd.so")
REPO_ROOT = os.path.dirname(PKG_ROOT)
HEADER_PATH = os.path.join(REPO_ROOT, "include", "sbs", "sbs.h")

SOLVER_JACOBI = 0
SOLVER_CUSOLVER_CHOLESKY = 1
SOLVER_EIGEN_CHOLESKY = 2
SOLVER_PCG = 3
SOLVER_DENSE_CHOLESKY = 4
SOLVER_BATCHED_PCG = 5

KERNELS = ["local_step", "pcg_solve", "jacobi_chebyshev", "direct_solve", "rhs_gather", "pcg_iteration"]


class LibraryNotBuilt(RuntimeError):
    pass


class SbsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sbs error {code}: {msg}")
        self.code = code


class SbsSolverData(ctypes.Structure):
    _fields_ = [("num_verts", ctypes.c_int), ("num_tets", ctypes.c_int), ("num_dbc", ctypes.c_int),
                ("X", ctypes.c_void_p), ("V", ctypes.c_void_p), ("X0", ctypes.c_void_p),
                ("XTilde", ctypes.c_void_p), ("DBCX", ctypes.c_void_p), ("Tet", ctypes.c_void_p),
                ("mass", ctypes.c_void_p), ("mu", ctypes.c_void_p), ("DBC", ctypes.c_void_p),
                ("moreDBC", ctypes.c_void_p), ("V0", ctypes.c_void_p), ("num_tris", ctypes.c_int),
                ("Tri", ctypes.c_void_p), ("TriFathers", ctypes.c_void_p), ("tIs", ctypes.c_void_p),
                ("Normals", ctypes.c_void_p), ("fixed_bodies", ctypes.c_void_p)]


class SbsSolverParams(ctypes.Structure):
    _fields_ = [("num_iterations", ctypes.c_int), ("dt", ctypes.c_double), ("gravity", ctypes.c_double),
                ("tol", ctypes.c_double), ("rho", ctypes.c_double), ("handle_collision", ctypes.c_int),
                ("pcg_tol", ctypes.c_double), ("pcg_floor", ctypes.c_double), ("pcg_max_iter", ctypes.c_int),
                ("pcg_chebyshev", ctypes.c_int), ("pcg_adaptive", ctypes.c_int), ("dense_precision", ctypes.c_int),
                ("dense_tol", ctypes.c_double), ("dense_max_iter", ctypes.c_int), ("pcg_halo", ctypes.c_int),
                ("mu_n", ctypes.c_double), ("mu_t", ctypes.c_double), ("use_graph", ctypes.c_int),
                ("batch_autotune", ctypes.c_int), ("polar_rotation", ctypes.c_int)]


class SbsPdStats(ctypes.Structure):
    _fields_ = [("pcg_iterations_last_step", ctypes.c_int), ("pcg_cap", ctypes.c_int),
                ("pcg_cap_hits", ctypes.c_longlong), ("nnz_offdiag", ctypes.c_longlong),
                ("pcg_iterations_total", ctypes.c_longlong), ("pcg_solves_total", ctypes.c_longlong),
                ("halo_groups", ctypes.c_int), ("graph_captures", ctypes.c_longlong),
                ("graph_replays", ctypes.c_longlong), ("partial_entries", ctypes.c_longlong)]


FIXED_SPHERE, FIXED_PLANE, FIXED_CYLINDER = 0, 1, 2


class SbsFixedBody(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("model", ctypes.c_float * 16), ("radius", ctypes.c_float)]


class SbsTetgenOptions(ctypes.Structure):
    _fields_ = [("start_index", ctypes.c_int), ("centralize", ctypes.c_int), ("pos", ctypes.c_float * 3),
                ("scale", ctypes.c_float * 3), ("rot", ctypes.c_float * 3), ("precision", ctypes.c_int)]


class SbsIpcData(ctypes.Structure):
    _fields_ = [("num_verts", ctypes.c_int), ("num_tets", ctypes.c_int), ("num_tris", ctypes.c_int),
                ("X", ctypes.c_void_p), ("V", ctypes.c_void_p), ("XTilde", ctypes.c_void_p),
                ("Tet", ctypes.c_void_p), ("Tri", ctypes.c_void_p), ("TriFathers", ctypes.c_void_p),
                ("mass", ctypes.c_void_p), ("mu", ctypes.c_void_p), ("lambda_", ctypes.c_void_p),
                ("DBC", ctypes.c_void_p), ("fixed_bodies", ctypes.c_void_p)]


class SbsIpcParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("damp", ctypes.c_double), ("dhat", ctypes.c_double),
                ("tol", ctypes.c_double), ("max_iterations", ctypes.c_int), ("lin_tol", ctypes.c_double),
                ("lin_max_iter", ctypes.c_int)]


class SbsIpcStats(ctypes.Structure):
    _fields_ = [("newton_iterations", ctypes.c_int), ("line_search_halvings", ctypes.c_int),
                ("linear_iterations", ctypes.c_int), ("failed", ctypes.c_int), ("queries", ctypes.c_longlong),
                ("kappa", ctypes.c_double), ("alpha0", ctypes.c_double), ("energy", ctypes.c_double)]


# Every function the header declares, with its ctypes signature.
_VP, _CP, _I, _D, _SZ = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
SIGNATURES = {
    "sbs_last_error": (_CP, []),
    "sbs_version": (_I, []),
    "sbs_pd_create": (_I, [_I, _I, ctypes.POINTER(SbsSolverData), ctypes.POINTER(_VP)]),
    "sbs_pd_destroy": (_I, [_VP]),
    "sbs_pd_set_global_solver": (_I, [_VP, _I]),
    "sbs_pd_update": (_I, [_VP, ctypes.POINTER(SbsSolverData), ctypes.POINTER(SbsSolverParams)]),
    "sbs_pd_reset": (_I, [_VP]),
    "sbs_pd_set_perf": (_I, [_VP, _I]),
    "sbs_pd_perf_count": (_I, [_VP, ctypes.POINTER(_I)]),
    "sbs_pd_perf_entry": (_I, [_VP, _I, ctypes.POINTER(_CP), ctypes.POINTER(_D)]),
    "sbs_pd_set_stream": (_I, [_VP, _VP]),
    "sbs_pd_stats": (_I, [_VP, ctypes.POINTER(SbsPdStats)]),
    "sbs_pd_set_kernel_timing": (_I, [_VP, ctypes.c_uint]),
    "sbs_pd_kernel_time": (_I, [_VP, _I, ctypes.POINTER(_CP), ctypes.POINTER(_I), ctypes.POINTER(_D)]),
    "sbs_pd_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SbsSolverParams), ctypes.POINTER(_I)]),
    "sbs_pd_local_step": (_I, [_VP, _VP, _VP, _I, _D, _VP]),
    "sbs_pd_system_nnz": (_I, [_VP, ctypes.POINTER(ctypes.c_longlong)]),
    "sbs_pd_export_system": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "sbs_pd_dense_info": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_longlong),
                               ctypes.POINTER(_I), ctypes.POINTER(_D)]),
    "sbs_pd_batch_info": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "sbs_pd_batch_class": (_I, [_VP, _I, _VP, _VP, _I, ctypes.POINTER(_I)]),
    "sbs_pd_batch_bodies": (_I, [_VP, _VP]),
    "sbs_collision_create": (_I, [_I, _I, _I, _VP, _VP, ctypes.POINTER(_VP)]),
    "sbs_collision_destroy": (_I, [_VP]),
    "sbs_collision_set_stream": (_I, [_VP, _VP]),
    "sbs_collision_detect": (_I, [_VP, _VP, _VP, _I, _VP, _VP]),
    "sbs_collision_min_step": (_I, [_VP, _VP, _VP, _I, ctypes.POINTER(_D)]),
    "sbs_collision_num_queries": (_I, [_VP, ctypes.POINTER(ctypes.c_longlong)]),
    "sbs_collision_get_queries": (_I, [_VP, _VP, _VP, _VP]),
    "sbs_collision_stats": (_I, [_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_I)]),
    "sbs_ccd_response": (_I, [_I, _I, _VP, _VP, _VP, _VP, _VP]),
    "sbs_mesh_load_tetgen": (_I, [_CP, _CP, _CP, ctypes.POINTER(SbsTetgenOptions), ctypes.POINTER(_VP)]),
    "sbs_mesh_counts": (_I, [_VP] + [ctypes.POINTER(_I)] * 5),
    "sbs_mesh_copy": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "sbs_mesh_destroy": (_I, [_VP]),
    "sbs_fixed_create": (_I, [ctypes.POINTER(SbsFixedBody), _I, ctypes.POINTER(_VP)]),
    "sbs_fixed_destroy": (_I, [_VP]),
    "sbs_fixed_handle_collisions": (_I, [_VP, _I, _VP, _VP, _I, _D, _D, _VP]),
    "sbs_fixed_params": (_I, [_VP, _VP]),
    "sbs_collision_set_build_type": (_I, [_VP, _I]),
    "sbs_collision_get_build_type": (_I, [_VP, ctypes.POINTER(_I)]),
    "sbs_collision_update_queries": (_I, [_VP, _VP, _D, ctypes.POINTER(ctypes.c_longlong)]),
    "sbs_collision_get_proximity": (_I, [_VP, _VP, _VP, _VP]),
    "sbs_linear_create": (_I, [_I, _I, _D, ctypes.POINTER(_VP)]),
    "sbs_linear_destroy": (_I, [_VP]),
    "sbs_linear_set_stream": (_I, [_VP, _VP]),
    "sbs_linear_solve": (_I, [_VP, _I, _VP, _VP, _VP, _I, _VP, _VP, _VP, ctypes.POINTER(_I)]),
    "sbs_linear_stats": (_I, [_VP, ctypes.POINTER(ctypes.c_longlong)]),
    "sbs_linear_solve_noalloc": (_I, [_VP, _I, _VP, _VP, _VP, _I, _VP, _VP, _VP, ctypes.POINTER(_I)]),
    "sbs_ipc_create": (_I, [ctypes.POINTER(SbsIpcData), ctypes.POINTER(_VP)]),
    "sbs_ipc_destroy": (_I, [_VP]),
    "sbs_ipc_update": (_I, [_VP, ctypes.POINTER(SbsIpcData), ctypes.POINTER(SbsIpcParams), ctypes.POINTER(_I)]),
    "sbs_ipc_reset": (_I, [_VP]),
    "sbs_ipc_set_elastic": (_I, [_VP, _I]),
    "sbs_ipc_set_linear_solver": (_I, [_VP, _I]),
    "sbs_ipc_set_stream": (_I, [_VP, _VP]),
    "sbs_ipc_stats": (_I, [_VP, ctypes.POINTER(SbsIpcStats)]),
    "sbs_ipc_energy": (_I, [_VP, _VP, ctypes.POINTER(_D), _VP]),
    "sbs_ipc_set_state": (_I, [_VP, _VP, _D]),
    "sbs_malloc": (_I, [ctypes.POINTER(_VP), _SZ]),
    "sbs_free": (_I, [_VP]),
    "sbs_memcpy_htod": (_I, [_VP, _VP, _SZ]),
    "sbs_memcpy_dtoh": (_I, [_VP, _VP, _SZ]),
    "sbs_memcpy_dtod": (_I, [_VP, _VP, _SZ]),
    "sbs_memset": (_I, [_VP, _I, _SZ]),
    "sbs_device_synchronize": (_I, []),
    "sbs_device_count": (_I, [ctypes.POINTER(_I)]),
    "sbs_stream_create": (_I, [ctypes.POINTER(_VP)]),
    "sbs_stream_destroy": (_I, [_VP]),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryNotBuilt(f"{LIB_PATH} not found: run __graft_entry__.build() (or make -C "
                              f"{PKG_ROOT}) first; there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code: int) -> None:
    if code != 0:
        msg = lib().sbs_last_error()
        raise SbsError(code, msg.decode() if msg else "")


def device_count() -> int:
    n = ctypes.c_int(0)
    code = lib().sbs_device_count(ctypes.byref(n))
    return n.value if code == 0 else 0


class DeviceArray:
    """A device allocation made through the library (sbs_malloc / sbs_free)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().sbs_malloc(ctypes.byref(p), max(self.nbytes, 16)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            check(lib().sbs_memcpy_htod(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().sbs_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        check(lib().sbs_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().sbs_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream made through the library (sbs_stream_create)."""

    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().sbs_stream_create(ctypes.byref(p)))
        self.ptr = p.value

    def close(self):
        if getattr(self, "ptr", None):
            lib().sbs_stream_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
