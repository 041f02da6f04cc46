"""Python mirror of the reference's general linear-solver interface.

LinearSolver<T>::Solve(N, d_b, d_x, A, nz, rowIdx, colIdx, d_guess)
(src/simulation/solver/linear/linear.h:52-70) with the PCGJacobiSolver
algorithm (pcgJacobi.h:7-10, pcgJacobi.cu:89-173).  Arguments are device
pointers (ints) or DeviceArray objects, as the reference's are device
pointers; the COO triplets may be unsorted and repeat entries (summed, as
sort_coo does, linear.cu:7-40).
"""
from __future__ import annotations

import ctypes

from . import capi


def _ptr(a):
    if a is None:
        return None
    return a.ptr if hasattr(a, "ptr") else int(a)


class PCGJacobiSolver:
    """PCGJacobiSolver<T>(N, max_iter = 2000, tolerance = 1e-5) on the GPU."""

    def __init__(self, N: int = 0, max_iter: int = 2000, tolerance: float = 1e-5, precision: int = 64):
        self._L = capi.lib()
        self.precision = precision
        h = ctypes.c_void_p()
        capi.check(self._L.sbs_linear_create(int(precision), int(max_iter), float(tolerance), ctypes.byref(h)))
        self.h = h.value
        self.iterations = 0

    def Solve(self, N: int, d_b, d_x, A=None, nz: int = 0, rowIdx=None, colIdx=None, d_guess=None) -> int:
        it = ctypes.c_int()
        capi.check(self._L.sbs_linear_solve(self.h, int(N), _ptr(d_b), _ptr(d_x), _ptr(A), int(nz), _ptr(rowIdx),
                                            _ptr(colIdx), _ptr(d_guess), ctypes.byref(it)))
        self.iterations = it.value
        return it.value

    def SolveNoAlloc(self, N: int, d_b, d_x, A=None, nz: int = 0, rowIdx=None, colIdx=None, d_guess=None) -> int:
        """Solve without device allocation (a fixed 32 MiB workspace per GPU)."""
        it = ctypes.c_int()
        capi.check(self._L.sbs_linear_solve_noalloc(self.h, int(N), _ptr(d_b), _ptr(d_x), _ptr(A), int(nz),
                                                    _ptr(rowIdx), _ptr(colIdx), _ptr(d_guess), ctypes.byref(it)))
        self.iterations = it.value
        return it.value

    def set_stream(self, stream) -> None:
        capi.check(self._L.sbs_linear_set_stream(self.h, _ptr(stream)))

    def nnz(self) -> int:
        n = ctypes.c_longlong()
        capi.check(self._L.sbs_linear_stats(self.h, ctypes.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "h", None):
            self._L.sbs_linear_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
