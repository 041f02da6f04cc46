"""MI355X-native projective-dynamics step (host mirror of the reference interface).

Modules: mesh (synthetic hex-grid scenes), capi (ctypes binding of
include/sbs/sbs.h), pd (SolverData / SolverParams / PdSolver).
"""
from . import mesh  # noqa: F401
