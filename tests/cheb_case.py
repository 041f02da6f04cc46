"""A PD scene on which the Jacobi + Chebyshev global step is a LINEAR iteration
with known contraction factors, and the closed form of its error.

Scene: one right-corner tet, vertex 0 at the origin and vertices 1, 2, 3 on
the x, y, z axes (Dm = I).  Vertices 0, 1, 2 are pinned, vertex 3 is free;
gravity is 0.  Every iterate keeps vertex 0 at the origin, vertex 1 on the x
axis, vertex 2 on the y axis and vertex 3 on the z axis, so F = diag(x1, y2, z3)
with x1 > y2 > z3 > 0 throughout, whose rotation R is exactly I.  The local
step's force w (R - F) Dm^-T G^T is then affine in the positions, and one
Jacobi update (pdUtil.cu:195-214) of each moving coordinate is an affine map
of that coordinate alone:

* a pinned coordinate (row overwritten with DBCX * 1e6/h^2, pdUtil.cu:147-166)
  contracts by g = diagL / (c + diagL) towards DBCX * (1e6/h^2) / c;
* the free z3 has g = 0: it jumps to z* = (c sn_old + w) / (c + w).

With the damping 0.9 (x - s) + s and Chebyshev extrapolation (pdUtil.cu:216-226,
pdSolver.cu:196-198: omega = 1 for i <= 10, 2/(2 - rho^2) at i = 11,
4/(4 - rho^2 omega) after), each coordinate's error after n iterations is
e_n = P_n(g') e_0 with g' = 0.9 g + 0.1 and

    P_n(x) = x^n                                   n <= 10,
    P_n(x) = x^10 T_{n-10}(x/rho) / T_{n-10}(1/rho)  n >= 10,

the Chebyshev semi-iteration started from the tenth iterate (Golub & Varga
1961; Wang 2015 for its use in projective dynamics), evaluated here through
cos/acos and cosh/acosh, not through the recurrence.
"""
from __future__ import annotations

import math

import numpy as np

from sbs_amd import mesh

DT = 1.0 / 60.0
RHO = 0.9992
W = 3.6e9                       # w = |V0| mu of the tet (V0 = 1/6)
MASS = (1.0, 1.0e4, 2.0e4, W * DT * DT)   # c of vertex 3 equals w: z* = (sn_old + 1) / 2
SHIFT = (0.3, 0.15, -0.5)       # displacement of vertex 1 (x), 2 (y), 3 (z) in the first prediction


def scene() -> mesh.Scene:
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    return mesh.Scene(X=X, tet=np.array([[0, 1, 2, 3]], dtype=np.uint32), tri=np.zeros((0, 3), np.uint32),
                      tri_fathers=np.zeros(0, np.uint32), mass=np.array(MASS), mu=np.array([6.0 * W]),
                      lam=np.zeros(1), dbc=np.array([1.0, 1.0, 1.0, 0.0]),
                      body_vert_start=np.array([0, 4]), body_tet_start=np.array([0, 1]))


def velocity() -> np.ndarray:
    """Initial velocities that put the first prediction s = X + h V at SHIFT."""
    V = np.zeros((4, 3))
    V[1, 0], V[2, 1], V[3, 2] = SHIFT[0] / DT, SHIFT[1] / DT, SHIFT[2] / DT
    return V


def chebyshev_t(m: int, x: float) -> float:
    """T_m(x) in closed form: cos(m acos x) on [-1, 1], cosh(m acosh |x|) outside."""
    if abs(x) <= 1.0:
        return math.cos(m * math.acos(x))
    return math.copysign(1.0, x) ** m * math.cosh(m * math.acosh(abs(x)))


def poly(n: int, x: float, rho: float = RHO, switch: int = 10, damping: float = 0.9) -> float:
    """P_n(g') for the base factor g (x = g): the error multiplier after n
    iterations.  `switch` and `damping` exist so a test can show that a wrong
    switch index or damping gives a different answer."""
    gp = damping * x + (1.0 - damping)
    if n <= switch:
        return gp ** n
    return gp ** switch * chebyshev_t(n - switch, gp / rho) / chebyshev_t(n - switch, 1.0 / rho)


def modes(c: np.ndarray, diagL: np.ndarray, weight: float):
    """(vertex, axis, g, fixed point, first iterate) of the three moving
    coordinates, from the assembled c and diagL and the pin weight 1e6/h^2."""
    out = []
    for v, ax in ((1, 0), (2, 1)):
        g = diagL[v] / (c[v] + diagL[v])
        out.append((v, ax, g, 1.0 * weight / c[v], 1.0 + SHIFT[v - 1]))
    s3 = 1.0 + SHIFT[2]
    out.append((3, 2, 0.0, (c[3] * s3 + W) / (c[3] + W), s3))
    return out
