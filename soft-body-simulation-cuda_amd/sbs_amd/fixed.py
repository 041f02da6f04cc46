"""Fixed (kinematic) bodies over the gfx950 library.

Mirrors FixedBody / Sphere / Plane / Cylinder (src/collision/rigid/rigid.h:8-28,
sphere.cpp, plane.cpp, cylinder.cpp), their construction from a scene
description (ReadFixedBodies, src/context.cpp:222-316, with
utilityCore::modelMatrix, src/utilities.cpp:141-149) and FixedBodyData with
HandleCollisions (src/simulation/fixedBodyData.h:10-30, fixedBodyData.cu).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import capi

SPHERE, PLANE, CYLINDER = capi.FIXED_SPHERE, capi.FIXED_PLANE, capi.FIXED_CYLINDER
TYPE_NAMES = {"sphere": SPHERE, "plane": PLANE, "cylinder": CYLINDER}


def _rotation(deg: float, axis) -> np.ndarray:
    """4x4 rotation by `deg` degrees about the unit `axis` (glm::rotate's matrix)."""
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    x, y, z = axis
    R = np.eye(4)
    R[:3, :3] = [[c + (1 - c) * x * x, (1 - c) * x * y - s * z, (1 - c) * x * z + s * y],
                 [(1 - c) * y * x + s * z, c + (1 - c) * y * y, (1 - c) * y * z - s * x],
                 [(1 - c) * z * x - s * y, (1 - c) * z * y + s * x, c + (1 - c) * z * z]]
    return R


def model_matrix(pos, rot, scale) -> np.ndarray:
    """utilityCore::modelMatrix (utilities.cpp:141-149): T(pos) Rx Ry Rz S(scale),
    rotations in degrees; composed in double and rounded to float32 (4x4, math
    layout: M[row, col])."""
    T = np.eye(4)
    T[:3, 3] = pos
    S = np.diag([scale[0], scale[1], scale[2], 1.0])
    M = T @ _rotation(rot[0], (1, 0, 0)) @ _rotation(rot[1], (0, 1, 0)) @ _rotation(rot[2], (0, 0, 1)) @ S
    return M.astype(np.float32)


@dataclass
class FixedBody:
    """One fixed body: type, model matrix (4x4, M[row, col]) and radius."""
    type: int
    model: np.ndarray
    radius: float = 0.0
    name: str = ""

    def to_c(self) -> capi.SbsFixedBody:
        b = capi.SbsFixedBody()
        b.type = int(self.type)
        col_major = np.asarray(self.model, np.float32).T.reshape(-1)   # glm: model[4*col + row]
        for i in range(16):
            b.model[i] = float(col_major[i])
        b.radius = float(self.radius)
        return b


def Sphere(pos=(0, 0, 0), rot=(0, 0, 0), radius: float = 1.0, name: str = "") -> FixedBody:
    """Sphere(modelMatrix(pos, rot, (r, r, r)), r) (context.cpp:266-288)."""
    return FixedBody(SPHERE, model_matrix(pos, rot, (radius, radius, radius)), float(radius), name)


def Cylinder(pos=(0, 0, 0), rot=(0, 0, 0), scale=(1, 1, 1), name: str = "") -> FixedBody:
    """Cylinder(modelMatrix(pos, rot, (s0, s1, s0)), scale): radius s0, height s1,
    infinite for the response (context.cpp:290-305, cylinder.cpp:9)."""
    return FixedBody(CYLINDER, model_matrix(pos, rot, (scale[0], scale[1], scale[0])), float(scale[0]), name)


def Plane(pos=(0, 0, 0), rot=(0, 0, 0), scale=(1, 1, 1), name: str = "") -> FixedBody:
    """Plane(modelMatrix(pos, rot, scale)): the y = 0 plane of the model frame
    (context.cpp:306-308, plane.cpp:9)."""
    return FixedBody(PLANE, model_matrix(pos, rot, scale), 0.0, name)


class FixedBodyData:
    """FixedBodyData (fixedBodyData.h:10-30): the bodies on the device and
    HandleCollisions (fixedBodyData.cu:136-148)."""

    def __init__(self, bodies: list[FixedBody]):
        self._L = capi.lib()
        self.bodies = list(bodies)
        arr = (capi.SbsFixedBody * max(1, len(self.bodies)))(*[b.to_c() for b in self.bodies])
        h = ctypes.c_void_p()
        capi.check(self._L.sbs_fixed_create(arr, len(self.bodies), ctypes.byref(h)))
        self.h = h.value

    def HandleCollisions(self, d_X: int, d_V: int, numVerts: int, muT: float, muN: float,
                         precision: int = 32) -> None:
        capi.check(self._L.sbs_fixed_handle_collisions(self.h, int(precision), d_X, d_V, int(numVerts),
                                                       float(muT), float(muN), None))

    def params(self) -> np.ndarray:
        """Derived parameters per body (creation order): p[3], u[3], radius, 0."""
        out = np.zeros((max(1, len(self.bodies)), 8), np.float32)
        capi.check(self._L.sbs_fixed_params(self.h, out.ctypes.data))
        return out[:len(self.bodies)]

    def close(self):
        if getattr(self, "h", None):
            self._L.sbs_fixed_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
