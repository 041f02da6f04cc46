"""Scene building from TetGen files over the library's loader.

Mirrors DataLoader (src/simulation/dataLoader.h:10-37): CollectData reads one
soft body (sbs_mesh_load_tetgen: loadNodeFile, loadEleFaceFile, CollectEdges
and the model transform, dataLoader.cu:39-237), AllocData concatenates the
bodies into one flat scene (dataLoader.cu:292-378): vertex offsets applied to
tets, triangles, edges and pinned indices, the triangle's body id
(dev_TriFathers), one mass per vertex and one mu / lambda per tet taken from
the body's SoftBodyAttribute (def.h:58-65).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .mesh import Scene


@dataclass
class SoftBodyAttribute:
    """SoftBodyAttribute (def.h:58-65): per-vertex mass, Lame parameters, pinned vertices."""
    mass: float = 1.0
    mu: float = 20000.0
    lam: float = 5000.0
    DBC: list = field(default_factory=list)


def load_tetgen(node_file: str, ele_file: str, face_file: str | None = None, pos=(0, 0, 0), scale=(1, 1, 1),
                rot=(0, 0, 0), centralize: bool = False, start_index: int = 0, precision: int = 32) -> dict:
    """One body from TetGen files: X (nv, 3) float64, tet (nt, 4), tri (ntri, 3),
    edges (ne, 2) as uint32, and whether the triangles came from the face file."""
    L = capi.lib()
    o = capi.SbsTetgenOptions()
    o.start_index, o.centralize, o.precision = int(start_index), int(bool(centralize)), int(precision)
    for k in range(3):
        o.pos[k], o.scale[k], o.rot[k] = float(pos[k]), float(scale[k]), float(rot[k])
    h = ctypes.c_void_p()
    enc = lambda s: os.fsencode(s) if s else None
    capi.check(L.sbs_mesh_load_tetgen(enc(node_file), enc(ele_file), enc(face_file), ctypes.byref(o),
                                      ctypes.byref(h)))
    try:
        n = [ctypes.c_int() for _ in range(5)]
        capi.check(L.sbs_mesh_counts(h, *[ctypes.byref(x) for x in n]))
        nv, nt, ntri, ne, from_file = (x.value for x in n)
        X = np.zeros((nv, 3), np.float64)
        tet = np.zeros((nt, 4), np.uint32)
        tri = np.zeros((ntri, 3), np.uint32)
        edges = np.zeros((ne, 2), np.uint32)
        capi.check(L.sbs_mesh_copy(h, X.ctypes.data, tet.ctypes.data, tri.ctypes.data, edges.ctypes.data))
    finally:
        L.sbs_mesh_destroy(h)
    return dict(X=X, tet=tet, tri=tri, edges=edges, tris_from_file=bool(from_file))


class DataLoader:
    """DataLoader<Scalar> (dataLoader.h:10-37); precision 32 or 64 is the Scalar."""

    def __init__(self, threadsPerBlock: int = 128, precision: int = 32):
        self.threadsPerBlock = threadsPerBlock
        self.precision = precision
        self.bodies: list[tuple[dict, SoftBodyAttribute, str]] = []

    def CollectData(self, nodeFileName: str, eleFileName: str, faceFileName: str | None, pos, scale, rot,
                    centralize: bool, startIndex: int, attrib: SoftBodyAttribute, name: str = "") -> None:
        m = load_tetgen(nodeFileName, eleFileName, faceFileName, pos, scale, rot, centralize, startIndex,
                        self.precision)
        if attrib.DBC and max(attrib.DBC) >= m["X"].shape[0]:
            raise ValueError(f"{name or nodeFileName}: pinned vertex index out of range")
        self.bodies.append((m, attrib, name or os.path.basename(nodeFileName)))

    def AllocData(self) -> Scene:
        """The flat scene; also sets self.edges (surface edges) and self.names."""
        if not self.bodies:
            raise ValueError("DataLoader: no soft body collected")
        X, tet, tri, fathers, edges, mass, mu, lam, dbc = [], [], [], [], [], [], [], [], []
        vstart, tstart = [0], [0]
        for i, (m, a, _) in enumerate(self.bodies):
            off = vstart[-1]
            nv, nt = m["X"].shape[0], m["tet"].shape[0]
            X.append(m["X"])
            tet.append(m["tet"].astype(np.int64) + off)
            tri.append(m["tri"].astype(np.int64) + off)
            edges.append(m["edges"].astype(np.int64) + off)
            fathers.append(np.full(m["tri"].shape[0], i, np.uint32))
            mass.append(np.full(nv, a.mass))
            mu.append(np.full(nt, a.mu))
            lam.append(np.full(nt, a.lam))
            d = np.zeros(nv)
            d[np.asarray(a.DBC, dtype=np.int64)] = 1.0
            dbc.append(d)
            vstart.append(off + nv)
            tstart.append(tstart[-1] + nt)
        self.edges = np.concatenate(edges).astype(np.uint32)
        self.names = [b[2] for b in self.bodies]
        return Scene(X=np.concatenate(X), tet=np.concatenate(tet).astype(np.uint32),
                     tri=np.concatenate(tri).astype(np.uint32).reshape(-1, 3), tri_fathers=np.concatenate(fathers),
                     mass=np.concatenate(mass), mu=np.concatenate(mu), lam=np.concatenate(lam),
                     dbc=np.concatenate(dbc), body_vert_start=np.asarray(vstart, np.int64),
                     body_tet_start=np.asarray(tstart, np.int64))
