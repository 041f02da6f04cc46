"""CPU restatement of the reference's TetGen loading for one body.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the library's
loader (sbs_mesh_load_tetgen); never imported by the product.

Restates (reference file:line):
  * loadNodeFile       src/simulation/dataLoader.cu:130-170: the count on the
    first line, then "index x y z" per vertex, coordinates as float32; with
    centralize, subtract the mean (accumulated in the scene precision, divided
    by float(n)) and swap y and z through a float32;
  * loadEleFaceFile    dataLoader.cu:39-127: "index a b c d" per tet and
    "index a b c marker" per face, minus the start index; without faces, the
    tet faces (0 1 2), (0 2 3), (0 3 1), (1 3 2) that occur an odd number of
    times, ordered by their sorted vertex triple, each with the winding of its
    last occurrence;
  * CollectEdges       dataLoader.cu:172-199: sorted unique (min, max) pairs;
  * the model transform of CollectData (dataLoader.cu:215-222) T S Rx Ry Rz with
    the column sum order (m0 x + m1 y) + (m2 z + m3) of glm's mat4 * vec4.
Comment text after '#' and blank lines are skipped.
"""
from __future__ import annotations

import math

import numpy as np


def _records(path):
    out = []
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if line:
                out.append(line.split())
    return out


def _rot(deg, axis):
    a = np.float32(deg) * np.float32(0.01745329251994329576923690768489)
    c, s = np.float32(math.cos(a)), np.float32(math.sin(a))
    ax = np.zeros(3, np.float32)
    ax[axis] = 1
    t = (np.float32(1) - c) * ax
    R = np.zeros((3, 3), np.float32)        # R[col, row]
    R[0] = [c + t[0] * ax[0], np.float32(0) + t[0] * ax[1] + s * ax[2], np.float32(0) + t[0] * ax[2] - s * ax[1]]
    R[1] = [np.float32(0) + t[1] * ax[0] - s * ax[2], c + t[1] * ax[1], np.float32(0) + t[1] * ax[2] + s * ax[0]]
    R[2] = [np.float32(0) + t[2] * ax[0] + s * ax[1], np.float32(0) + t[2] * ax[1] - s * ax[0], c + t[2] * ax[2]]
    return R


def model_matrix(pos, scale, rot) -> np.ndarray:
    """Columns of T(pos) S(scale) Rx Ry Rz in float32: M[col] is a 4-vector."""
    M = np.eye(4, dtype=np.float32)
    M[3] = M[0] * np.float32(pos[0]) + M[1] * np.float32(pos[1]) + M[2] * np.float32(pos[2]) + M[3]
    for k in range(3):
        M[k] = M[k] * np.float32(scale[k])
    for axis in range(3):
        R = _rot(rot[axis], axis)
        cols = [M[0] * R[c, 0] + M[1] * R[c, 1] + M[2] * R[c, 2] for c in range(3)]
        M[0], M[1], M[2] = cols
    return M


def boundary_faces(tet: np.ndarray) -> np.ndarray:
    seen = {}
    for t in np.asarray(tet, dtype=np.int64).tolist():
        for f in ((t[0], t[1], t[2]), (t[0], t[2], t[3]), (t[0], t[3], t[1]), (t[1], t[3], t[2])):
            key = tuple(sorted(f))
            n, _ = seen.get(key, (0, None))
            seen[key] = (n + 1, f)
    faces = [seen[k][1] for k in sorted(seen) if seen[k][0] % 2 == 1]
    return np.asarray(faces, dtype=np.uint32).reshape(-1, 3)


def edges_of(tri: np.ndarray) -> np.ndarray:
    e = set()
    for a, b, c in np.asarray(tri, dtype=np.int64).tolist():
        for u, v in ((a, b), (b, c), (c, a)):
            e.add((min(u, v), max(u, v)))
    return np.asarray(sorted(e), dtype=np.uint32).reshape(-1, 2)


def load(node_file, ele_file, face_file=None, pos=(0, 0, 0), scale=(1, 1, 1), rot=(0, 0, 0), centralize=False,
         start_index=0, precision=32) -> dict:
    S = np.float32 if precision == 32 else np.float64
    nodes = _records(node_file)
    nv = int(nodes[0][0])
    X = np.array([[np.float32(r[1]), np.float32(r[2]), np.float32(r[3])] for r in nodes[1:1 + nv]], dtype=S)
    if centralize:
        c = np.zeros(3, S)
        for p in X:
            c = c + p
        c = c / S(np.float32(nv))
        X = X - c
        y = X[:, 1].astype(np.float32).astype(S)
        X[:, 1] = X[:, 2]
        X[:, 2] = y
    eles = _records(ele_file)
    nt = int(eles[0][0])
    tet = np.array([[int(v) - start_index for v in r[1:5]] for r in eles[1:1 + nt]], dtype=np.int64)
    tri = None
    if face_file:
        try:
            faces = _records(face_file)
        except OSError:
            faces = None
        if faces and int(faces[0][0]) > 0:
            nf = int(faces[0][0])
            tri = np.array([[int(v) - start_index for v in r[1:4]] for r in faces[1:1 + nf]], dtype=np.int64)
    from_file = tri is not None
    if tri is None:
        tri = boundary_faces(tet)
    M = model_matrix(pos, scale, rot).astype(S)
    out = np.empty_like(X)
    for r in range(3):
        a0 = M[0, r] * X[:, 0] + M[1, r] * X[:, 1]
        a1 = M[2, r] * X[:, 2] + M[3, r] * S(1)
        out[:, r] = a0 + a1
    return dict(X=out.astype(np.float64), tet=tet.astype(np.uint32), tri=np.asarray(tri, np.uint32).reshape(-1, 3),
                edges=edges_of(tri), tris_from_file=from_file)
