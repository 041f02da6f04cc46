"""Synthetic tetrahedral scenes: hex grids split into Kuhn tetrahedra.

The reference loads TetGen ``.node/.ele/.face`` files and concatenates every
soft body into one flat scene (``src/simulation/dataLoader.cu:204-237`` and
``AllocData`` at ``dataLoader.cu:292-378``).  The workloads measured here are
hex grids, so this module builds the same flat arrays directly:

* vertices in x-fastest order, ``idx = i + (nx+1)*(j + (ny+1)*k)``;
* each hex split into the 6 Kuhn tetrahedra that share the hex diagonal
  (0,0,0)-(1,1,1); the split is conforming across neighbouring hexes;
* every tetrahedron stored with positive orientation (det Dm > 0);
* surface triangles chosen by the reference's face-parity rule
  (``dataLoader.cu:92-127``): a tet face that appears an odd number of times
  is a boundary face, emitted in ascending order of its sorted vertex triple
  with the winding of the tet face that produced it;
* a counter-based seeded perturbation (splitmix64) so any language can
  reproduce the same scene.

Per-vertex mass follows the recorded decision m_v = rho * V_body / N_v (the
reference fills one scalar mass per body, ``dataLoader.cu:357``).
"""
from __future__ import annotations

import dataclasses
import numpy as np

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, counter: np.ndarray) -> np.ndarray:
    """splitmix64 of (seed + golden*counter); vectorised, wraps mod 2**64."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (counter.astype(np.uint64) + np.uint64(1)))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, counter: np.ndarray) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, counter) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# Kuhn split: for each permutation (a, b, c) of the axes the tet is
# 0 -> e_a -> e_a+e_b -> (1,1,1).  Corners are numbered bit0=x, bit1=y, bit2=z.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _kuhn_corner_table() -> np.ndarray:
    tets = []
    for a, b, c in _PERMS:
        c0 = 0
        c1 = 1 << a
        c2 = c1 | (1 << b)
        c3 = 7
        # orientation of (e_a, e_a+e_b, 1,1,1): det = sign of the permutation
        e = np.eye(3)
        d = np.linalg.det(np.stack([e[a], e[a] + e[b], np.ones(3)], axis=1))
        tets.append((c0, c1, c2, c3) if d > 0 else (c0, c2, c1, c3))
    return np.asarray(tets, dtype=np.int64)


KUHN = _kuhn_corner_table()


@dataclasses.dataclass
class Scene:
    """Flat multi-body scene, same arrays as the reference's SolverData."""

    X: np.ndarray          # (N_v, 3) float64 rest positions (after perturbation / offsets)
    tet: np.ndarray        # (N_t, 4) uint32
    tri: np.ndarray        # (N_tri, 3) uint32 surface triangles
    tri_fathers: np.ndarray  # (N_tri,) uint32 body id per triangle
    mass: np.ndarray       # (N_v,) float64
    mu: np.ndarray         # (N_t,) float64
    lam: np.ndarray        # (N_t,) float64
    dbc: np.ndarray        # (N_v,) float64, 1 for pinned vertices
    body_vert_start: np.ndarray  # (n_bodies+1,) vertex offsets
    body_tet_start: np.ndarray   # (n_bodies+1,) tet offsets

    @property
    def nv(self) -> int:
        return int(self.X.shape[0])

    @property
    def nt(self) -> int:
        return int(self.tet.shape[0])

    @property
    def ntri(self) -> int:
        return int(self.tri.shape[0])


def lame(E: float, nu: float) -> tuple[float, float]:
    """mu = E/(2(1+nu)), lambda = E nu/((1+nu)(1-2nu))."""
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def hex_grid_tets(nx: int, ny: int, nz: int) -> np.ndarray:
    """(6*nx*ny*nz, 4) vertex indices, hex-major then the 6 Kuhn tets."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # hex order: x fastest, then y, then z (matches the vertex order)
    i = i.transpose(2, 1, 0).ravel()
    j = j.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    corner = np.empty((i.size, 8), dtype=np.int64)
    for c in range(8):
        corner[:, c] = (i + (c & 1)) + sx * (j + ((c >> 1) & 1)) + sy * (k + ((c >> 2) & 1))
    tets = corner[:, KUHN]            # (nhex, 6, 4)
    return tets.reshape(-1, 4).astype(np.uint32)


def surface_triangles(tet: np.ndarray) -> np.ndarray:
    """Boundary faces by the reference's parity rule (dataLoader.cu:92-127)."""
    t = tet.astype(np.int64)
    faces = np.concatenate([
        t[:, [0, 1, 2]], t[:, [0, 2, 3]], t[:, [0, 3, 1]], t[:, [1, 3, 2]]], axis=0)
    keys = np.sort(faces, axis=1)
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    ks = keys[order]
    # group identical keys; keep groups with odd multiplicity
    new = np.ones(len(ks), dtype=bool)
    new[1:] = np.any(ks[1:] != ks[:-1], axis=1)
    starts = np.flatnonzero(new)
    counts = np.diff(np.append(starts, len(ks)))
    keep = starts[(counts % 2) == 1]
    # the reference keeps the orientation of the last face inserted for a key;
    # boundary keys occur once, so that is the only face with the key
    last = np.append(starts[1:], len(ks)) - 1
    keep_last = last[(counts % 2) == 1]
    del keep
    return faces[order[keep_last]].astype(np.uint32)


def hex_grid_body(nx: int, ny: int, nz: int, spacing: float = 0.1, origin=(0.0, 0.0, 0.0),
                  jitter: float = 1e-3, seed: int = 0):
    """Vertices (float64), tets and surface triangles of one hex-grid body."""
    gi, gj, gk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    P = np.stack([gi.transpose(2, 1, 0).ravel(), gj.transpose(2, 1, 0).ravel(),
                  gk.transpose(2, 1, 0).ravel()], axis=1).astype(np.float64) * spacing
    P += np.asarray(origin, dtype=np.float64)
    if jitter > 0.0:
        n = P.size
        u = uniform01(seed, np.arange(n, dtype=np.uint64)).reshape(P.shape)
        P += (2.0 * u - 1.0) * jitter
    tet = hex_grid_tets(nx, ny, nz)
    tri = surface_triangles(tet)
    return P, tet, tri


def make_scene(bodies: list[dict]) -> Scene:
    """Concatenate bodies like DataLoader::AllocData (dataLoader.cu:292-378).

    Each body dict: nx, ny, nz, spacing, origin, jitter, seed, density, E, nu,
    pin_top (bool: pin vertices with the maximum grid y index), pin_bottom
    (bool: pin vertices with grid y index 0).
    """
    Xs, tets, tris, fathers, masses, mus, lams, dbcs = [], [], [], [], [], [], [], []
    vstart, tstart = [0], [0]
    for b_id, b in enumerate(bodies):
        nx, ny, nz = b["nx"], b["ny"], b["nz"]
        sp = b.get("spacing", 0.1)
        P, tet, tri = hex_grid_body(nx, ny, nz, sp, b.get("origin", (0.0, 0.0, 0.0)),
                                    b.get("jitter", 1e-3), b.get("seed", 0))
        off = vstart[-1]
        nvb = P.shape[0]
        vol = (nx * sp) * (ny * sp) * (nz * sp)
        mu, lam = lame(b.get("E", 1e5), b.get("nu", 0.3))
        dbc = np.zeros(nvb)
        jj = (np.arange(nvb) // (nx + 1)) % (ny + 1)
        if b.get("pin_top", False):
            dbc[jj == ny] = 1.0
        if b.get("pin_bottom", False):
            dbc[jj == 0] = 1.0
        Xs.append(P)
        tets.append(tet.astype(np.int64) + off)
        tris.append(tri.astype(np.int64) + off)
        fathers.append(np.full(tri.shape[0], b_id, dtype=np.uint32))
        masses.append(np.full(nvb, b.get("density", 1000.0) * vol / nvb))
        mus.append(np.full(tet.shape[0], mu))
        lams.append(np.full(tet.shape[0], lam))
        dbcs.append(dbc)
        vstart.append(off + nvb)
        tstart.append(tstart[-1] + tet.shape[0])
    return Scene(X=np.concatenate(Xs), tet=np.concatenate(tets).astype(np.uint32),
                 tri=np.concatenate(tris).astype(np.uint32), tri_fathers=np.concatenate(fathers),
                 mass=np.concatenate(masses), mu=np.concatenate(mus), lam=np.concatenate(lams),
                 dbc=np.concatenate(dbcs), body_vert_start=np.asarray(vstart, dtype=np.int64),
                 body_tet_start=np.asarray(tstart, dtype=np.int64))


def cube_scene(n: int, spacing: float = 0.1, jitter: float = 1e-3, seed: int = 0,
               E: float = 1e5, nu: float = 0.3, density: float = 1000.0, pin_top: bool = True) -> Scene:
    """One n^3-hex cube, top face pinned: the single-body PD workloads."""
    return make_scene([dict(nx=n, ny=n, nz=n, spacing=spacing, jitter=jitter, seed=seed,
                            E=E, nu=nu, density=density, pin_top=pin_top)])


def small_bodies_scene(nbodies: int = 256, n: int = 6, seed: int = 0) -> Scene:
    """The mixed-precision small-body workload: nbodies n^3-hex cubes on a
    grid 1 m apart, E per body uniform in [5e4, 5e5] Pa from the seed,
    nu = 0.3, top faces pinned, +-1 mm jitter."""
    E = uniform01(seed, np.arange(nbodies, dtype=np.uint64)) * 4.5e5 + 5e4
    side = int(np.ceil(np.sqrt(nbodies)))
    bodies = [dict(nx=n, ny=n, nz=n, spacing=0.1, origin=(1.0 * (i % side), 0.0, 1.0 * (i // side)),
                   jitter=1e-3, seed=seed + 1 + i, E=float(E[i]), nu=0.3, density=1000.0, pin_top=True)
              for i in range(nbodies)]
    return make_scene(bodies)


def scalar_laplacian_nnz(tet: np.ndarray, nv: int) -> int:
    """N_v + 2*(number of distinct mesh edges): the scalar system's nonzeros."""
    t = tet.astype(np.int64)
    pairs = np.concatenate([t[:, [a, b]] for a in range(4) for b in range(a + 1, 4)], axis=0)
    pairs.sort(axis=1)
    e = np.unique(pairs[:, 0] * nv + pairs[:, 1])
    return nv + 2 * int(e.size)


def stack_scene(n: int, spacing: float = 0.1, gap: float = 2e-3, jitter: float = 5e-4, seed: int = 0,
                E: float = 1e5, nu: float = 0.3, density: float = 1000.0) -> Scene:
    """Two n^3-hex cubes, the lower one pinned on its bottom face, the upper one
    `gap` above it: the contact workload of the IPC step."""
    top = n * spacing + gap + 2 * jitter
    return make_scene([dict(nx=n, ny=n, nz=n, spacing=spacing, jitter=jitter, seed=seed, E=E, nu=nu,
                            density=density, pin_bottom=True),
                       dict(nx=n, ny=n, nz=n, spacing=spacing, jitter=jitter, seed=seed + 1, E=E, nu=nu,
                            density=density, origin=(0.0, top, 0.0))])
