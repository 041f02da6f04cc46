"""Host half of the halo-staged PCG (csrc/halo_tiles.hpp), compiled alone with g++.

The tiles are checked against what they must describe: every row once, in
blocks of the Morton order, each workgroup's halo exactly its rows and their neighbours,
every entry's halo-local column naming the right row and its value index the
right CSR position, and the greedy cut (a workgroup stops early only when the
next row would overflow its halo).
"""
import os
import subprocess

import numpy as np
import pytest

from sbs_amd import mesh

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "soft-body-simulation-cuda_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("ht") / "halo_tiles_driver")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, os.path.join(HERE, "host", "halo_tiles_driver.cpp"),
                    "-o", exe], check=True)
    return exe


def pattern(tet, nv, diagonal=True):
    """CSR of the scalar pattern: each vertex's neighbours through shared tets (and itself)."""
    t = tet.astype(np.int64)
    pairs = np.concatenate([t[:, [a, b]] for a in range(4) for b in range(4) if a != b], axis=0)
    if diagonal:
        pairs = np.concatenate([pairs, np.stack([np.arange(nv)] * 2, axis=1)])
    key = np.unique(pairs[:, 0] * nv + pairs[:, 1])
    rows, cols = key // nv, key % nv
    rowstart = np.zeros(nv + 1, np.int64)
    np.add.at(rowstart, rows + 1, 1)
    return np.cumsum(rowstart), cols


def run(exe, rowstart, cols, X, rpg=256, stride=768):
    nv = len(rowstart) - 1
    text = f"{nv} {len(cols)} {rpg} {stride}\n" + " ".join(map(str, rowstart)) + "\n" + \
        " ".join(map(str, cols)) + "\n" + " ".join(repr(float(x)) for x in np.asarray(X).ravel()) + "\n"
    p = subprocess.run([exe], input=text, capture_output=True, text=True)
    if p.returncode != 0:
        return p.returncode, None
    out = {}
    for line in p.stdout.strip().split("\n"):
        name, *vals = line.split(" ")
        out[name] = np.array(list(map(int, vals)), dtype=np.int64)
    return 0, out


def morton(X):
    """10-bit-per-axis Morton code of the positions in their bounding box (x highest)."""
    lo, hi = X.min(axis=0), X.max(axis=0)
    ext = np.where(hi > lo, hi - lo, 1.0)
    cell = np.clip(np.where(hi > lo, (X - lo) / ext, 0.0) * 1024.0, 0, 1023).astype(np.int64)
    code = np.zeros(len(X), np.int64)
    for b in range(10):
        for k in range(3):
            code |= ((cell[:, k] >> b) & 1) << (3 * b + (2 - k))
    return code


def check_tiles(h, rowstart, cols, X, rpg, stride):
    nv = len(rowstart) - 1
    perm, row_of = h["row_perm"], h["row_of"]
    assert np.array_equal(np.sort(perm), np.arange(nv))
    assert np.array_equal(row_of[perm], np.arange(nv))
    # workgroups tile the rows in order: each takes the next block of the
    # Morton order (ties by vertex id), its rows numbered by ascending vertex id
    key = morton(np.asarray(X))
    assert np.array_equal(h["keys"], key)
    morton_order = np.lexsort((np.arange(nv), key))
    row0, rows, hn, width, eptr = h["g_row0"], h["g_rows"], h["g_halo_n"], h["g_width"], h["g_eptr"]
    ng = len(row0)
    assert row0[0] == 0 and np.array_equal(row0[1:], np.cumsum(rows)[:-1]) and rows.sum() == nv
    for g in range(ng):
        blk = slice(row0[g], row0[g] + rows[g])
        assert np.array_equal(perm[blk], np.sort(morton_order[blk]))
    assert (rows >= 1).all() and (rows <= rpg).all() and (hn <= stride).all()
    assert h["max_halo"][0] == hn.max()
    halo = h["halo"].reshape(ng, stride)
    assert np.array_equal(h["halo_vertex"], perm[h["halo"]])
    col, src = h["col"], h["src"]
    rl = np.diff(rowstart)
    for g in range(ng):
        r = np.arange(row0[g], row0[g] + rows[g])
        verts = perm[r]
        nb = np.unique(np.concatenate([r] + [row_of[cols[rowstart[v]:rowstart[v + 1]]] for v in verts]))
        assert np.array_equal(halo[g, :hn[g]], nb), g                      # exactly rows + neighbours, ascending
        assert (halo[g, hn[g]:] == row0[g]).all()
        own = h["g_own0"][g] + np.arange(rows[g])                        # the rows' own halo-local columns
        assert np.array_equal(halo[g, own], r)
        assert width[g] == rl[verts].max()
        base = eptr[g]
        for t in range(rpg):
            e = base + np.arange(width[g]) * rpg + t
            if t < rows[g]:
                v = verts[t]
                n = rl[v]
                assert np.array_equal(halo[g, col[e[:n]]], row_of[cols[rowstart[v]:rowstart[v + 1]]])
                assert np.array_equal(src[e[:n]], np.arange(rowstart[v], rowstart[v + 1]))
                pad = e[n:]
                assert (src[pad] == -1).all() and (halo[g, col[pad]] == row0[g] + t).all()
            else:
                assert (src[e] == -1).all() and (halo[g, col[e]] == row0[g]).all()
        # greedy cut: a short workgroup (not the last) could not take the next
        # row of the Morton order
        if rows[g] < rpg and g + 1 < ng:
            nv_next = morton_order[row0[g] + rows[g]]
            more = np.concatenate([nb, [row_of[nv_next]], row_of[cols[rowstart[nv_next]:rowstart[nv_next + 1]]]])
            assert len(np.unique(more)) > stride, g
    assert eptr[-1] + width[-1] * rpg == len(col) == len(src)


def test_cube_tiles(driver):
    s = mesh.cube_scene(12)
    rs, cols = pattern(s.tet, s.nv)
    rc, h = run(driver, rs, cols, s.X, 256, 768)
    assert rc == 0
    check_tiles(h, rs, cols, s.X, 256, 768)
    # Morton blocks of a grid keep the halo near 256 rows' 1-ring (a 8x8x4 block: 10x10x6)
    assert h["max_halo"][0] <= 700


def test_multi_body_tiles_tight_halo(driver):
    """Several bodies, a halo capacity that forces early cuts, 64-row workgroups."""
    s = mesh.make_scene([dict(nx=5, ny=4, nz=3, seed=1, pin_top=True),
                         dict(nx=3, ny=6, nz=2, origin=(1.0, 0.0, 0.0), seed=2, pin_top=True),
                         dict(nx=4, ny=4, nz=4, origin=(0.0, 1.0, 0.0), seed=3, pin_top=True)])
    rs, cols = pattern(s.tet, s.nv)
    rc, h = run(driver, rs, cols, s.X, 64, 90)
    assert rc == 0
    check_tiles(h, rs, cols, s.X, 64, 90)
    assert (h["g_rows"][:-1] < 64).any()          # the capacity did cut some workgroups


def test_off_diagonal_pattern(driver):
    s = mesh.cube_scene(5)
    rs, cols = pattern(s.tet, s.nv, diagonal=False)
    rc, h = run(driver, rs, cols, s.X, 128, 400)
    assert rc == 0
    check_tiles(h, rs, cols, s.X, 128, 400)


def test_row_larger_than_halo_is_refused(driver):
    s = mesh.cube_scene(3)
    rs, cols = pattern(s.tet, s.nv)
    rc, _ = run(driver, rs, cols, s.X, 64, int(np.diff(rs).max()) - 1)
    assert rc == 3
