"""Host half of the dense small-body solver (csrc/dense_factor.hpp), compiled alone with g++."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from sbs_amd import mesh

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "soft-body-simulation-cuda_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("df") / "dense_factor_driver")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, os.path.join(HERE, "host", "dense_factor_driver.cpp"),
                    "-o", exe], check=True)
    return exe


def run(exe, rowptr, col, val):
    n = len(rowptr) - 1
    text = f"{n} {len(col)}\n" + " ".join(map(str, rowptr)) + "\n" + " ".join(map(str, col)) + "\n" + \
        " ".join(repr(float(v)) for v in val) + "\n"
    p = subprocess.run([exe], input=text, capture_output=True, text=True)
    lines = p.stdout.split("\n")
    nc, contig = map(int, lines[0].split())
    comp = np.array(list(map(int, lines[1].split())))
    invs, k = [], 2
    while p.returncode == 0 and k + 1 < len(lines) and lines[k].strip():
        m = int(lines[k])
        invs.append(np.array(list(map(float, lines[k + 1].split()))).reshape(m, m))
        k += 2
    return p.returncode, nc, bool(contig), comp, invs


def dense(rowptr, col, val, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    return A


def test_bodies_and_inverses_of_a_multi_body_pd_matrix(driver):
    s = mesh.make_scene([dict(nx=2, ny=2, nz=2, seed=1, pin_top=True, E=1e5),
                         dict(nx=3, ny=2, nz=2, origin=(1.0, 0, 0), seed=2, pin_top=True, E=3e5),
                         dict(nx=2, ny=2, nz=3, origin=(2.0, 0, 0), seed=3, pin_top=True, E=5e4)])
    ex = O.PdOracle(s, O.make_params(), fp64=True).export()
    rc, nc, contig, comp, invs = run(driver, ex["rowptr"], ex["col"], ex["val"])
    assert rc == 0 and nc == 3 and contig
    assert np.array_equal(np.bincount(comp), np.diff(s.body_vert_start))
    A = dense(ex["rowptr"], ex["col"], ex["val"], s.nv)
    for b, inv in enumerate(invs):
        r0, r1 = s.body_vert_start[b], s.body_vert_start[b + 1]
        Ab = A[r0:r1, r0:r1]
        np.testing.assert_allclose(Ab @ inv, np.eye(r1 - r0), atol=1e-10)
        np.testing.assert_allclose(inv, np.linalg.inv(Ab), rtol=1e-9, atol=1e-15 * np.abs(inv).max())


def test_non_contiguous_bodies_are_reported(driver):
    # rows 0 and 2 coupled, 1 alone: components {0,2}, {1} are not contiguous ranges
    rowptr, col, val = [0, 2, 3, 5], [0, 2, 1, 0, 2], [2.0, -1.0, 1.0, -1.0, 2.0]
    rc, nc, contig, comp, invs = run(driver, rowptr, col, val)
    assert rc == 0 and nc == 2 and not contig
    assert comp[0] == comp[2] != comp[1]


def test_indefinite_block_is_rejected(driver):
    rowptr, col, val = [0, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 2.0, 1.0]
    rc, *_ = run(driver, rowptr, col, val)
    assert rc == 3
