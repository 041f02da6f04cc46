"""Host half of the direct solver (csrc/band_factor.hpp), compiled alone with g++."""
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
    exe = str(tmp_path_factory.mktemp("bf") / "band_factor_driver")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, os.path.join(HERE, "host", "band_factor_driver.cpp"),
                    "-o", exe], check=True)
    return exe


def run(exe, rowptr, col, val):
    n = len(rowptr) - 1
    text = f"{n} {len(col)}\n" + " ".join(map(str, rowptr)) + "\n" + " ".join(map(str, col)) + "\n" + \
        " ".join(repr(float(v)) for v in val) + "\n"
    p = subprocess.run([exe], input=text, capture_output=True, text=True)
    if p.returncode != 0:
        return p.returncode, None
    lines = p.stdout.split("\n")
    n_, bw = map(int, lines[0].split())
    perm = np.array(list(map(int, lines[1].split())))
    Lb = np.array(list(map(float, lines[2].split()))).reshape(n, bw + 1)
    L = np.zeros((n, n))
    for i in range(n):
        for d in range(bw + 1):
            if i - d >= 0:
                L[i, i - d] = Lb[i, d]
    run.systems = list(map(int, lines[3].split()))
    run.bw_estimate = int(lines[4])
    return 0, (bw, perm, L)


def dense(rowptr, col, val, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    return A


def test_factor_of_the_pd_matrix(driver):
    s = mesh.cube_scene(5)
    ex = O.PdOracle(s, O.make_params(), fp64=True).export()
    rc, (bw, perm, L) = run(driver, ex["rowptr"], ex["col"], ex["val"])
    assert rc == 0
    A = dense(ex["rowptr"], ex["col"], ex["val"], s.nv)
    Ap = A[np.ix_(perm, perm)]
    np.testing.assert_allclose(L @ L.T, Ap, rtol=1e-12, atol=1e-12 * np.abs(A).max())
    assert sorted(perm) == list(range(s.nv))
    assert run.systems == [0, s.nv] and run.bw_estimate == bw
    nat_bw = max(abs(r - c) for r in range(s.nv) for c in ex["col"][ex["rowptr"][r]:ex["rowptr"][r + 1]])
    assert bw <= nat_bw


def test_rcm_narrows_a_scrambled_band(driver):
    rng = np.random.default_rng(0)
    s = mesh.cube_scene(4)
    ex = O.PdOracle(s, O.make_params(), fp64=True).export()
    A = dense(ex["rowptr"], ex["col"], ex["val"], s.nv)
    q = rng.permutation(s.nv)
    B = A[np.ix_(q, q)]
    rows, cols = np.nonzero(B)
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=s.nv))])
    rc, (bw, perm, L) = run(driver, rowptr, cols, B[rows, cols])
    assert rc == 0
    assert bw < 0.5 * max(abs(r - c) for r, c in zip(rows, cols))
    np.testing.assert_allclose(L @ L.T, B[np.ix_(perm, perm)], rtol=1e-12, atol=1e-12 * np.abs(B).max())


def test_two_components_and_errors(driver):
    # block-diagonal SPD with two components
    A = np.array([[4.0, 1, 0, 0], [1, 3, 0, 0], [0, 0, 2, -1], [0, 0, -1, 2]])
    rows, cols = np.nonzero(A)
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=4))])
    rc, (bw, perm, L) = run(driver, rowptr, cols, A[rows, cols])
    assert rc == 0
    np.testing.assert_allclose(L @ L.T, A[np.ix_(perm, perm)], rtol=1e-14)
    assert run.systems == [0, 2, 4]
    # indefinite -> not positive definite
    A[0, 0] = -1.0
    assert run(driver, rowptr, cols, A[rows, cols])[0] == 3
    # malformed CSR -> invalid argument
    bad = cols.copy()
    bad[0] = 9
    assert run(driver, rowptr, bad, A[rows, cols])[0] == 2


def test_independent_systems_of_a_multi_body_scene(driver):
    """Three bodies: three systems, each a contiguous block of whole bodies, and
    no coupling across a cut (the solve runs each system in its own workgroup)."""
    s = mesh.make_scene([dict(nx=3, ny=2, nz=2, seed=1), dict(nx=2, ny=2, nz=2, seed=2, origin=(3.0, 0, 0)),
                         dict(nx=4, ny=2, nz=1, seed=3, origin=(0, 3.0, 0))])
    ex = O.PdOracle(s, O.make_params(), fp64=True).export()
    rowptr, col = np.asarray(ex["rowptr"]), np.asarray(ex["col"])
    rc, (bw, perm, L) = run(driver, rowptr, col, ex["val"])
    assert rc == 0
    sysv = run.systems
    assert len(sysv) == 4 and sysv[0] == 0 and sysv[-1] == s.nv
    pos = np.empty(s.nv, int)
    pos[perm] = np.arange(s.nv)
    which = np.searchsorted(sysv, pos, side="right") - 1
    for r in range(s.nv):
        assert np.all(which[col[rowptr[r]:rowptr[r + 1]]] == which[r])
    assert run.bw_estimate == bw
