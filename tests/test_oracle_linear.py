"""Pins of the general linear-solve oracle (oracle/linear_oracle.py), CPU only.

* sort_coo against scipy.sparse's COO -> CSR conversion (an independent
  implementation that also sums repeated entries);
* the PCG solution against numpy's dense solve;
* closed forms of Jacobi-preconditioned CG: a diagonal matrix is solved in one
  iteration, a matrix whose preconditioned form has two distinct eigenvalues
  in at most two, the exact solution as the guess in none; max_iter holds.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import linear_oracle as LO


def random_spd_coo(n, seed, dup=True):
    """An SPD matrix (graph Laplacian + diagonal) as shuffled COO triplets, with
    every entry split in two when `dup`."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in rng.choice(n, size=min(4, n), replace=False):
            if i == j:
                continue
            w = rng.uniform(0.1, 1.0)
            rows += [i, j, i, j]
            cols += [j, i, i, j]
            vals += [-w, -w, w, w]
    rows += list(range(n))
    cols += list(range(n))
    vals += list(rng.uniform(0.5, 2.0, n))
    rows, cols, vals = np.array(rows), np.array(cols), np.array(vals)
    if dup:
        f = rng.uniform(0.2, 0.8, vals.size)
        rows, cols, vals = np.concatenate([rows, rows]), np.concatenate([cols, cols]), np.concatenate([vals * f, vals * (1 - f)])
    perm = rng.permutation(vals.size)
    return rows[perm].astype(np.int32), cols[perm].astype(np.int32), vals[perm]


def test_sort_coo_matches_scipy():
    n = 40
    r, c, v = random_spd_coo(n, 1)
    rowptr, col, val = LO.sort_coo(n, v, r, c)
    ref = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    ref.sort_indices()
    np.testing.assert_array_equal(rowptr, ref.indptr)
    np.testing.assert_array_equal(col, ref.indices)
    np.testing.assert_allclose(val, ref.data, rtol=1e-14, atol=1e-15)


def test_empty_rows_and_bad_indices():
    rowptr, col, val = LO.sort_coo(5, np.array([1.0, 2.0]), np.array([3, 1]), np.array([3, 1]))
    np.testing.assert_array_equal(rowptr, [0, 0, 1, 1, 2, 2])
    with pytest.raises(ValueError):
        LO.sort_coo(3, np.array([1.0]), np.array([3]), np.array([0]))


def test_solution_matches_dense_solve():
    n = 60
    r, c, v = random_spd_coo(n, 2)
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).toarray()
    b = np.random.default_rng(3).normal(size=n)
    lmin = np.linalg.eigvalsh(A).min()
    # stopped by ||r|| < tol: the error is at most ||A^-1|| tol
    x, k = LO.pcg_jacobi(n, b, v, r, c, tol=1e-6, max_iter=500)
    assert 0 < k < 500 and np.linalg.norm(A @ x - b) < 1.01e-6
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1.01e-6 / lmin
    # with a tiny tol the absolute |r.z| < 1e-15 test stops it first
    x, k2 = LO.pcg_jacobi(n, b, v, r, c, tol=1e-14, max_iter=500)
    res = np.linalg.norm(A @ x - b)
    assert k2 > k and 1e-14 < res < 1e-6
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1.01 * res / lmin


def test_diagonal_in_one_iteration():
    n = 10
    d = np.arange(1.0, n + 1)
    b = np.linspace(-1, 1, n)
    x, k = LO.pcg_jacobi(n, b, d, np.arange(n), np.arange(n), tol=1e-12)
    assert k == 1
    np.testing.assert_allclose(x, b / d, rtol=1e-15)


def test_two_eigenvalues_in_two_iterations():
    n, s, a = 12, 3.0, 0.7
    A = s * (np.eye(n) + a * np.ones((n, n)))    # D^-1 A has eigenvalues 1/(1+a) and (1+n a)/(1+a)
    r, c = np.nonzero(np.ones((n, n)))
    b = np.random.default_rng(4).normal(size=n)
    x, k = LO.pcg_jacobi(n, b, A[r, c], r, c, tol=1e-10)
    assert k == 2
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-12)


def test_exact_guess_and_iteration_limit():
    n = 30
    r, c, v = random_spd_coo(n, 5)
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).toarray()
    xs = np.random.default_rng(6).normal(size=n)
    x, k = LO.pcg_jacobi(n, A @ xs, v, r, c, guess=xs, tol=1e-9)
    assert k == 0 and np.array_equal(x, xs)
    _, k = LO.pcg_jacobi(n, A @ xs, v, r, c, tol=1e-30, max_iter=7)
    assert k == 7
