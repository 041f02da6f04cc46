"""CPU oracle of the general linear solve (test infrastructure only).

Restates, in numpy, what LinearSolver<T>::Solve does with the PCGJacobiSolver
algorithm; only tests/ use it, as the checker.  The product never imports it.

  sort_coo                   src/simulation/solver/linear/linear.cu:7-40
      sort the (row, col, value) triplets by (row, col) and sum repeats
  ExtractInverseDiagonalKernel  pcgJacobi.cu:6-24
      1/diag per row, 1 where the diagonal is missing or |d| < 1e-9
  PCGJacobiSolver::Solve     pcgJacobi.cu:89-173
      x = guess or 0; r = b - A x; for k < max_iter: stop if ||r||_2 < tol;
      z = D^-1 r; rho = r.z; stop if |rho| < 1e-15; p = z (k = 0) or
      z + (rho / rho_prev) p; q = A p; alpha = rho / p.q; x += alpha p;
      r -= alpha q

Arithmetic: vectors in the solve's precision, dot products accumulated in
float64 (the product does the same; cuBLAS's reduction order is not public,
so its rounding is not restated).
"""
from __future__ import annotations

import numpy as np


def sort_coo(n: int, A, row, col):
    """(rowptr, col, val) of the summed triplets, rows and columns ascending."""
    row = np.asarray(row, dtype=np.int64)
    col = np.asarray(col, dtype=np.int64)
    A = np.asarray(A)
    if row.size and (row.min() < 0 or row.max() >= n or col.min() < 0 or col.max() >= n):
        raise ValueError("index outside [0, N)")
    key = row * n + col
    order = np.argsort(key, kind="stable")
    ks = key[order]
    vs = A[order]
    if ks.size == 0:
        return np.zeros(n + 1, np.int64), np.zeros(0, np.int64), np.zeros(0, A.dtype)
    start = np.ones(ks.size, bool)
    start[1:] = ks[1:] != ks[:-1]
    idx = np.flatnonzero(start)
    # sum each run in its stored order (the order the repeats were given)
    val = np.array([vs[a:b].sum(dtype=np.float64) for a, b in zip(idx, np.append(idx[1:], ks.size))],
                   dtype=A.dtype)
    uk = ks[idx]
    urow, ucol = uk // n, uk % n
    rowptr = np.searchsorted(urow, np.arange(n + 1), side="left")
    return rowptr, ucol, val


def inv_diag(n, rowptr, col, val):
    d = np.ones(n, dtype=val.dtype)
    for r in range(n):
        seg = slice(rowptr[r], rowptr[r + 1])
        hit = np.flatnonzero(col[seg] == r)
        if hit.size:
            d[r] = val[seg][hit[0]]
    d[np.abs(d.astype(np.float64)) < 1e-9] = 1
    return (1 / d).astype(val.dtype)


def csr_matvec(rowptr, col, val, v):
    out = np.zeros(rowptr.size - 1, dtype=val.dtype)
    for r in range(rowptr.size - 1):
        seg = slice(rowptr[r], rowptr[r + 1])
        out[r] = (val[seg] * v[col[seg]]).sum(dtype=val.dtype)
    return out


def pcg_jacobi(n, b, A, row, col, guess=None, max_iter=2000, tol=1e-5, dtype=np.float64):
    """x and the iteration count k at which the loop stopped."""
    b = np.asarray(b, dtype=dtype)
    rowptr, c, val = sort_coo(n, np.asarray(A, dtype=dtype), row, col)
    dinv = inv_diag(n, rowptr, c, val)
    if guess is not None:
        x = np.asarray(guess, dtype=dtype).copy()
        r = b - csr_matvec(rowptr, c, val, x)
    else:
        x = np.zeros(n, dtype)
        r = b.copy()
    p = np.zeros(n, dtype)
    rho_prev = 0.0
    k = 0
    for k in range(max_iter + 1):
        if k == max_iter:
            break
        if np.sqrt(np.dot(r.astype(np.float64), r.astype(np.float64))) < tol:
            break
        z = (dinv * r).astype(dtype)
        rho = float(np.dot(r.astype(np.float64), z.astype(np.float64)))
        if abs(rho) < 1e-15:
            break
        p = z.copy() if k == 0 else (z + dtype(rho / rho_prev) * p).astype(dtype)
        rho_prev = rho
        q = csr_matvec(rowptr, c, val, p)
        alpha = dtype(rho / float(np.dot(p.astype(np.float64), q.astype(np.float64))))
        x = (x + alpha * p).astype(dtype)
        r = (r - alpha * q).astype(dtype)
    return x, k
