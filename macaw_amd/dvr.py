"""Harmonic-oscillator DVR (restates Macaw's ``harm_DVR``,
source/dvr.f90:263-344).

The grid is the spectrum of the position operator in the first ``n``
harmonic-oscillator functions (tridiagonal, off-diagonal sqrt(i/(2 m w))).
Eigenvector signs are fixed so that the first component is positive, which
makes the DVR matrices independent of the eigensolver used.  Derivative
matrices are the FBR matrices of d/dx, x d/dx + d/dx x and d^2/dx^2
transformed to the DVR, then (anti)symmetrised exactly as the reference does.
"""
from __future__ import annotations

import numpy as np


def harm_dvr(n: int, x0: float, mw: float):
    """Return (grid, weights, d1, d2, xd1, trafo) as float64 arrays.

    Matrices are (n, n) with the reference's index meaning: ``d1[i, j]`` is
    the DVR matrix element between grid points i and j.
    """
    if n < 2:
        raise ValueError("too few grid points selected for a DVR")
    xmat = np.zeros((n, n))
    for i in range(1, n):
        xmat[i - 1, i] = np.sqrt(0.5 * i / mw)
        xmat[i, i - 1] = xmat[i - 1, i]
    grid, trafo = np.linalg.eigh(xmat)
    grid = grid.copy()
    for i in range(1, n // 2 + 1):
        dummy = 0.5 * (-grid[i - 1] + grid[n - i])
        grid[i - 1] = -dummy
        grid[n - i] = dummy
    if n % 2 == 1:
        grid[n // 2] = 0.0
    grid = grid + x0
    trafo = trafo.copy()
    for i in range(n):
        if trafo[0, i] < 0.0:
            trafo[:, i] = -trafo[:, i]
    w = trafo[0, :] ** 2 * np.exp(mw * (grid - x0) ** 2) * np.sqrt(np.pi / mw)

    aux = np.zeros((n, n))
    for i in range(1, n):
        aux[i - 1, i] = np.sqrt(float(i))
        aux[i, i - 1] = -np.sqrt(float(i))
    aux *= np.sqrt(0.5 * mw)
    d1 = trafo.T @ (aux @ trafo)

    aux = np.zeros((n, n))
    for i in range(1, n - 1):
        aux[i - 1, i + 1] = np.sqrt(float(i * (i + 1)))
        aux[i + 1, i - 1] = -np.sqrt(float(i * (i + 1)))
    xd1 = trafo.T @ (aux @ trafo)

    for m in (d1, xd1):
        anti = 0.5 * (m - m.T)
        m[:, :] = anti
        m[np.diag_indices(n)] = 0.0

    aux = np.zeros((n, n))
    for i in range(1, n + 1):
        aux[i - 1, i - 1] = -float(2 * i - 1)
    for i in range(1, n - 1):
        aux[i - 1, i + 1] = np.sqrt(float(i * (i + 1)))
        aux[i + 1, i - 1] = aux[i - 1, i + 1]
    d2 = 0.5 * mw * (trafo.T @ aux @ trafo)
    d2 = np.triu(d2) + np.triu(d2, 1).T
    return grid, w, d1, d2, xd1, trafo
