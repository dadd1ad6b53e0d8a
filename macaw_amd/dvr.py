"""Harmonic-oscillator DVR (restates Macaw's ``harm_DVR``,
source/dvr.f90:263-344) and sine DVR (``sin_DVR``, source/dvr.f90:107-170).

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


def sin_dvr(n: int, xi: float, xf: float):
    """Sine (particle-in-a-box) DVR on [xi, xf] with n points: return (grid,
    weights, d1, d2, xd1, trafo) as harm_dvr does.

    The grid is equally spaced from xi to xf (symmetrised when centred at
    0), the transformation T_ij = sqrt(2/(n+1)) sin(i j pi/(n+1)); d1 =
    (4/(xf-xi)) T B T with B_ij = (i-j odd) i j/(i^2-j^2); xd1_ij = d1_ij
    (x_i + x_j); d2 from the closed form of the box's second derivative,
    diagonal (2n^2+4n+3)/3 - 1/sin^2(pi i/(n+1)), off-diagonal
    (-1)^(j-i) (1/sin^2(pi(i-j)/(2(n+1))) - 1/sin^2(pi(i+j)/(2(n+1)))),
    all times -(1/2)(pi/(xf-xi+2dx))^2 (the box includes one spacing beyond
    each end point)."""
    if n < 2:
        raise ValueError("too few grid points selected for a DVR")
    dx = (xf - xi) / (n - 1)
    grid = xi + dx * np.arange(n, dtype=np.float64)
    # the reference accumulates the points by repeated addition
    acc = xi
    for i in range(n):
        grid[i] = acc
        acc = acc + dx
    w = np.full(n, dx)
    if abs(abs(xf) - abs(xi)) < np.finfo(float).eps:
        for i in range(1, n // 2 + 1):
            dummy = 0.5 * (-grid[i - 1] + grid[n - i])
            grid[i - 1] = -dummy
            grid[n - i] = dummy
        if n % 2 == 1:
            grid[n // 2] = 0.0
    np1 = float(n + 1)
    idx = np.arange(1, n + 1, dtype=np.float64)
    trafo = np.sin(np.outer(idx, idx) * np.pi / np1) * np.sqrt(2.0 / np1)
    ii, jj = np.meshgrid(idx, idx, indexing="ij")
    with np.errstate(divide="ignore", invalid="ignore"):
        aux = np.where(((ii - jj).astype(int) % 2) != 0, ii * jj / (ii ** 2 - jj ** 2), 0.0)
    d1 = trafo @ ((4.0 / (xf - xi)) * (aux @ trafo))
    xd1 = d1 * (grid[:, None] + grid[None, :])
    d2 = np.zeros((n, n))
    for i in range(1, n + 1):
        d2[i - 1, i - 1] = (2.0 * n * n + 4.0 * n + 3.0) / 3.0 - 1.0 / np.sin(np.pi * i / np1) ** 2
    for i in range(1, n):
        sgn = -1.0
        for j in range(i + 1, n + 1):
            d2[i - 1, j - 1] = (1.0 / np.sin(0.5 * np.pi * (i - j) / np1) ** 2
                                - 1.0 / np.sin(0.5 * np.pi * (i + j) / np1) ** 2) * sgn
            sgn = -sgn
    dummy = 0.5 * (np.pi / (xf - xi + 2.0 * dx)) ** 2
    d2 = -np.triu(d2) * dummy
    d2 = d2 + np.triu(d2, 1).T
    return grid, w, d1, d2, xd1, trafo


def spline(x, y, xx):
    """Cubic spline of the table (x, y) evaluated at xx, as the reference's
    Spline (maths.f90:651-728): on [x_i, x_i+1] the function is y_i +
    a_i d + b_i d^2 + c_i d^3 (d = x - x_i), continuous with its first two
    derivatives at the inner nodes, through y_n, and with c = 0 in the first
    and the last interval (so quadratic extrapolation outside the table).
    The 3(n-1) coefficients solve one linear system."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    n = x.size
    if n < 3:
        raise ValueError("an interpolation table needs at least 3 points")
    m = 3 * n - 3
    M = np.zeros((m, m))
    rhs = np.zeros(m)
    M[0, 2] = 1.0
    ie, ia = 0, 0
    for i in range(n - 2):
        dx = x[i + 1] - x[i]
        M[ie + 1, ia:ia + 3] = (dx, dx ** 2, dx ** 3)
        rhs[ie + 1] = y[i + 1] - y[i]
        M[ie + 2, ia:ia + 4] = (1.0, 2 * dx, 3 * dx ** 2, -1.0)
        M[ie + 3, ia + 1] = 1.0
        M[ie + 3, ia + 2] = 3 * dx
        M[ie + 3, ia + 4] = -1.0
        ie += 3
        ia += 3
    ie += 1
    dx = x[n - 1] - x[n - 2]
    M[ie, ia:ia + 3] = (dx, dx ** 2, dx ** 3)
    rhs[ie] = y[n - 1] - y[n - 2]
    M[ie + 1, ia + 2] = 1.0
    coef = np.linalg.solve(M, rhs)
    out = np.zeros(len(xx))
    for k, xt in enumerate(np.asarray(xx, float)):
        j = 0
        while j < n - 1 and not xt < x[j + 1]:
            j += 1
        j = min(j, n - 2)
        if xt > x[n - 2]:
            j = n - 2
        d = xt - x[j]
        c = 3 * j
        out[k] = y[j] + coef[c] * d + coef[c + 1] * d ** 2 + coef[c + 2] * d ** 3
    return out


def read_table(path):
    """(x, y) columns of an interpolation file (list-directed reads of two
    numbers per line until the first line that does not hold them)."""
    xs, ys = [], []
    with open(path) as fh:
        for line in fh:
            toks = line.replace(",", " ").split()
            if len(toks) < 2:
                break
            try:
                xs.append(float(toks[0].replace("d", "e").replace("D", "e")))
                ys.append(float(toks[1].replace("d", "e").replace("D", "e")))
            except ValueError:
                break
    return np.array(xs), np.array(ys)


def param_operator_on_grid(op, grid, d2):
    """Values of a parametrised operator on a DVR grid (hamilton.f90:212-300):
    the diagonal function for funcs 0-5, the eigenvector for the projector
    (func 6: eigenstate p3 of -p2 d2 + V, V the spline of the file; ascending
    order, eigenvector sign irrelevant to the projector)."""
    g = np.asarray(grid, float)
    f = op.func
    if f == 0:
        return np.exp(op.p1 * g + op.p2 * g ** 2)
    if f == 1:
        x, y = read_table(op.file)
        return spline(x, y, g)
    if f == 2:
        return np.where(np.abs(g - op.p1) < 0.5 * op.p2, 1.0, 0.0)
    if f == 3:
        return g ** int(op.p3) * np.exp(op.p1 * g + op.p2 * g ** 2)
    if f == 4:
        return np.sin(op.p1 * g) ** op.p2
    if f == 5:
        return np.cos(op.p1 * g) ** op.p2
    if f == 6:
        x, y = read_table(op.file)
        H = -op.p2 * np.asarray(d2, float) + np.diag(spline(x, y, g))
        w, U = np.linalg.eigh(H)
        return U[:, int(op.p3)].copy()
    raise ValueError(f"unknown parametrised operator function {f}")
