"""Brute-force numpy restatements used to pin the oracle on tiny problems.

Everything here is built from the mathematics (dense Kronecker products,
einsum contractions, eigh), not from the oracle's loop structure:

* primitive operators as dense grid matrices (p = -i d/dx etc.);
* the operator basis as a dict keyed by (labels, codes);
* the full multi-set Hamiltonian in the SPF configuration basis;
* density matrices and mean fields as einsum hole contractions.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg


def prim_matrix(prob, dof: int, code: int) -> np.ndarray:
    """Dense grid matrix of a primitive operator on DOF ``dof`` (1-based)."""
    n = int(prob.ngp[dof - 1])
    if code == 0:
        return np.eye(n, dtype=complex)
    if code > 200:
        return np.diag(prob.gridoper[:n, code - 201, dof - 1]).astype(complex)
    if code <= -200:
        phi = prob.gridoper[:n, -code - 201, dof - 1]
        return np.outer(phi, phi).astype(complex)
    if code >= 1:
        return np.diag(prob.grid[:n, dof - 1] ** code).astype(complex)
    if code == -1:
        return -1j * prob.d1[:n, :n, dof - 1]
    if code == -2:
        return prob.d2[:n, :n, dof - 1].astype(complex)
    if code == -3:
        return -1j * prob.xd1[:n, :n, dof - 1]
    raise ValueError(code)


def mode_grid_op(prob, m: int, codes) -> np.ndarray:
    """Operator on the product grid of mode m (0-based), first DOF fastest."""
    nmd = int(prob.nmodeDOF[m])
    out = np.ones((1, 1), dtype=complex)
    for k in range(nmd):
        dof = int(prob.iDOF[k, m])
        out = np.kron(prim_matrix(prob, dof, codes[k]), out)
    return out


def spf_block(prob, psi, m: int, s: int) -> np.ndarray:
    ip, _ = prob.ipsiSPF()
    ngpm = prob.ngpm()
    n = int(prob.nSPF[m, s])
    start = ip[m, s] - 1
    return psi[start:start + ngpm[m] * n].reshape(ngpm[m], n, order="F")


def term_codes(prob, m: int, t: int):
    nmd = int(prob.nmodeDOF[m])
    return [int(prob.iOpDof[prob.iDOF[k, m] - 1, t]) for k in range(nmd)]


def is_identity(prob, m: int, t: int) -> bool:
    a, b = int(prob.lhs_el[t]), int(prob.rhs_el[t])
    if not prob.MultiSet[m]:
        a = b = 1
    return a == b and all(c == 0 for c in term_codes(prob, m, t))


def term_mode_matrix(prob, psi, m: int, t: int) -> np.ndarray:
    """<phi_lhs | O_t^(m) | phi_rhs> on mode m (identity when skipped)."""
    a, b = int(prob.lhs_el[t]) - 1, int(prob.rhs_el[t]) - 1
    if is_identity(prob, m, t):
        return np.eye(int(prob.nSPF[m, b]), dtype=complex)
    pa = spf_block(prob, psi, m, a)
    pb = spf_block(prob, psi, m, b)
    return pa.conj().T @ mode_grid_op(prob, m, term_codes(prob, m, t)) @ pb


def kron_modes(mats) -> np.ndarray:
    """Kronecker product acting on vec(A) with mode 0 fastest."""
    out = np.ones((1, 1), dtype=complex)
    for M in mats:
        out = np.kron(M, out)
    return out


def full_hamiltonian(prob, psi) -> np.ndarray:
    off = prob.AStateOffset()
    nconf = prob.nAConf()
    N = prob.nAConftot
    H = np.zeros((N, N), dtype=complex)
    for t in range(prob.nHam):
        a, b = int(prob.lhs_el[t]) - 1, int(prob.rhs_el[t]) - 1
        mats = [term_mode_matrix(prob, psi, m, t) for m in range(prob.nmodes)]
        H[off[a]:off[a] + nconf[a], off[b]:off[b] + nconf[b]] += prob.ae[t] * kron_modes(mats)
    return H


def state_tensor(prob, A, s: int) -> np.ndarray:
    off = prob.AStateOffset()[s]
    n = [int(prob.nSPF[m, s]) for m in range(prob.nmodes)]
    return A[off:off + int(np.prod(n))].reshape(n, order="F")


def density(prob, A, m: int, s: int) -> np.ndarray:
    T = state_tensor(prob, A, s)
    T = np.moveaxis(T, m, 0).reshape(T.shape[m], -1)
    return T.conj() @ T.T


def apply_mode(T: np.ndarray, M: np.ndarray, m: int) -> np.ndarray:
    return np.moveaxis(np.tensordot(M, T, axes=([1], [m])), 0, m)


def mean_fields(prob, A, psi):
    """MF[(m, key)] with key = operator labels+codes, summed over terms."""
    mf = {}
    for t in range(prob.nHam):
        a, b = int(prob.lhs_el[t]) - 1, int(prob.rhs_el[t]) - 1
        Ta = state_tensor(prob, A, a)
        Tb = state_tensor(prob, A, b)
        mats = [term_mode_matrix(prob, psi, m, t) for m in range(prob.nmodes)]
        for m in range(prob.nmodes):
            if is_identity(prob, m, t):
                continue
            X = Tb
            for q in range(prob.nmodes):
                if q != m:
                    X = apply_mode(X, mats[q], q)
            La = np.moveaxis(Ta, m, 0).reshape(Ta.shape[m], -1)
            Rb = np.moveaxis(X, m, 0).reshape(X.shape[m], -1)
            G = prob.ae[t] * (La.conj() @ Rb.T)
            key = (m, op_key(prob, m, t))
            mf[key] = mf.get(key, 0) + G
    return mf


def op_key(prob, m: int, t: int):
    if prob.MultiSet[m]:
        lab = (int(prob.lhs_el[t]), int(prob.rhs_el[t]))
    else:
        lab = (1, 1)
    return lab + tuple(term_codes(prob, m, t))


def basis(prob):
    """Operator basis per mode: ordered unique keys (hamilton.f90:91-147)."""
    out = []
    for m in range(prob.nmodes):
        keys = []
        for t in range(prob.nHam):
            k = op_key(prob, m, t)
            if k not in keys:
                keys.append(k)
        out.append(keys)
    return out


def reg_inv(A: np.ndarray, eps: float, kind: str) -> np.ndarray:
    w, U = np.linalg.eigh(A)
    if kind == "tikh":
        f = w / (w ** 2 + eps ** 2)
    else:
        w = np.where(w < eps, 0.5 * (eps + w ** 2 / eps), w)
        f = 1.0 / w
    return (U * f) @ U.conj().T


def lanczos_expm(H: np.ndarray, b: np.ndarray, dt: float) -> np.ndarray:
    v = scipy.linalg.expm(-1j * dt * H) @ b
    return v / np.linalg.norm(v)
