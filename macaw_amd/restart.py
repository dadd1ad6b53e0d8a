"""Macaw restart files and the INIT-WF ``restart`` initial state.

File format (output.f90:230-241): Fortran unformatted sequential, as the
reference writes it with gfortran and ``-fdefault-integer-8
-freal-4-real-8`` (Makefile:18): every record is framed by a 4-byte
little-endian byte count before and after it; integers are 8 bytes, reals
8 bytes, complex 16 bytes.  Records, in order:

1. ``nstates, nmodes``
2. ``nSPF(1:nmodes, s)``, one record per electronic state
3. ``time``
4. ``AVector`` (every state's coefficients, first mode fastest)
5. ``psi`` (the SPF vector)
6. ``GWPa(ndofGWP, nS00max, ndofGWP)``, the Gaussian width parameters
   (element (j, l, i): DOF j of the i-th Gaussian mode, SPF l; empty when
   there are no Gaussian modes).

Records longer than 2^31 - 9 bytes would need gfortran's subrecords; they
are refused here.

Reading a restart into a problem follows init_calc.f90:290-481: the file's
states are associated with the problem's (``associated[s]``); grid-mode SPFs
are copied up to the smaller count, and further SPFs are made from the
previous one multiplied by 0.001 (i - n/2) on the grid and orthonormalised
twice against those before; Gaussian modes copy their parameters (their SPF
counts must agree); a coefficient is copied when its configuration exists in
the associated state of the file and the state is the chosen one (or
``iel_start`` = 0), else it is zero.  init_calc.f90:215-220 then divides A by
the square root of the wavefunction's norm.
"""
from __future__ import annotations

import struct

import numpy as np

_MAXREC = 2**31 - 9


class RestartError(ValueError):
    pass


def _write_record(fh, payload: bytes) -> None:
    if len(payload) > _MAXREC:
        raise RestartError(f"record of {len(payload)} bytes needs Fortran subrecords (not supported)")
    mark = struct.pack("<i", len(payload))
    fh.write(mark)
    fh.write(payload)
    fh.write(mark)


def _read_record(fh) -> bytes:
    head = fh.read(4)
    if len(head) != 4:
        raise RestartError("unexpected end of file")
    (n,) = struct.unpack("<i", head)
    if n < 0:
        raise RestartError("Fortran subrecords are not supported")
    payload = fh.read(n)
    tail = fh.read(4)
    if len(payload) != n or len(tail) != 4 or struct.unpack("<i", tail)[0] != n:
        raise RestartError("corrupt record markers")
    return payload


def _gwp_modes(prob):
    return [m for m in range(prob.nmodes) if prob.gwp_mode(m)]


def gwpa_array(prob) -> np.ndarray:
    """The reference's GWPa(ndofGWP, nS00max, ndofGWP) from the problem's
    per-mode width table (init_calc.f90:69-83, 1074)."""
    gm = _gwp_modes(prob)
    ndof_gwp = int(sum(int(prob.nmodeDOF[m]) for m in gm))
    ns00 = max([int(prob.nSPF[m, :].max()) for m in gm], default=0)
    g = np.zeros((ndof_gwp, ns00, ndof_gwp), order="F")
    for i, m in enumerate(gm):
        nmd = int(prob.nmodeDOF[m])
        nl = int(prob.nSPF[m, :].max())
        g[:nmd, :nl, i] = prob.gwp_a[:nmd, :nl, m]
    return g


def write_restart(path: str, prob, time: float, A: np.ndarray, psi: np.ndarray) -> None:
    """Write the state (A, psi) at `time` (atomic units) as Macaw's `restart`."""
    A = np.ascontiguousarray(A, np.complex128)
    psi = np.ascontiguousarray(psi, np.complex128)
    if A.size != prob.nAConftot or psi.size != prob.dimpsi:
        raise RestartError("A or psi does not match the problem's layout")
    with open(path, "wb") as fh:
        _write_record(fh, struct.pack("<2q", prob.nstates, prob.nmodes))
        for s in range(prob.nstates):
            _write_record(fh, np.asarray(prob.nSPF[:, s], dtype="<i8").tobytes())
        _write_record(fh, struct.pack("<d", float(time)))
        _write_record(fh, A.astype("<c16").tobytes())
        _write_record(fh, psi.astype("<c16").tobytes())
        _write_record(fh, gwpa_array(prob).astype("<f8").tobytes(order="F"))


def write_psi_record(fh, time: float, A: np.ndarray, psi: np.ndarray) -> None:
    """One record of Macaw's `psi` trajectory file: time, AVector, psi in a
    single unformatted record (propagator.f90:40-45, output.f90:248-252)."""
    _write_record(fh, struct.pack("<d", float(time)) + np.ascontiguousarray(A, np.complex128).astype("<c16").tobytes()
                  + np.ascontiguousarray(psi, np.complex128).astype("<c16").tobytes())


def read_psi_file(path: str, nconf: int, dimpsi: int) -> list:
    """Records (time, A, psi) of a `psi` trajectory file."""
    import os
    out = []
    size = os.path.getsize(path)
    with open(path, "rb") as fh:
        while fh.tell() < size:
            rec = _read_record(fh)
            if len(rec) != 8 + 16 * (nconf + dimpsi):
                raise RestartError("psi record does not match the layout")
            (t,) = struct.unpack_from("<d", rec, 0)
            A = np.frombuffer(rec, dtype="<c16", count=nconf, offset=8).astype(np.complex128)
            psi = np.frombuffer(rec, dtype="<c16", count=dimpsi, offset=8 + 16 * nconf).astype(np.complex128)
            out.append((t, A, psi))
    return out


def read_restart(path: str) -> dict:
    """Records of a restart file: nstates, nmodes, nSPF[nmodes, nstates],
    time, A, psi, GWPa (flat, Fortran order)."""
    def unpack(fmt, rec, what):
        if len(rec) != struct.calcsize(fmt):
            raise RestartError(f"{what} record has the wrong length")
        return struct.unpack(fmt, rec)

    with open(path, "rb") as fh:
        ns, nm = unpack("<2q", _read_record(fh), "nstates/nmodes")
        if ns < 1 or nm < 1:
            raise RestartError(f"invalid header: {ns} states, {nm} modes")
        rows = []                                   # sized by the file, not by the header
        for s in range(ns):
            v = np.frombuffer(_read_record(fh), dtype="<i8")
            if v.size != nm:
                raise RestartError("nSPF record has the wrong length")
            rows.append(v.astype(np.int64))
        nspf = np.stack(rows, axis=1)
        if nspf.min() < 1:
            raise RestartError("nSPF record holds a count below 1")
        (time,) = unpack("<d", _read_record(fh), "time")
        A = np.frombuffer(_read_record(fh), dtype="<c16").astype(np.complex128)
        psi = np.frombuffer(_read_record(fh), dtype="<c16").astype(np.complex128)
        gwpa = np.frombuffer(_read_record(fh), dtype="<f8").astype(np.float64)
    if A.size != int(sum(np.prod(nspf[:, s]) for s in range(ns))):
        raise RestartError("AVector record does not match the SPF counts")
    return {"nstates": ns, "nmodes": nm, "nSPF": nspf, "time": time, "A": A, "psi": psi, "GWPa": gwpa}


def _psi_layout(prob, nspf):
    """1-based SPF block starts (init_calc.f90:339-364) for SPF counts `nspf`
    of the problem's modes; returns (ip[j, m, s], dimpsi)."""
    ns = nspf.shape[1]
    npar = prob.npar()
    ip = np.zeros((int(nspf.max()), prob.nmodes, ns), dtype=np.int64)
    dim = 0
    for s in range(ns):
        for m in range(prob.nmodes):
            if not prob.MultiSet[m] and s > 0:
                ip[:, m, s] = ip[:, m, 0]
                continue
            for k in range(int(nspf[m, s])):
                ip[k, m, s] = dim + 1
                dim += int(npar[m])
    return ip, dim


def trace_norm(prob, A: np.ndarray, psi: np.ndarray) -> float:
    """<Psi|Psi> = sum_s A_s^H (prod_m S_m^(s)) A_s with S the SPF overlap
    matrices of each mode in state s (grid modes: the SPF Gram matrix;
    Gaussian modes: normalised Gaussian overlaps)."""
    ip, _ = prob.ipsiSPF()
    off = prob.AStateOffset()
    npar = prob.npar()
    total = 0.0
    for s in range(prob.nstates):
        dims = [int(prob.nSPF[m, s]) for m in range(prob.nmodes)]
        a = A[off[s]: off[s] + int(np.prod(dims))].reshape(dims, order="F")
        sa = a
        for m in range(prob.nmodes):
            n = dims[m]
            start = int(ip[m, s]) - 1
            if prob.gwp_mode(m):
                S = gaussian_overlap(prob, m, psi[start:start + n * int(npar[m])].reshape(int(npar[m]), n, order="F"))
            else:
                blk = psi[start:start + n * int(npar[m])].reshape(int(npar[m]), n, order="F")
                S = blk.conj().T @ blk
            sa = np.moveaxis(np.tensordot(S, sa, axes=([1], [m])), 0, m)
        total += float(np.vdot(a.reshape(-1, order="F"), sa.reshape(-1, order="F")).real)
    return total


def gaussian_overlap(prob, m: int, xi: np.ndarray) -> np.ndarray:
    """Overlaps of the normalised Gaussians exp(a x^2 + xi x) of mode m (a < 0
    real per DOF and SPF, xi[d, l]): per DOF,
    int g_k* g_l dx = sqrt(pi / -(a_k + a_l)) exp(-(xi_k* + xi_l)^2 / (4 (a_k + a_l))),
    divided by the square roots of the two self-overlaps (evaluated in
    logarithms, so large xi do not overflow)."""
    nmd, n = xi.shape
    a = prob.gwp_a[:nmd, :n, m]
    logS = np.zeros((n, n), dtype=np.complex128)
    for d in range(nmd):
        ak, al = a[d][:, None], a[d][None, :]
        xk, xl = xi[d][:, None], xi[d][None, :]
        lkl = -(np.conj(xk) + xl) ** 2 / (4.0 * (ak + al)) + 0.5 * np.log(np.pi / -(ak + al))
        lkk = -(2.0 * xk.real) ** 2 / (8.0 * ak) + 0.5 * np.log(np.pi / (-2.0 * ak))
        lll = -(2.0 * xl.real) ** 2 / (8.0 * al) + 0.5 * np.log(np.pi / (-2.0 * al))
        logS = logS + lkl - 0.5 * (lkk + lll)
    return np.exp(logS)


def restart_state(prob, data: dict, iel_start: int, associated) -> tuple:
    """Initial (A, psi) of `prob` from restart records `data`
    (init_calc.f90:290-481, then the normalisation of :215-220).  The
    problem's Gaussian widths are replaced by the file's.  `associated[s]`
    is the 1-based file state for problem state s + 1."""
    ns, nm = prob.nstates, prob.nmodes
    if data["nmodes"] != nm:
        raise RestartError(f"the wavefunction on the file should consist of {nm} modes instead of {data['nmodes']}")
    assoc = [int(a) - 1 for a in associated]
    if len(assoc) != ns or min(assoc) < 0 or max(assoc) >= data["nstates"]:
        raise RestartError("electronic state associations must name a state of the file for every state")
    nspf_t = data["nSPF"]
    gm = _gwp_modes(prob)
    for s in range(ns):
        for m in gm:
            if nspf_t[m, assoc[s]] != prob.nSPF[m, s]:
                raise RestartError(f"mode {m + 1}: for GWPs the number of SPFs should be the same")
    ip_t, dim_t = _psi_layout(prob, nspf_t)
    if data["psi"].size != dim_t:
        raise RestartError("psi record does not match the SPF counts and grids")
    psi_t = data["psi"]
    if gm:
        g = data["GWPa"]
        ndof_gwp = int(sum(int(prob.nmodeDOF[m]) for m in gm))
        ns00 = g.size // (ndof_gwp * ndof_gwp) if ndof_gwp else 0
        g = g.reshape((ndof_gwp, ns00, ndof_gwp), order="F")
        for i, m in enumerate(gm):
            nmd = int(prob.nmodeDOF[m])
            nl = min(ns00, prob.gwp_a.shape[1])
            prob.gwp_a[:nmd, :nl, m] = g[:nmd, :nl, i]
    ip, dimpsi = prob.ipsiSPF()
    npar = prob.npar()
    psi = np.zeros(dimpsi, dtype=np.complex128)
    for s in range(ns):
        st = assoc[s]
        for m in range(nm):
            nG = int(npar[m])
            base = int(ip[m, s]) - 1
            if prob.gwp_mode(m):
                for j in range(int(prob.nSPF[m, s])):
                    src = int(ip_t[j, m, st]) - 1
                    psi[base + j * nG: base + (j + 1) * nG] = psi_t[src: src + nG]
                continue
            k = min(int(prob.nSPF[m, s]), int(nspf_t[m, st]))
            for j in range(k):
                src = int(ip_t[j, m, st]) - 1
                psi[base + j * nG: base + (j + 1) * nG] = psi_t[src: src + nG]
            ramp = 0.001 * (np.arange(1, nG + 1) - nG // 2)
            for j in range(k, int(prob.nSPF[m, s])):
                prev = psi[base + (j - 1) * nG: base + j * nG]
                phi = ramp * prev
                for _ in range(2):
                    for i2 in range(j):
                        v = psi[base + i2 * nG: base + (i2 + 1) * nG]
                        phi = phi - v * np.vdot(v, phi)
                    phi = phi / np.sqrt(np.vdot(phi, phi).real)
                psi[base + j * nG: base + (j + 1) * nG] = phi
    # coefficients: configurations present in both layouts
    A = np.zeros(prob.nAConftot, dtype=np.complex128)
    off = prob.AStateOffset()
    off_t = np.concatenate([[0], np.cumsum([int(np.prod(nspf_t[:, t])) for t in range(data["nstates"])])])
    for s in range(ns):
        if iel_start not in (0, s + 1):
            continue
        st = assoc[s]
        dims = tuple(int(prob.nSPF[m, s]) for m in range(nm))
        dims_t = tuple(int(nspf_t[m, st]) for m in range(nm))
        multi = np.unravel_index(np.arange(int(np.prod(dims))), dims, order="F")
        inside = np.ones(int(np.prod(dims)), dtype=bool)
        for m in range(nm):
            inside &= multi[m] < dims_t[m]
        lin_t = np.ravel_multi_index(tuple(mi[inside] for mi in multi), dims_t, order="F")
        a = np.zeros(int(np.prod(dims)), dtype=np.complex128)
        a[inside] = data["A"][off_t[st] + lin_t]
        A[off[s]: off[s] + a.size] = a
    cnorm = trace_norm(prob, A, psi)
    if not cnorm > 0.0:
        raise RestartError("the restart state has zero norm on the chosen states")
    return A / np.sqrt(cnorm), psi
