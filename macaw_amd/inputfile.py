"""Reader for the Macaw input-file and operator-file formats (the subset the
CMF2 engine runs).

Keyword grammar follows Macaw's ``read_input`` (source/input.f90:30-426):
case-insensitive ``name = value`` lines, ``#`` comments, the ``SPF`` and
``GRID`` blocks, ``INIT-WF ... END-INIT-WF`` and ``integration = CMF2``
followed by the two integrator lines.  Defaults are those of
source/input.f90:44-82.  Operator lines follow ``read_declare_ham``
(source/input.f90:812-1144): ``coeff | elec | dof op | dof op ...`` up to a
line starting with ``----``.

Times are converted to atomic units with Macaw's constant
``au2fs = 0.02418884326505`` (source/globals.f90:17).
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

AU2EV = 27.2113961
AU2FS = 0.02418884326505

# primitive operator codes (source/input.f90:1098-1137, hamilton.f90:24-31)
OP_CODES = {"p": -1, "dq^2": -2, "qp": -3, "q": 1}
for _n in range(2, 11):
    OP_CODES[f"q^{_n}"] = _n


class InputError(ValueError):
    """Raised for input the reference rejects with ``stop``."""


# Parametrised one-DOF operators (input.f90:950-1090; OpParamFunc of
# hamilton.f90:212-300): 0 Gaussian exp(p1 x + p2 x^2), 1 interpolation of a
# (x, y) table, 2 indicator of [p1 - p2/2, p1 + p2/2], 3 Hermite function
# x^r exp(p1 x + p2 x^2), 4 sin(p1 x)^p2, 5 cos(p1 x)^p2, 6 projector on
# eigenstate r of -p2 d^2/dx^2 + V(x) with V interpolated from a file.
# Operator codes: 200 + p (diagonal functions) and -200 - p (projector),
# p = 1-based index in the table of the problem.
@dataclass
class OpParam:
    func: int
    p1: float = 0.0
    p2: float = 0.0
    p3: float = 0.0
    file: str = ""


class OpParams:
    """The parametrised operators of a problem's operator files, in reading
    order (the Hamiltonian's, then the expectation operators'); an
    interpolation file read twice is one entry (input.f90:969-979)."""

    def __init__(self):
        self.items = []

    def add(self, op: OpParam) -> int:
        if op.func == 1:
            for i, o in enumerate(self.items):
                if o.func == 1 and o.file == op.file:
                    return i + 1
        self.items.append(op)
        return len(self.items)


def _param_operator(name: str, registry: OpParams, basedir: str):
    """Code of a parametrised operator name, or None (input.f90:950-1090)."""
    def after(key):
        return name[name.index(key) + len(key):].strip()

    def nums(text, n, default=None):
        toks = text.replace(",", " ").split()
        if not toks and default is not None:
            return default
        if len(toks) < n:
            raise InputError(f"Operator {name}: expected {n} parameters")
        return [_fortran_float(t) for t in toks[:n]]

    def fname(text):
        t = text.strip().split()[0] if text.strip() else ""
        t = t.strip("'\"")
        return t if os.path.isabs(t) else os.path.join(basedir, t)

    if "gau" in name:
        p1, p2 = nums(after("gau"), 2)
        return 200 + registry.add(OpParam(0, p1, p2))
    if "interp" in name:
        return 200 + registry.add(OpParam(1, file=fname(after("interp"))))
    if "square" in name:
        p1, p2 = nums(after("square"), 2)
        return 200 + registry.add(OpParam(2, p1, p2))
    if "hermite" in name:
        r, p1, p2 = nums(after("hermite"), 3)
        return 200 + registry.add(OpParam(3, p1, p2, float(int(r))))
    if "sin^-2" in name:
        raise InputError("ERROR: sin^-2 operator allowed only for Leg-DVR")
    if "sin" in name:
        p1, p2 = nums(after("sin"), 2, default=[1.0, 1.0])
        return 200 + registry.add(OpParam(4, p1, p2))
    if "cos" in name:
        p1, p2 = nums(after("cos"), 2, default=[1.0, 1.0])
        return 200 + registry.add(OpParam(5, p1, p2))
    if "eig_proj" in name:
        rest = after("eig_proj").split()
        if len(rest) < 3:
            raise InputError(f"Operator {name}: expected the potential file, the kinetic prefactor and the level")
        return -200 - registry.add(OpParam(6, 0.0, _fortran_float(rest[1]), float(int(rest[2])), fname(rest[0])))
    return None


@dataclass
class SPFDecl:
    multiset: bool
    dofs: list          # 1-based DOF ids of the logical mode
    nspf: list          # per electronic state


@dataclass
class GridDecl:
    kind: str           # 'HO' (x0, m w), 'sin' (x0 = first, mw = last point), or 'gwp' (plot grid)
    npts: int
    x0: float
    mw: float


@dataclass
class InitSPF:
    kind: str           # 'gauss' or 'gwp'
    x0: float
    p0: float
    dx: float
    dist: int = 0       # gwp: lattice type (0 = neighbour shells, 1 = gwpgrid_<mode> file)
    ovl: float = 0.7    # gwp: overlap between neighbouring Gaussians
    scw: float = 1.0    # gwp: width scaling per lattice step


@dataclass
class InputDeck:
    nstates: int = 1
    ndof: int = -1
    nmodes: int = 0
    spf: list = field(default_factory=list)
    grid: dict = field(default_factory=dict)
    hamiltonian: str = ""
    tinit: float = 0.0
    tfinal: float = 0.0
    tout: float = 1.0 / AU2FS
    eps_rho: float = 1e-8
    eps_c: float = 1e-8
    eps_ovl: float = 1e-8
    eps_int: float = 1e-8
    eps_a: float = 1e-8
    eps_psi: float = 1e-8
    min_stepsize: float = 1e-9 / AU2FS
    int_type: int = 1
    int_type_a: int = 0
    int_type_psi: int = 0
    steps: bool = False
    normalize: int = 0
    densmat: list = field(default_factory=list)   # modes (1-based) of `densmat = m`
    density: bool = False
    wp_ovl: bool = False
    expect: list = field(default_factory=list)   # operator files of `expect = file`
    init_state: int = 1
    init_spf: dict = field(default_factory=dict)
    # INIT-WF restart: file, chosen state (0 = all), file state per state
    restart_file: str = ""
    restart_state: int = 0
    restart_assoc: list = field(default_factory=list)
    basedir: str = "."


def _strip(line: str) -> str:
    i = line.find("#")
    return line[:i] if i >= 0 else line


def _fortran_float(s: str) -> float:
    return float(s.strip().lower().replace("d", "e"))


def parse_input(path: str) -> InputDeck:
    """Parse a Macaw input file (source/input.f90:113-354)."""
    deck = InputDeck(basedir=os.path.dirname(os.path.abspath(path)))
    with open(path) as fh:
        lines = fh.read().splitlines()
    i = 0
    n = len(lines)
    while i < n:
        line = _strip(lines[i])
        i += 1
        if line.strip() == "INIT-WF":
            block = []
            while i < n and _strip(lines[i]).strip() != "END-INIT-WF":
                block.append(_strip(lines[i]))
                i += 1
            if i >= n:
                raise InputError("Closing instruction END-INIT-WF missing")
            i += 1
            _parse_init_wf(deck, block)
            continue
        if "=" in line:
            name, value = line.split("=", 1)
            name = name.strip().lower()
            value = value.strip()
        else:
            name, value = line.strip().lower(), "blank"
        if name == "":
            continue
        if name == "nstates":
            deck.nstates = int(value)
        elif name == "ndof":
            deck.ndof = int(value)
            if deck.ndof > 99:
                raise InputError("the maximum number of degrees of freedom is 99")
        elif name == "nmodes":
            deck.nmodes = int(value)
        elif name == "spf":
            if deck.nmodes == 0 or deck.ndof <= 0:
                raise InputError("with the keyword SPF, nmodes and ndof must be > 0")
            for _ in range(deck.nmodes):
                deck.spf.append(_parse_spf(_strip(lines[i]), deck.nstates))
                i += 1
        elif name == "grid":
            for _ in range(deck.ndof):
                k, g = _parse_grid(_strip(lines[i]))
                deck.grid[k] = g
                i += 1
        elif name == "hamiltonian":
            deck.hamiltonian = value
        elif name == "expect":
            # an operator whose expectation value is written at every output
            # time (input.f90:196-215; output.f90:449-498)
            deck.expect.append(value)
        elif name == "tinit":
            deck.tinit = _fortran_float(value) / AU2FS
        elif name == "tfinal":
            deck.tfinal = _fortran_float(value) / AU2FS
        elif name == "tout":
            deck.tout = _fortran_float(value) / AU2FS
        elif name == "eps_rho":
            deck.eps_rho = _fortran_float(value)
        elif name == "eps_c":
            deck.eps_c = _fortran_float(value)
        elif name == "eps_int":
            deck.eps_int = _fortran_float(value)
        elif name == "eps_s":
            deck.eps_ovl = _fortran_float(value)
        elif name == "minstep":
            deck.min_stepsize = _fortran_float(value) / AU2FS
        elif name == "steps":
            deck.steps = True
        elif name == "densmat":
            # input.f90:277-282: one mode per keyword, its density matrices
            # and SPFs written at every output time
            deck.densmat.append(int(value))
        elif name == "density":
            deck.density = True
        elif name == "wp_ovl":
            deck.wp_ovl = True
        elif name == "normalize":
            deck.normalize = int(value)
        elif name == "integration":
            v = value.lower()
            if v.startswith("cmf2"):
                deck.int_type = 10
                a_name, a_eps = lines[i].split()[:2]
                p_name, p_eps = lines[i + 1].split()[:2]
                i += 2
                deck.int_type_a = {"lanczos": 5, "rk45": 1, "dopri": 2, "bs": 3}.get(a_name.lower(), -1)
                if deck.int_type_a < 0:
                    raise InputError(f"Integrator {a_name} not possible for the A coefficients")
                deck.eps_a = _fortran_float(a_eps)
                deck.int_type_psi = {"rk45": 1, "dopri": 2, "bs": 3}.get(p_name.lower(), -1)
                if deck.int_type_psi < 0:
                    raise InputError(f"Integrator {p_name} not possible for the SPFs")
                deck.eps_psi = _fortran_float(p_eps)
            elif v == "rk45":
                deck.int_type = 1
            elif v == "dopri":
                deck.int_type = 2
            else:
                raise InputError(f"unknown integrator: {value}")
        else:
            raise InputError(f"Input instruction: {line.strip()} cannot be understood")
    if not deck.hamiltonian:
        raise InputError("Hamiltonian file not given")
    return deck


def _parse_spf(line: str, nstates: int) -> SPFDecl:
    """One line of the SPF block (source/input.f90:451-512)."""
    if "=" not in line:
        raise InputError(f"Wrong SPF declaration: {line.strip()}")
    sdof, sspf = line.split("=", 1)
    parts = [p.strip() for p in sdof.split(",")]
    multiset = parts[0][:1] != "S"
    dofs = [int(p) for p in parts[1:]]
    nspf = [int(p) for p in sspf.split(",")]
    if not multiset:
        nspf = [nspf[0]] * nstates
    elif len(nspf) < nstates:
        raise InputError("Wrong SPF declaration (SPF)")
    return SPFDecl(multiset, dofs, nspf[:nstates])


def _parse_grid(line: str):
    """One line of the GRID block (source/dvr.f90:32-103, sin, HO and gwp branches)."""
    j = line.find("sin")
    if j >= 0:
        k = int(line[:j])
        toks = line[j + 3:].split()
        return k, GridDecl("sin", int(toks[0]), _fortran_float(toks[1]), _fortran_float(toks[2]))
    j = line.find("gwp")
    if j >= 0:
        k = int(line[:j])
        toks = line[j + 3:].split()
        if not toks:
            return k, GridDecl("gwp", 101, -10.0, 10.0)
        return k, GridDecl("gwp", int(toks[0]), _fortran_float(toks[1]), _fortran_float(toks[2]))
    j = line.find("HO")
    if j < 0:
        raise InputError(f"only sin, HO and gwp grids are supported by this engine: {line.strip()}")
    k = int(line[:j])
    toks = line[j + 2:].split()
    return k, GridDecl("HO", int(toks[0]), _fortran_float(toks[1]), _fortran_float(toks[2]))


def _parse_init_wf(deck: InputDeck, block: list) -> None:
    """INIT-WF block: ``restart`` (source/init_calc.f90:290-333: file name,
    chosen state, then one ``state <-- file state`` line per state) or
    ``single-conf`` (source/init_calc.f90:482-752, el/gauss branches)."""
    rows = [b for b in block if b.strip()]
    if rows and rows[0].strip() == "restart":
        if len(rows) < 3 + deck.nstates:
            raise InputError("INIT-WF restart: expected the file, the initial state and one association per state")
        deck.restart_file = rows[1].strip()
        deck.restart_state = int(rows[2].split()[0])
        assoc = {}
        for row in rows[3:3 + deck.nstates]:
            j = row.find("<--")
            if j < 0:
                raise InputError("Error in INIT-WF: Electronic state associations must be specified")
            assoc[int(row[:j])] = int(row[j + 3:].split()[0])
        if sorted(assoc) != list(range(1, deck.nstates + 1)):
            raise InputError("INIT-WF restart: every electronic state needs one association")
        deck.restart_assoc = [assoc[s] for s in range(1, deck.nstates + 1)]
        return
    if not rows or rows[0].strip() != "single-conf":
        raise InputError("The INIT-WF block must start with the keyword single-conf or restart")
    m = re.search(r"el\s+(\d+)", rows[1])
    if not m:
        raise InputError("single-conf: expected 'el <state>'")
    deck.init_state = int(m.group(1))
    for row in rows[2:]:
        j = row.find("gauss")
        if j >= 0:
            k = int(row[:j])
            x0, p0, dx = (_fortran_float(t) for t in row[j + 5:].split()[:3])
            deck.init_spf[k] = InitSPF("gauss", x0, p0, dx)
            continue
        j = row.find("gwp")
        if j < 0:
            raise InputError(f"The chosen initial SPF is unknown: {row.strip()}")
        # k gwp <dist> x0 p0 dx [ovl = v] [scw = v] (source/init_calc.f90:676-714)
        k = int(row[:j])
        toks = row[j + 3:].split()
        dist = int(toks[0])
        x0, p0, dx = (_fortran_float(t) for t in toks[1:4])
        opts = {}
        for key in ("ovl", "scw"):
            m = re.search(key + r"\s*=\s*([-+0-9.eEdD]+)", row)
            if m:
                opts[key] = _fortran_float(m.group(1))
        deck.init_spf[k] = InitSPF("gwp", x0, p0, dx, dist, opts.get("ovl", 0.7), opts.get("scw", 1.0))


@dataclass
class HamTerm:
    coeff: complex
    lhs: int
    rhs: int
    ops: dict           # dof -> primitive operator code (identity omitted)


def parse_operator_file(path: str, nstates: int, registry: OpParams | None = None) -> list:
    """Hamiltonian terms of an operator file (source/input.f90:752-1144).

    ``1`` and a bare coefficient give one term per electronic state, ``Sa&b``
    with a != b gives both directions with the same coefficient, ``Za&b`` one
    direction.  Parametrised operators are entered in `registry` (a new one
    when None) and get the codes 200 + p / -200 - p.
    """
    registry = registry if registry is not None else OpParams()
    basedir = os.path.dirname(os.path.abspath(path))
    if not os.path.exists(path):
        # source/input.f90:529-536
        raise InputError(f"Input error: Operator file, {path} not found")
    terms = []
    with open(path) as fh:
        for raw in fh:
            if raw[:4] == "----":
                break
            if not raw.strip():
                continue
            line = _strip(raw)
            if not line.strip():
                continue
            fields = line.split("|")
            cstr = fields[0]
            coeff = _fortran_float(cstr.replace("i", "")) * (1j if "i" in cstr else 1.0)
            pairs = []
            if len(fields) == 1 or fields[1].strip() == "1":
                pairs = [(s, s) for s in range(1, nstates + 1)]
            else:
                el = fields[1].strip()
                m = re.match(r"([SZ])\s*(\d+)\s*&\s*(\d+)", el)
                if not m:
                    raise InputError('Only operator labels "1", "Sa&b" or "Za&b" are allowed')
                a, b = int(m.group(2)), int(m.group(3))
                if m.group(1) == "S":
                    pairs = [(a, b)] if a == b else [(a, b), (b, a)]
                else:
                    if a == b:
                        raise InputError('Electronic operators "Za&b" are allowed only between different states')
                    pairs = [(a, b)]
            ops = {}
            for f in fields[2:]:
                f = f.rstrip()
                kdof = int(f[:2]) if f[:2].strip() else int(f.split()[0])
                name = f[2:].strip()
                if kdof == 0:
                    # time-dependent coefficient (input.f90:924-947)
                    raise InputError("time-dependent coefficients (dof 0 'gau') are not supported by this engine")
                if name == "1":
                    continue
                code = _param_operator(name, registry, basedir)
                if code is None:
                    if name not in OP_CODES:
                        raise InputError(f"Operator {name} not implemented")
                    code = OP_CODES[name]
                ops[kdof] = code
            for a, b in pairs:
                terms.append(HamTerm(complex(coeff), a, b, dict(ops)))
    return terms
