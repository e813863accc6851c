"""input.ini handling and the normalisation chain of the reference's main().

parse_ini      follows INIParser::parse (iniparser.cpp:4-50): lines starting
               with ';' or '#' and empty lines are skipped, "[name]" opens a
               section, "[species]" lines are kept in order, other lines are
               "key = value" trimmed of spaces and tabs.  Inline comments are
               not stripped (getInt/getDouble parse the numeric prefix, as
               std::stoi/std::stod do).
derive         follows main.cpp:151-261 (densities, Debye lengths, plasma
               frequencies, dx, DT, spwt) and Domain::set_normscheme
               (domain.cpp:123-183), returning the C-ABI descriptors.
"""
import math
import re
from dataclasses import dataclass, field
from typing import Dict, List

from . import _lib

# include/domain.h:204-215
EPS_0 = 8.85418782e-12
QE = 1.602176565e-19
AMU = 1.660538921e-27
ME = 9.10938215e-31
K_b = 1.380648e-23
PI = 3.141592653
EV_to_K = QE / K_b
eV = 1.602176565e-19

_NUM = re.compile(r"^\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?|inf|nan)", re.I)
_INT = re.compile(r"^\s*[+-]?\d+")


def stod(s: str) -> float:
    m = _NUM.match(s)
    if not m:
        raise ValueError(f"stod: no conversion of {s!r}")
    return float(m.group(0))


def stoi(s: str) -> int:
    m = _INT.match(s)
    if not m:
        raise ValueError(f"stoi: no conversion of {s!r}")
    return int(m.group(0))


def string_to_bool(s: str) -> bool:
    """include/extrafun.h:6-17"""
    return s in ("true", "True", "T", "t", "1", "yes", "Yes", "Y", "y")


def parse_ini_text(text: str) -> Dict[str, Dict[str, str]]:
    sections: Dict[str, Dict[str, str]] = {}
    cur = ""
    for line in text.split("\n"):
        if not line or line[0] in ";#":
            continue
        if line[0] == "[" and line[-1] == "]":
            cur = line[1:-1]
            continue
        sec = sections.setdefault(cur, {})
        if cur == "species":
            sec[str(len(sec))] = line
        else:
            if "=" not in line:
                continue
            k, v = line.split("=", 1)
            if v == "":
                continue  # getline(iss, value) fails on an empty remainder
            sections[cur][k.strip(" \t")] = v.strip(" \t")
    return sections


def parse_ini(path: str) -> Dict[str, Dict[str, str]]:
    with open(path) as f:
        return parse_ini_text(f.read())


def loadtypeextract(s: str):
    """INIParser::loadtypeextract (iniparser.cpp:100-113): "sin(3)" -> ("sin", 3)."""
    if "(" in s and ")" in s.split("(", 1)[1]:
        a, b = s.split("(", 1)
        return a, stoi(b.split(")", 1)[0])
    return s, 0


@dataclass
class SpeciesSpec:
    name: str
    mass: float
    N: int
    temp: float
    sign: int
    frac: float
    vsx: float = 0.0
    vsy: float = 0.0
    load: str = "random"


@dataclass
class Setup:
    domain: "_lib.DomainDesc"
    species: List["_lib.SpeciesDesc"]
    info: Dict[str, float] = field(default_factory=dict)
    specs: List[SpeciesSpec] = field(default_factory=list)


def derive(Nx, Ny, x0, species: List[SpeciesSpec], density, DT_coeff, norm_scheme=1, L_scale=0.005,
           T_scale="omegape", bc="pbc", shape="CIC", solver=_lib.SOLVER_PCG, poisson_op=None, tolerance=1e-8,
           tol_relative=0, max_iteration=10000, precision=_lib.FP64, sort_interval=0, seed=1, ionfixed=0,
           capacity_factor=1.0, sor_omega=1.4, precond=0) -> Setup:
    """main.cpp:151-261 for the given ini values."""
    nsp = len(species)
    if nsp < 2:
        raise ValueError("the reference's normalisation needs at least two species (main.cpp:159-160)")
    k = 0.0
    for sp in species:
        k += (-sp.sign) * sp.frac
    normden = [0.0] * nsp
    normden[1] = density
    normden[0] = density / k
    for i in range(2, nsp):
        normden[i] = species[i].frac * normden[0]
    t0, t1 = species[0].temp, species[1].temp
    m0, m1 = species[0].mass, species[1].mass
    LDe = math.sqrt((EPS_0 * K_b * t0 * EV_to_K) / (normden[0] * QE * QE))
    wpe = math.sqrt((normden[0] * QE * QE) / (m0 * EPS_0))
    wpi = math.sqrt((normden[1] * QE * QE) / (m1 * EPS_0))
    LDi = math.sqrt((EPS_0 * K_b * t1 * EV_to_K) / (normden[1] * QE * QE))
    CS = math.sqrt(t0 * K_b * EV_to_K / m1)
    if T_scale == "omegape":
        time_scale = wpe
    elif T_scale == "omegapi":
        time_scale = wpi
    else:
        time_scale = stod(str(T_scale))
    if norm_scheme in (2, 4):
        stepSize = LDi
        dx = stepSize / LDi
        dy = stepSize / LDi
        DT = DT_coeff * (1.0 / wpi)
        DT = wpi * DT
        L = LDi
        W = wpi if norm_scheme == 2 else wpe
    elif norm_scheme in (1, 3):
        stepSize = LDe
        dx = stepSize / LDe
        dy = stepSize / LDe
        DT = DT_coeff * (1.0 / wpe)
        DT = wpe * DT
        L, W = LDe, wpe
    elif norm_scheme == 5:
        stepSize = L_scale
        dx = stepSize / L_scale
        dy = stepSize / L_scale
        DT = DT_coeff * (1.0 / time_scale)
        DT = time_scale * DT
        L, W = L_scale, time_scale
    else:
        raise ValueError("Invalid normscheme flag")
    vel_norm = L * W  # domain.cpp:125-146: the trailing else wins
    nx, ny = Nx + 1, Ny + 1
    Lx = (nx - 1) * dx
    Ly = (ny - 1) * dy
    if poisson_op is None:
        poisson_op = (_lib.OP_DIRICHLET_WALL if solver in (_lib.SOLVER_GS, _lib.SOLVER_GS_BOLTZMANN)
                      else _lib.OP_REF_GHOST)
    dom = _lib.DomainDesc(x0=x0, y0=x0, dx=dx, dy=dy, nx=nx, ny=ny, L=L, W=W, LDe=LDe, wpe=wpe, vel_norm=vel_norm,
                          density=density, DT=DT, bc=_lib.BC[bc], shape=_lib.SHAPE[shape], solver=solver,
                          poisson_op=poisson_op, tolerance=tolerance, tol_relative=tol_relative,
                          max_iteration=max_iteration, sor_omega=sor_omega, precision=precision,
                          sort_interval=sort_interval, seed=seed, precond=precond)
    descs = []
    for i, sp in enumerate(species):
        spwt = normden[i] * Lx * Ly * L * L / sp.N
        cap = int(math.ceil(sp.N * capacity_factor))
        d = _lib.SpeciesDesc(name=sp.name.encode()[:31], mass=sp.mass, charge=sp.sign * QE, spwt=spwt, temp=sp.temp,
                             numparticle=sp.N, capacity=cap, vsx=sp.vsx, vsy=sp.vsy, fract_den=sp.frac,
                             pushed=0 if (sp.name == "ion" and ionfixed == 1) else 1)
        descs.append(d)
    info = dict(LDe=LDe, wpe=wpe, LDi=LDi, wpi=wpi, CS=CS, L=L, W=W, vel_norm=vel_norm, dx=dx, dy=dy, DT=DT, Lx=Lx,
                Ly=Ly, k=k, normden=normden)
    return Setup(domain=dom, species=descs, info=info, specs=list(species))


def setup_from_ini(sections, precision=_lib.FP64, seed=1, capacity_factor=1.0):
    """Reads the keys main.cpp:42-149 reads (missing keys raise KeyError, as
    map::at throws) and derives the C-ABI descriptors."""
    g = sections
    Nx = stoi(g["domain"]["Nx"])
    Ny = stoi(g["domain"]["Ny"])
    x0 = stod(g["domain"]["x0"])
    NUM_TS = stoi(g["time"]["NUM_TS"])
    DT_coeff = stod(g["time"]["DT_coeff"])
    shape = g["simulation"]["shapefunction"]
    den = stod(g["simulation"]["density"])
    bc = g["simulation"]["bc"]
    ionfixed = stoi(g["simulation"]["ionfixed"])
    norm_scheme = stoi(g["normalization"]["norm_scheme"])
    L_scale = stod(g["normalization"]["lenght_scale"])
    T_scale = g["normalization"]["time_scale"]
    solvertype = g["solver"]["solvertype"]
    tol = stod(g["solver"]["tolerance"])
    maxit = stoi(g["solver"]["max_iteration"])
    solver = {"gs": _lib.SOLVER_GS, "cg": _lib.SOLVER_PCG, "iter": _lib.SOLVER_PCG, "iterative": _lib.SOLVER_PCG,
              "pcg_mixed": _lib.SOLVER_PCG_MIXED, "spectral": _lib.SOLVER_SPECTRAL,
              "gshybrid": _lib.SOLVER_GS_BOLTZMANN, "hybrid": _lib.SOLVER_NEWTON_CG}[solvertype]
    specs = []
    for key in sorted(g.get("species", {}), key=int):
        tok = g["species"][key].split(",")
        if len(tok) != 9:
            continue
        specs.append(SpeciesSpec(tok[0], stod(tok[1]), stoi(tok[2]), stod(tok[3]), stoi(tok[4]), stod(tok[5]),
                                 stod(tok[6]), stod(tok[7]), tok[8]))
    st = derive(Nx, Ny, x0, specs, den, DT_coeff, norm_scheme, L_scale, T_scale, bc, shape, solver,
                tolerance=tol, max_iteration=maxit, precision=precision, seed=seed, ionfixed=ionfixed,
                capacity_factor=capacity_factor)
    st.info["NUM_TS"] = NUM_TS
    coll = g.get("collision", {})
    st.info["collision"] = {k: coll[k] for k in coll}
    return st
