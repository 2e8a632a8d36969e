"""EAM and LJ potential set-up on the host.

* ``johnson_cu_funcfl`` tabulates an analytic Johnson-type EAM for Cu
  (exponential density, exponential pair term, Banerjea-Smith embedding) to
  the reference's hybrid funcfl layout: F(rho), then rho(r), then r*phi(r)
  (pair_eam.cpp:197-268, the z2r block used as-is at pair_eam.cpp:486-487).
* ``read_funcfl`` / ``write_funcfl`` read and write that file
  (pair_eam.cpp:197-268: line 1 comment, line 2 ``Z mass``, line 3
  ``nrho drho nr dr cut``, then nrho + 2*nr values, any number per line).
* ``read_setfl`` / ``write_setfl`` read and write LAMMPS setfl files
  (pair_style eam/alloy): 3 comment lines, ``N el1 .. elN``,
  ``nrho drho nr dr cut``, then per element ``Z mass a lattice``, F(rho)
  [nrho] and rho(r) [nr], then r*phi(r) [nr] for every pair i >= j (i, j in
  file order, j fastest), values in any number per line.  The reference
  reads funcfl only (pair_eam.cpp:197-268); its tables come from
  ``cac_eam_tables_from_setfl``.
* ``lj_coeffs`` gives lj1..lj4 exactly as PairLJCut::init_one computes them
  (pair_lj_cut.cpp:140-198, libm ``pow``).

The spline tables themselves (file2array + array2spline) are built by the
native library (``cac_eam_tables_from_funcfl``), so the device and the host
share one set of bits.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

# Johnson, Phys. Rev. B 37, 3924 (1988) functional forms; parameters for Cu:
# cohesive energy E_c, nearest-neighbour distance r_e = a/sqrt(2) at
# a = 3.615 A, alpha = sqrt(9*Omega*B/E_c) with Omega = a^3/4 and
# B = 1.38 Mbar, beta and gamma the density and pair decay exponents,
# phi_e the pair energy at r_e, f_e the density scale.
JOHNSON_CU = dict(E_c=3.54, r_e=3.615 / math.sqrt(2.0), alpha=5.08, beta=5.85,
                  gamma=8.00, phi_e=0.59, f_e=0.30)
CU_MASS = 63.546
CU_A0 = 3.615


def _taper(r, r_s, r_c):
    """C2 switch: 1 below r_s, 0 above r_c, 1-10x^3+15x^4-6x^5 between."""
    x = np.clip((np.asarray(r, dtype=np.float64) - r_s) / (r_c - r_s), 0.0, 1.0)
    return 1.0 - x ** 3 * (10.0 - 15.0 * x + 6.0 * x * x)


@dataclass
class Funcfl:
    """Contents of one hybrid funcfl file (pair_eam.cpp:197-268)."""
    comment: str
    z: int
    mass: float
    nrho: int
    drho: float
    nr: int
    dr: float
    cut: float
    frho: np.ndarray  # [nrho], F at rho = m*drho
    rhor: np.ndarray  # [nr], rho at r = m*dr
    zr: np.ndarray    # [nr], r*phi(r) at r = m*dr


# Taper window [rc - 0.25, rc] = [4.70, 4.95] A: every fcc shell stays out of
# it under |strain| <= 2% (3rd shell a*sqrt(3/2) = 4.427 A -> at most 4.52 A;
# 4th shell a*sqrt(2) = 5.112 A -> at least 5.01 A), so the force constants of
# the lattice are analytic over that range.  A 0.5 A window would start at
# 4.45 A, on top of the 3rd shell, where the switch's third derivative jumps.
TAPER_WIDTH = 0.25


def johnson_cu_functions(rc: float = 4.95, taper_width: float = TAPER_WIDTH, params: dict | None = None):
    """Analytic density f(r), pair phi(r) and embedding F(rho) of the
    Johnson-type Cu model (with the C2 taper on f and phi), plus rho_e."""
    p = dict(JOHNSON_CU)
    if params:
        p.update(params)
    r_s = rc - taper_width

    def f_of(rr):
        rr = np.asarray(rr, dtype=np.float64)
        return p["f_e"] * np.exp(-p["beta"] * (rr / p["r_e"] - 1.0)) * _taper(rr, r_s, rc)

    def phi_of(rr):
        rr = np.asarray(rr, dtype=np.float64)
        return p["phi_e"] * np.exp(-p["gamma"] * (rr / p["r_e"] - 1.0)) * _taper(rr, r_s, rc)

    # equilibrium density of the perfect fcc lattice at a = CU_A0 (shells
    # inside rc): the Banerjea-Smith reference density
    rho_e = float(sum(n * float(f_of(d)) for d, n in fcc_shells(CU_A0, rc)))
    ab = p["alpha"] / p["beta"]
    gb = p["gamma"] / p["beta"]

    def F_of(rho):
        rho = np.asarray(rho, dtype=np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            x = rho / rho_e
            pos = x > 0
            xs = np.where(pos, x, 1.0)
            val = -p["E_c"] * (1.0 - ab * np.log(xs)) * xs ** ab - 6.0 * p["phi_e"] * xs ** gb
        return np.where(pos, val, 0.0)

    return f_of, phi_of, F_of, rho_e


def johnson_cu_funcfl(nrho: int = 10000, nr: int = 10000, rc: float = 4.95,
                      taper_width: float = TAPER_WIDTH, rho_max_factor: float = 2.0,
                      params: dict | None = None) -> Funcfl:
    """Tabulate the analytic Johnson Cu EAM with a C2 taper on f and phi."""
    f_of, phi_of, F_of, rho_e = johnson_cu_functions(rc, taper_width, params)
    dr = rc / (nr - 1)
    r = np.arange(nr, dtype=np.float64) * dr
    rhor = f_of(r)
    zr = r * phi_of(r)
    drho = rho_max_factor * rho_e / (nrho - 1)
    frho = F_of(np.arange(nrho, dtype=np.float64) * drho)
    return Funcfl("Johnson-type analytic EAM for Cu (exp density, exp pair, Banerjea-Smith F)",
                  29, CU_MASS, nrho, drho, nr, dr, rc, frho, rhor, zr)


def fcc_shells(a: float, rc: float):
    """(distance, count) of fcc neighbour shells with distance < rc."""
    n = int(math.ceil(rc / (a / 2.0))) + 1
    rng = np.arange(-n, n + 1)
    i, j, k = np.meshgrid(rng, rng, rng, indexing="ij")
    keep = ((i + j + k) % 2 == 0) & ~((i == 0) & (j == 0) & (k == 0))
    d2 = (i[keep] ** 2 + j[keep] ** 2 + k[keep] ** 2).astype(np.int64)
    out = []
    for s in np.unique(d2):
        d = a / 2.0 * math.sqrt(float(s))
        if d < rc:
            out.append((d, int(np.count_nonzero(d2 == s))))
    return out


def write_funcfl(path: str, ff: Funcfl) -> None:
    with open(path, "w") as fh:
        fh.write(ff.comment.replace("\n", " ") + "\n")
        fh.write(f"{ff.z} {ff.mass!r}\n")
        fh.write(f"{ff.nrho} {ff.drho!r} {ff.nr} {ff.dr!r} {ff.cut!r}\n")
        for arr in (ff.frho, ff.rhor, ff.zr):
            vals = [repr(float(v)) for v in arr]
            for s in range(0, len(vals), 5):
                fh.write(" ".join(vals[s:s + 5]) + "\n")


def read_funcfl(path: str) -> Funcfl:
    """PairEAM::read_file + grab, pair_eam.cpp:197-268 (values parsed with
    the same decimal-to-double conversion as atof)."""
    with open(path) as fh:
        comment = fh.readline().rstrip("\n")
        w = fh.readline().split()
        z, mass = int(w[0]), float(w[1])
        w = fh.readline().split()
        nrho, drho, nr, dr, cut = int(w[0]), float(w[1]), int(w[2]), float(w[3]), float(w[4])
        vals = np.array(fh.read().split(), dtype=np.float64)
    need = nrho + 2 * nr
    if vals.size < need:
        raise ValueError(f"{path}: expected {need} table values, found {vals.size}")
    return Funcfl(comment, z, mass, nrho, drho, nr, dr, cut,
                  vals[:nrho].copy(), vals[nrho:nrho + nr].copy(), vals[nrho + nr:need].copy())


@dataclass
class Setfl:
    """Contents of one LAMMPS setfl file."""
    comments: List[str]
    elements: List[str]
    nrho: int
    drho: float
    nr: int
    dr: float
    cut: float
    number: List[int]
    mass: List[float]
    lattice_constant: List[float]
    lattice: List[str]
    frho: List[np.ndarray]                       # per element, [nrho]
    rhor: List[np.ndarray]                       # per element, [nr]
    z2r: Dict[Tuple[int, int], np.ndarray] = field(default_factory=dict)  # (i, j), i >= j: r*phi [nr]

    def element_index(self, name: str) -> int:
        if name not in self.elements:
            raise ValueError(f"No matching element {name} in EAM potential file")
        return self.elements.index(name)


def read_setfl(path: str) -> Setfl:
    """A LAMMPS setfl file (values parsed as atof parses them)."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    if len(lines) < 5:
        raise ValueError(f"{path}: not a setfl file")
    comments = lines[:3]
    w = lines[3].split()
    nel = int(w[0])
    if nel < 1 or len(w) < 1 + nel:
        raise ValueError(f"{path}: incorrect element names in EAM potential file")
    elements = w[1:1 + nel]
    w = lines[4].split()
    nrho, drho, nr, dr, cut = int(w[0]), float(w[1]), int(w[2]), float(w[3]), float(w[4])
    pos = 5

    def values(n):
        nonlocal pos
        out = []
        while len(out) < n:
            if pos >= len(lines):
                raise ValueError(f"{path}: premature end of EAM potential file")
            out.extend(lines[pos].split())
            pos += 1
        if len(out) != n:
            raise ValueError(f"{path}: table values do not end at a line end")
        return np.array(out, dtype=np.float64)

    number, mass, alat, lat, frho, rhor = [], [], [], [], [], []
    for _ in range(nel):
        while pos < len(lines) and not lines[pos].split():
            pos += 1
        if pos >= len(lines):
            raise ValueError(f"{path}: premature end of EAM potential file")
        w = lines[pos].split()
        pos += 1
        number.append(int(w[0]))
        mass.append(float(w[1]))
        alat.append(float(w[2]) if len(w) > 2 else 0.0)
        lat.append(w[3] if len(w) > 3 else "")
        frho.append(values(nrho))
        rhor.append(values(nr))
    z2r = {}
    for i in range(nel):
        for j in range(i + 1):
            z2r[(i, j)] = values(nr)
    return Setfl(comments, elements, nrho, drho, nr, dr, cut, number, mass, alat, lat, frho, rhor, z2r)


def write_setfl(path: str, sf: Setfl) -> None:
    """Write sf (17 significant digits, 5 values per line)."""
    def block(fh, arr):
        vals = [format(float(v), ".17g") for v in arr]
        for s in range(0, len(vals), 5):
            fh.write(" ".join(vals[s:s + 5]) + "\n")

    with open(path, "w") as fh:
        for c in (list(sf.comments) + ["", "", ""])[:3]:
            fh.write(c.replace("\n", " ") + "\n")
        fh.write(f"{len(sf.elements)} " + " ".join(sf.elements) + "\n")
        fh.write(f"{sf.nrho} {sf.drho!r} {sf.nr} {sf.dr!r} {sf.cut!r}\n")
        for e in range(len(sf.elements)):
            fh.write(f"{sf.number[e]} {sf.mass[e]!r} {sf.lattice_constant[e]!r} {sf.lattice[e]}\n")
            block(fh, sf.frho[e])
            block(fh, sf.rhor[e])
        for i in range(len(sf.elements)):
            for j in range(i + 1):
                block(fh, sf.z2r[(i, j)])


def setfl_from_funcfl(ff: Funcfl, element: str = "Cu", lattice_constant: float = CU_A0,
                      lattice: str = "fcc") -> Setfl:
    """A one-element setfl holding a funcfl's arrays on its own grid (the
    funcfl's third block is r*phi(r) already, pair_eam.cpp:486-487)."""
    return Setfl([ff.comment, "converted from funcfl", ""], [element], ff.nrho, ff.drho, ff.nr, ff.dr, ff.cut,
                 [ff.z], [ff.mass], [lattice_constant], [lattice], [ff.frho.copy()], [ff.rhor.copy()],
                 {(0, 0): ff.zr.copy()})


def lj_coeffs(epsilon: float, sigma: float, cut: float):
    """PairLJCut::init_one, pair_lj_cut.cpp:150-153 (offset off)."""
    lj1 = 48.0 * epsilon * math.pow(sigma, 12.0)
    lj2 = 24.0 * epsilon * math.pow(sigma, 6.0)
    lj3 = 4.0 * epsilon * math.pow(sigma, 12.0)
    lj4 = 4.0 * epsilon * math.pow(sigma, 6.0)
    return lj1, lj2, lj3, lj4, cut * cut
