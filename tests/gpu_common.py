"""Shared helpers for the GPU parity tests: small systems, the oracle's view of
them, and sorted-set comparisons of neighbor lists."""
from __future__ import annotations

import numpy as np

from cac_code_amd.lattice import build_cac_system, corner_quadrature, CACSystem

CUT_EAM = 4.95
CUTNEIGH_EAM = 5.95


def jitter(s: CACSystem, sigma: float, seed: int) -> CACSystem:
    """Displace atoms and nodes by Gaussian noise (nonaffine elements)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s.x = np.ascontiguousarray(s.x + sigma * rng.standard_normal(s.x.shape))
    s.x = np.mod(s.x - s.box_lo, s.box_hi - s.box_lo) + s.box_lo
    s.nodex = np.ascontiguousarray(s.nodex + sigma * rng.standard_normal(s.nodex.shape))
    return s


def small_system(seed=7, L=10, nae=5, core=4.0, sigma=0.05, temperature=300.0):
    return jitter(build_cac_system(L, nae, core, temperature=temperature, seed=seed), sigma, seed + 1)


def sorted_rows(num, nb):
    return [np.sort(nb[i, :num[i]]) for i in range(len(num))]


def assert_same_sets(num_a, nb_a, num_b, nb_b, what):
    assert np.array_equal(num_a, num_b), f"{what}: neighbor counts differ at {np.nonzero(num_a != num_b)[0][:10]}"
    ra, rb = sorted_rows(num_a, nb_a), sorted_rows(num_b, nb_b)
    bad = [i for i in range(len(ra)) if not np.array_equal(ra[i], rb[i])]
    assert not bad, f"{what}: neighbor sets differ for rows {bad[:10]}"


def rel_err(a, b):
    """max |a-b| / max |b| (norm-relative; per-component ratios are
    meaningless where the reference value crosses zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b))) / scale


def rms_rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = float(np.sqrt(np.mean(b * b))) if b.size else 0.0
    return float(np.sqrt(np.mean((a - b) ** 2))) / nb if nb > 0 else 0.0


def atoms_u64(nb):
    """Oracle atom entries (uint32 j | img << 27) in the library's uint64
    form (site field CAC_ATOM_SITE)."""
    return nb.astype(np.uint64) | (np.uint64(0xFFFFFFFF) << np.uint64(32))


def oracle_state(O, s: CACSystem, cutneigh, tables=None, lj=None, nsplit=2, maxneigh=200, maxnip=200,
                 coupled=False):
    """Oracle neighbor lists and forces for system s (positions as given).
    Atom entries are returned in the library's uint64 form."""
    es = O.ElementSet(s.nodex, s.nae, s.intp_ijk, s.intp_w)
    out = {"es": es}
    if coupled:
        out["anum"], out["anb"], out["inum"], out["inb"] = O.coupled_neighbors(
            s.x, es, s.box_lo, s.box_hi, cutneigh, nsplit, maxneigh, maxnip)
        args = (s.x, es, s.box_lo, s.box_hi, out["anum"], out["anb"], out["inum"], out["inb"])
        r = None
        if tables is not None:
            r = O.eam_coupled(*args, tables)
        if lj is not None:
            r = O.lj_coupled(*args, lj)
        if r is not None:
            out["atoms"] = {k: r[k] for k in ("rho", "fp", "f", "eatom", "vatom") if k in r}
            out["elems"] = {k: r[k] for k in ("erho", "nodef", "ipf", "eip", "vip") if k in r}
        return out
    if s.n_atoms:
        out["anum"], anb = O.atom_neighbors(s.x, s.box_lo, s.box_hi, cutneigh, maxneigh)
        out["anb"] = atoms_u64(anb)
        out["anb32"] = anb
    if s.n_elements:
        out["inum"], out["inb"] = O.ip_neighbors(es, s.box_lo, s.box_hi, cutneigh, nsplit, maxnip)
    if tables is not None:
        if s.n_atoms:
            out["atoms"] = O.eam_atoms(s.x, s.box_lo, s.box_hi, out["anum"], out["anb32"], tables)
        if s.n_elements:
            out["elems"] = O.eam_elements(es, s.box_lo, s.box_hi, out["inum"], out["inb"], tables)
    if lj is not None:
        if s.n_atoms:
            out["atoms"] = O.lj_atoms(s.x, s.box_lo, s.box_hi, out["anum"], out["anb32"], lj)
        if s.n_elements:
            out["elems"] = O.lj_elements(es, s.box_lo, s.box_hi, out["inum"], out["inb"], lj)
    return out
