"""ctypes front end of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline.  The product path
(cac_code_amd, libcac_hip.so) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)


class Box(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class EAM(C.Structure):
    _fields_ = [("nrho", C.c_int), ("nr", C.c_int), ("rdrho", C.c_double), ("rdr", C.c_double),
                ("cutforcesq", C.c_double), ("frho_spline", _d), ("rhor_spline", _d),
                ("z2r_spline", _d)]


class LJ(C.Structure):
    _fields_ = [("lj1", C.c_double), ("lj2", C.c_double), ("lj3", C.c_double),
                ("lj4", C.c_double), ("cutsq", C.c_double)]


class Elems(C.Structure):
    _fields_ = [("ne", C.c_int), ("nodex", _d), ("nae", C.c_int * 3), ("nsite", C.c_int),
                ("spa_site", _d), ("nint", C.c_int), ("spa_int", _d), ("weight", _d),
                ("node_list", C.c_int * 8)]


def build() -> str:
    """Compile liboracle.so with oracle/Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_element_max_size.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _box(lo, hi):
    b = Box()
    b.lo[:] = [float(v) for v in lo]
    b.hi[:] = [float(v) for v in hi]
    return b


# ---------------------------------------------------------------- tables
def shape(p):
    N = np.zeros(8)
    lib().orc_shape(C.c_double(p[0]), C.c_double(p[1]), C.c_double(p[2]), _p(N))
    return N


def site_shapes(nae):
    nae_c = (C.c_int * 3)(*nae)
    out = np.zeros((int(np.prod(nae)), 8))
    lib().orc_site_shapes(nae_c, _p(out))
    return out


def site_pcoords(nae):
    nae_c = (C.c_int * 3)(*nae)
    out = np.zeros((int(np.prod(nae)), 3))
    lib().orc_site_pcoords(nae_c, _p(out))
    return out


def intp_shapes(nae, ijk):
    ijk = np.ascontiguousarray(ijk, dtype=np.int32)
    out = np.zeros((ijk.shape[0], 8))
    lib().orc_intp_shapes((C.c_int * 3)(*nae), C.c_int(ijk.shape[0]), _p(ijk, _i), _p(out))
    return out


def node_list(spa_int):
    nl = (C.c_int * 8)()
    rc = lib().orc_node_list(C.c_int(spa_int.shape[0]), _p(np.ascontiguousarray(spa_int)), nl)
    if rc:
        raise ValueError("integration points do not cover every node")
    return list(nl)


def eam_tables(ff):
    """file2array + array2spline for one funcfl (pair_eam.cpp:311-580)."""
    n_out = max(ff.nrho, ff.nr) + 1
    frho = np.zeros(n_out)
    rhor = np.zeros(n_out)
    z2r = np.zeros(n_out)
    nrho, nr = C.c_int(), C.c_int()
    drho, dr = C.c_double(), C.c_double()
    rc = lib().orc_funcfl_regrid(C.c_int(ff.nrho), C.c_double(ff.drho), C.c_int(ff.nr), C.c_double(ff.dr),
                                 _p(np.ascontiguousarray(ff.frho)), _p(np.ascontiguousarray(ff.rhor)),
                                 _p(np.ascontiguousarray(ff.zr)), C.byref(nrho), C.byref(drho), C.byref(nr),
                                 C.byref(dr), _p(frho), _p(rhor), _p(z2r))
    if rc:
        raise ValueError("bad funcfl")
    out = dict(nrho=nrho.value, nr=nr.value, drho=drho.value, dr=dr.value,
               rdrho=1.0 / drho.value, rdr=1.0 / dr.value, cutforcesq=ff.cut * ff.cut,
               frho=frho[:nrho.value + 1].copy(), rhor=rhor[:nr.value + 1].copy(), z2r=z2r[:nr.value + 1].copy())
    for name, n, d in (("frho", nrho.value, drho.value), ("rhor", nr.value, dr.value), ("z2r", nr.value, dr.value)):
        sp = np.zeros((n + 1, 7))
        lib().orc_spline(C.c_int(n), C.c_double(d), _p(out[name]), _p(sp))
        out[name + "_spline"] = sp
    return out


def eam_tables_setfl(sf, element=0):
    """Tables of one element of a LAMMPS setfl file (PairEAMAlloy::file2array:
    the file's arrays are the knot values as they are, no resample), through
    array2spline (pair_eam.cpp:527-580)."""
    out = dict(nrho=sf.nrho, nr=sf.nr, drho=sf.drho, dr=sf.dr, rdrho=1.0 / sf.drho, rdr=1.0 / sf.dr,
               cutforcesq=sf.cut * sf.cut)
    for name, arr, n, d in (("frho", sf.frho[element], sf.nrho, sf.drho), ("rhor", sf.rhor[element], sf.nr, sf.dr),
                            ("z2r", sf.z2r[(element, element)], sf.nr, sf.dr)):
        f = np.zeros(n + 1)
        f[1:] = arr
        out[name] = f
        sp = np.zeros((n + 1, 7))
        lib().orc_spline(C.c_int(n), C.c_double(d), _p(f), _p(sp))
        out[name + "_spline"] = sp
    return out


def _eam_struct(t):
    e = EAM()
    e.nrho, e.nr = t["nrho"], t["nr"]
    e.rdrho, e.rdr, e.cutforcesq = t["rdrho"], t["rdr"], t["cutforcesq"]
    e.frho_spline = _p(t["frho_spline"])
    e.rhor_spline = _p(t["rhor_spline"])
    e.z2r_spline = _p(t["z2r_spline"])
    return e


class ElementSet:
    """Keeps the numpy arrays of one element population alive for ctypes."""

    def __init__(self, nodex, nae, intp_ijk, intp_w):
        self.nodex = np.ascontiguousarray(nodex, dtype=np.float64)
        self.nae = tuple(int(v) for v in nae)
        self.spa_site = site_shapes(self.nae)
        self.spa_int = intp_shapes(self.nae, intp_ijk)
        self.weight = np.ascontiguousarray(intp_w, dtype=np.float64)
        self.node_list = node_list(self.spa_int)
        self.refresh()

    def refresh(self):
        s = Elems()
        s.ne = self.nodex.shape[0]
        s.nodex = _p(self.nodex)
        s.nae[:] = list(self.nae)
        s.nsite = self.spa_site.shape[0]
        s.spa_site = _p(self.spa_site)
        s.nint = self.spa_int.shape[0]
        s.spa_int = _p(self.spa_int)
        s.weight = _p(self.weight)
        s.node_list[:] = self.node_list
        self.c = s
        return s


# ------------------------------------------------------------- neighbors
def atom_neighbors(x, lo, hi, cutneigh, maxneigh=160, brute=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    num = np.zeros(n, dtype=np.int32)
    nb = np.zeros((max(n, 1), maxneigh), dtype=np.uint32)
    fn = lib().orc_atom_neighbors_brute if brute else lib().orc_atom_neighbors
    b = _box(lo, hi)
    rc = fn(C.c_int(n), _p(x), C.byref(b), C.c_double(cutneigh), C.c_int(maxneigh), _p(num, _i), _p(nb, _u32))
    if rc:
        raise OverflowError("atom neighbor overflow")
    return num, nb[:n]


def ip_neighbors(es: ElementSet, lo, hi, cutneigh, nsplit=2, maxnip=160, brute=False):
    nip = es.nodex.shape[0] * es.spa_int.shape[0]
    num = np.zeros(nip, dtype=np.int32)
    nb = np.zeros((max(nip, 1), maxnip), dtype=np.uint64)
    b = _box(lo, hi)
    es.refresh()
    if brute:
        rc = lib().orc_ip_neighbors_brute(C.byref(es.c), C.byref(b), C.c_double(cutneigh), C.c_int(maxnip),
                                          _p(num, _i), _p(nb, _u64))
    else:
        rc = lib().orc_ip_neighbors(C.byref(es.c), C.byref(b), C.c_double(cutneigh), C.c_int(nsplit),
                                    C.c_int(maxnip), _p(num, _i), _p(nb, _u64))
    if rc:
        raise OverflowError(f"ip neighbor build failed ({rc})")
    return num, nb[:nip]


# ---------------------------------------------------------------- forces
def eam_atoms(x, lo, hi, num, nb, tables):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    rho, fp = np.zeros(n), np.zeros(n)
    f = np.zeros((n, 3))
    e = np.zeros(n)
    v = np.zeros((n, 6))
    t = _eam_struct(tables)
    b = _box(lo, hi)
    nb = np.ascontiguousarray(nb)
    lib().orc_eam_atoms(C.c_int(n), _p(x), C.byref(b), _p(num, _i), _p(nb, _u32), C.c_int(nb.shape[1]),
                        C.byref(t), _p(rho), _p(fp), _p(f), _p(e), _p(v))
    return dict(rho=rho, fp=fp, f=f, eatom=e, vatom=v)


def eam_elements(es: ElementSet, lo, hi, num, nb, tables):
    ne, nint = es.nodex.shape[0], es.spa_int.shape[0]
    erho = np.zeros((ne, 8))
    nodef = np.zeros((ne, 8, 3))
    ipf = np.zeros((ne * nint, 3))
    eip = np.zeros(ne * nint)
    vip = np.zeros((ne * nint, 6))
    t = _eam_struct(tables)
    b = _box(lo, hi)
    es.refresh()
    nb = np.ascontiguousarray(nb)
    lib().orc_eam_elements(C.byref(es.c), C.byref(b), _p(num, _i), _p(nb, _u64), C.c_int(nb.shape[1]),
                           C.byref(t), _p(erho), _p(nodef), _p(ipf), _p(eip), _p(vip))
    return dict(erho=erho, nodef=nodef, ipf=ipf, eip=eip, vip=vip)


def lj_atoms(x, lo, hi, num, nb, lj):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    f = np.zeros((n, 3))
    e = np.zeros(n)
    v = np.zeros((n, 6))
    p = LJ(*lj)
    b = _box(lo, hi)
    nb = np.ascontiguousarray(nb)
    lib().orc_lj_atoms(C.c_int(n), _p(x), C.byref(b), _p(num, _i), _p(nb, _u32), C.c_int(nb.shape[1]),
                       C.byref(p), _p(f), _p(e), _p(v))
    return dict(f=f, eatom=e, vatom=v)


def lj_elements(es: ElementSet, lo, hi, num, nb, lj):
    ne, nint = es.nodex.shape[0], es.spa_int.shape[0]
    nodef = np.zeros((ne, 8, 3))
    ipf = np.zeros((ne * nint, 3))
    eip = np.zeros(ne * nint)
    vip = np.zeros((ne * nint, 6))
    p = LJ(*lj)
    b = _box(lo, hi)
    es.refresh()
    nb = np.ascontiguousarray(nb)
    lib().orc_lj_elements(C.byref(es.c), C.byref(b), _p(num, _i), _p(nb, _u64), C.c_int(nb.shape[1]),
                          C.byref(p), _p(nodef), _p(ipf), _p(eip), _p(vip))
    return dict(nodef=nodef, ipf=ipf, eip=eip, vip=vip)


# ------------------------------------------------------------ integration
def nve_atoms_initial(x, v, f, dtv, dtf, mass):
    lib().orc_nve_atoms_initial(C.c_int(x.shape[0]), _p(x), _p(v), _p(np.ascontiguousarray(f)),
                                C.c_double(dtv), C.c_double(dtf), C.c_double(mass))


def nve_atoms_final(v, f, dtf, mass):
    lib().orc_nve_atoms_final(C.c_int(v.shape[0]), _p(v), _p(np.ascontiguousarray(f)),
                              C.c_double(dtf), C.c_double(mass))


def nve_nodes_initial(nodex, nodev, nodef, xc, dtv, dtf, mass_e):
    lib().orc_nve_nodes_initial(C.c_int(nodex.shape[0]), _p(nodex), _p(nodev), _p(np.ascontiguousarray(nodef)),
                                _p(xc) if xc is not None else None, C.c_double(dtv), C.c_double(dtf),
                                C.c_double(mass_e))


def nve_nodes_final(nodev, nodef, dtf, mass_e):
    lib().orc_nve_nodes_final(C.c_int(nodev.shape[0]), _p(nodev), _p(np.ascontiguousarray(nodef)),
                              C.c_double(dtf), C.c_double(mass_e))


def pbc_atoms(x, lo, hi):
    b = _box(lo, hi)
    lib().orc_pbc_atoms(C.c_int(x.shape[0]), _p(x), None, C.byref(b))


def pbc_elements(xc, nodex, lo, hi):
    b = _box(lo, hi)
    lib().orc_pbc_elements(C.c_int(nodex.shape[0]), _p(xc), _p(nodex), None, C.byref(b))


def element_centers(nodex):
    xc = np.zeros((nodex.shape[0], 3))
    lib().orc_element_centers(C.c_int(nodex.shape[0]), _p(np.ascontiguousarray(nodex)), _p(xc))
    return xc


def element_max_size(nodex):
    return float(lib().orc_element_max_size(C.c_int(nodex.shape[0]), _p(np.ascontiguousarray(nodex))))


def displacement_exceeds(x, xhold, triggersq):
    return bool(lib().orc_max_displacement_exceeds(C.c_int(x.shape[0]), _p(np.ascontiguousarray(x)),
                                                  _p(np.ascontiguousarray(xhold)), C.c_double(triggersq)))


# ------------------------------------------------------ atom-element coupling
ATOM_SITE = 0xFFFFFFFF


def coupled_neighbors(x, es: ElementSet, lo, hi, cutneigh, nsplit=2, maxneigh=160, maxnip=160, brute=False):
    """Coupled lists (uint64 entries, site field ATOM_SITE for atoms): atom
    rows hold atoms and element sites, integration-point rows sites and
    atoms, each row sorted."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    nip = es.nodex.shape[0] * es.spa_int.shape[0]
    anum = np.zeros(max(n, 1), dtype=np.int32)
    anb = np.zeros((max(n, 1), maxneigh), dtype=np.uint64)
    inum = np.zeros(max(nip, 1), dtype=np.int32)
    inb = np.zeros((max(nip, 1), maxnip), dtype=np.uint64)
    b = _box(lo, hi)
    es.refresh()
    rc = lib().orc_coupled_neighbors(C.c_int(n), _p(x), C.byref(es.c), C.byref(b), C.c_double(cutneigh),
                                     C.c_int(nsplit), C.c_int(1 if brute else 0), C.c_int(maxneigh), _p(anum, _i),
                                     _p(anb, _u64), C.c_int(maxnip), _p(inum, _i), _p(inb, _u64))
    if rc:
        raise OverflowError(f"coupled neighbor build failed ({rc})")
    return anum[:n], anb[:n], inum[:nip], inb[:nip]


def eam_coupled(x, es: ElementSet, lo, hi, anum, anb, inum, inb, tables):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    ne, nint = es.nodex.shape[0], es.spa_int.shape[0]
    out = dict(rho=np.zeros(n), fp=np.zeros(n), f=np.zeros((n, 3)), eatom=np.zeros(n), vatom=np.zeros((n, 6)),
               erho=np.zeros((ne, 8)), nodef=np.zeros((ne, 8, 3)), ipf=np.zeros((ne * nint, 3)),
               eip=np.zeros(ne * nint), vip=np.zeros((ne * nint, 6)))
    t = _eam_struct(tables)
    b = _box(lo, hi)
    es.refresh()
    anb = np.ascontiguousarray(anb)
    inb = np.ascontiguousarray(inb)
    lib().orc_eam_coupled(C.c_int(n), _p(x), C.byref(es.c), C.byref(b), _p(anum, _i), _p(anb, _u64),
                          C.c_int(anb.shape[1] if anb.ndim == 2 else 0), _p(inum, _i), _p(inb, _u64),
                          C.c_int(inb.shape[1] if inb.ndim == 2 else 0), C.byref(t),
                          *[_p(out[k]) for k in ("rho", "fp", "f", "eatom", "vatom", "erho", "nodef", "ipf",
                                                  "eip", "vip")])
    return out


def centro_atoms(x, lo, hi, num, nb, cutneigh, nnn=12, es: ElementSet | None = None):
    """Centro-symmetry parameter of every atom from its list (uint32 atom
    entries, or the uint64 coupled entries with es for the sites)."""
    if not (nnn % 2 == 0 and 2 <= nnn <= 16):
        raise ValueError("nnn must be even, 2..16")
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    nb = np.asarray(nb)
    if nb.dtype == np.uint32:
        nb = nb.astype(np.uint64) | (np.uint64(ATOM_SITE) << np.uint64(32))
    out = np.zeros(n)
    if n == 0:
        return out
    nb = np.ascontiguousarray(nb, dtype=np.uint64).reshape(n, -1)
    b = _box(lo, hi)
    ec = None
    if es is not None:
        es.refresh()
        ec = C.byref(es.c)
    num = np.ascontiguousarray(num, dtype=np.int32)
    lib().orc_centro_atoms(C.c_int(n), _p(x), ec, C.byref(b), _p(num, _i), _p(nb, _u64), C.c_int(nb.shape[1]),
                           C.c_double(cutneigh), C.c_int(nnn), _p(out))
    return out


def lj_coupled(x, es: ElementSet, lo, hi, anum, anb, inum, inb, lj):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    ne, nint = es.nodex.shape[0], es.spa_int.shape[0]
    out = dict(f=np.zeros((n, 3)), eatom=np.zeros(n), vatom=np.zeros((n, 6)), nodef=np.zeros((ne, 8, 3)),
               ipf=np.zeros((ne * nint, 3)), eip=np.zeros(ne * nint), vip=np.zeros((ne * nint, 6)))
    p = LJ(*lj)
    b = _box(lo, hi)
    es.refresh()
    anb = np.ascontiguousarray(anb)
    inb = np.ascontiguousarray(inb)
    lib().orc_lj_coupled(C.c_int(n), _p(x), C.byref(es.c), C.byref(b), _p(anum, _i), _p(anb, _u64),
                         C.c_int(anb.shape[1] if anb.ndim == 2 else 0), _p(inum, _i), _p(inb, _u64),
                         C.c_int(inb.shape[1] if inb.ndim == 2 else 0), C.byref(p),
                         *[_p(out[k]) for k in ("f", "eatom", "vatom", "nodef", "ipf", "eip", "vip")])
    return out


def num_threads():
    return int(lib().orc_num_threads())


def set_num_threads(n: int):
    lib().orc_set_num_threads(C.c_int(n))
