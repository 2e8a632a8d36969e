"""Host-side mirror of the reference's pair-style and run interfaces on top of
the C-ABI (include/cac_hip.h).

* ``PairEAM`` / ``PairLJCut`` follow ``pair_style eam`` and ``pair_style
  lj/cut`` of the reference: ``settings(args)`` (pair_eam.cpp:109-114,
  pair_lj_cut.cpp:74-88), ``coeff(args)`` (pair_eam.cpp:120-160,
  pair_lj_cut.cpp:94-121), ``init_one(i, j)`` (pair_eam.cpp:288-304,
  pair_lj_cut.cpp:140-198) and ``compute``/``ecompute`` (pair.h:96-97),
  with the same argument meaning and the same error conditions.
* ``Engine`` owns one device context: box, atoms, elements, integration
  points, neighbor settings (``neighbor`` / ``neigh_modify``), force calls
  and the Verlet/FixNVE loop (verlet.cpp:149-219, fix_nve.cpp:40-242).

Everything numeric runs in libcac_hip.so; this module only moves arguments.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .lattice import BOLTZ_METAL, CACSystem, FTM2V_METAL, MVV2E_METAL, NKTV2P_METAL, vertex_clusters
from .potential import Funcfl, Setfl, lj_coeffs, read_funcfl, read_setfl

SURR_NFEAT, SURR_NOUT, SURR_NQUAD = 279, 24, 561   # CAC_SURR_* in cac_hip.h
SURR_NFEAT_CONFIG = 102                            # CAC_SURR_NFEAT_CONFIG


def _p(a, t=_lib._d):
    return a.ctypes.data_as(t)


def _check(ctx, rc):
    if rc != _lib.CAC_OK:
        msg = _lib.lib().cac_last_error(ctx).decode() if ctx else "no context"
        raise _lib.CacError(rc, msg)


def _bounds(s: str, ntypes: int):
    """Force::bounds, force.cpp:158-180: 'N', '*', 'N*', '*N', 'M*N'."""
    if s == "*":
        return 1, ntypes
    if "*" not in s:
        v = int(s)
        return v, v
    a, b = s.split("*", 1)
    lo = int(a) if a else 1
    hi = int(b) if b else ntypes
    return lo, hi


class PairError(ValueError):
    pass


class PairEAM:
    """pair_style eam: one funcfl file per atom type (pair_eam.cpp:120-160)."""
    style = "eam"

    def __init__(self, ntypes: int = 1):
        self.ntypes = ntypes
        self.funcfl: Optional[Funcfl] = None
        self.mass: Optional[float] = None
        self.tables = None

    def settings(self, args):
        if len(args) > 0:
            raise PairError("Illegal pair_style command")          # pair_eam.cpp:111

    def coeff(self, args, funcfl: Optional[Funcfl] = None):
        if len(args) != 3:
            raise PairError("Incorrect args for pair coefficients")  # pair_eam.cpp:124
        ilo, ihi = _bounds(args[0], self.ntypes)
        jlo, jhi = _bounds(args[1], self.ntypes)
        if not (ilo <= 1 <= ihi and jlo <= 1 <= jhi) or ihi > self.ntypes:
            raise PairError("Incorrect args for pair coefficients")
        ff = funcfl if funcfl is not None else read_funcfl(args[2])
        self.funcfl = ff
        self.mass = ff.mass                                          # pair_eam.cpp:159-160
        self.tables = eam_tables(ff)

    def init_one(self, i: int, j: int) -> float:
        if self.funcfl is None:
            raise PairError("All pair coeffs are not set")
        return self.funcfl.cut                                       # pair_eam.cpp:297-301

    def install(self, eng: "Engine"):
        t = self.tables
        cut = self.init_one(1, 1)
        _check(eng.ctx, _lib.lib().cac_set_eam(eng.ctx, t["nrho"], t["drho"], _p(t["frho_spline"]), t["nr"],
                                               t["dr"], _p(t["rhor_spline"]), _p(t["z2r_spline"]), cut * cut))


class PairEAMAlloy(PairEAM):
    """pair_style eam/alloy: one LAMMPS setfl file, `pair_coeff * * file el`
    mapping atom type 1 to element `el` of the file (LAMMPS' PairEAMAlloy).
    No counterpart in the reference, which reads funcfl only
    (pair_eam.cpp:197-268); the tables come from cac_eam_tables_from_setfl
    and run through the same device passes as eam."""
    style = "eam/alloy"

    def __init__(self, ntypes: int = 1):
        super().__init__(ntypes)
        self.setfl: Optional[Setfl] = None
        self.element = None

    def coeff(self, args, setfl: Optional[Setfl] = None):
        if len(args) != 3 + self.ntypes or args[0] != "*" or args[1] != "*":
            raise PairError("Incorrect args for pair coefficients")
        sf = setfl if setfl is not None else read_setfl(args[2])
        e = sf.element_index(args[3])
        self.setfl, self.element = sf, e
        self.mass = sf.mass[e]
        self.tables = eam_tables_setfl(sf, e)

    def init_one(self, i: int, j: int) -> float:
        if self.setfl is None:
            raise PairError("All pair coeffs are not set")
        return self.setfl.cut


class PairLJCut:
    """pair_style lj/cut cut (pair_lj_cut.cpp:74-198)."""
    style = "lj/cut"

    def __init__(self, ntypes: int = 1):
        self.ntypes = ntypes
        self.cut_global = None
        self.eps = self.sigma = self.cut = None

    def settings(self, args):
        if len(args) != 1:
            raise PairError("Illegal pair_style command")
        self.cut_global = float(args[0])

    def coeff(self, args):
        if len(args) < 4 or len(args) > 5:
            raise PairError("Incorrect args for pair coefficients")  # pair_lj_cut.cpp:96-97
        _bounds(args[0], self.ntypes)
        _bounds(args[1], self.ntypes)
        self.eps = float(args[2])
        self.sigma = float(args[3])
        self.cut = float(args[4]) if len(args) == 5 else self.cut_global

    def init_one(self, i: int, j: int):
        if self.eps is None:
            raise PairError("All pair coeffs are not set")
        return lj_coeffs(self.eps, self.sigma, self.cut)

    def install(self, eng: "Engine"):
        lj1, lj2, lj3, lj4, cutsq = self.init_one(1, 1)
        _check(eng.ctx, _lib.lib().cac_set_lj(eng.ctx, lj1, lj2, lj3, lj4, cutsq))


def eam_tables(ff: Funcfl):
    """file2array + array2spline in the native library (pair_eam.cpp:311-580)."""
    L = _lib.lib()
    nrho, nr = C.c_int(), C.c_int()
    drho, dr = C.c_double(), C.c_double()
    fr, rr, zr = (np.ascontiguousarray(a, dtype=np.float64) for a in (ff.frho, ff.rhor, ff.zr))
    rc = L.cac_eam_tables_from_funcfl(ff.nrho, ff.drho, ff.nr, ff.dr, _p(fr), _p(rr), _p(zr), C.byref(nrho),
                                      C.byref(drho), C.byref(nr), C.byref(dr), None, None, None)
    if rc:
        raise PairError("bad EAM file")
    out = dict(nrho=nrho.value, nr=nr.value, drho=drho.value, dr=dr.value, cutforcesq=ff.cut * ff.cut)
    out["frho_spline"] = np.zeros((nrho.value + 1, 7))
    out["rhor_spline"] = np.zeros((nr.value + 1, 7))
    out["z2r_spline"] = np.zeros((nr.value + 1, 7))
    rc = L.cac_eam_tables_from_funcfl(ff.nrho, ff.drho, ff.nr, ff.dr, _p(fr), _p(rr), _p(zr), C.byref(nrho),
                                      C.byref(drho), C.byref(nr), C.byref(dr), _p(out["frho_spline"]),
                                      _p(out["rhor_spline"]), _p(out["z2r_spline"]))
    if rc:
        raise PairError("bad EAM file")
    return out


def eam_tables_setfl(sf: Setfl, element: int = 0):
    """Spline tables of one element of a setfl file: its arrays are the knot
    values as they are (cac_eam_tables_from_setfl)."""
    L = _lib.lib()
    fr = np.ascontiguousarray(sf.frho[element], dtype=np.float64)
    rr = np.ascontiguousarray(sf.rhor[element], dtype=np.float64)
    zr = np.ascontiguousarray(sf.z2r[(element, element)], dtype=np.float64)
    out = dict(nrho=sf.nrho, nr=sf.nr, drho=sf.drho, dr=sf.dr, cutforcesq=sf.cut * sf.cut)
    out["frho_spline"] = np.zeros((sf.nrho + 1, 7))
    out["rhor_spline"] = np.zeros((sf.nr + 1, 7))
    out["z2r_spline"] = np.zeros((sf.nr + 1, 7))
    rc = L.cac_eam_tables_from_setfl(sf.nrho, sf.drho, sf.nr, sf.dr, _p(fr), _p(rr), _p(zr),
                                     _p(out["frho_spline"]), _p(out["rhor_spline"]), _p(out["z2r_spline"]))
    if rc:
        raise PairError("bad EAM file")
    return out


class _MlpMixin:
    """Device MLP and surrogate entry points of Engine (cac_mlp_*, cac_surrogate_*)."""

    def mlp_create(self, nin=None, nhidden=256, nlayers=3, nout=24, seed=0):
        nin = SURR_NFEAT if nin is None else nin
        _check(self.ctx, self.L.cac_mlp_create(self.ctx, nin, nhidden, nlayers, nout, seed))
        self.mlp_shape = (nin, nhidden, nlayers, nout)

    def _layer_dims(self):
        nin, nh, nl, nout = self.mlp_shape
        dims = [nin] + [nh] * nl + [nout]
        return [(dims[l + 1], dims[l]) for l in range(nl + 1)]

    def mlp_get_weights(self):
        out = []
        for l, (o, i) in enumerate(self._layer_dims()):
            w = np.zeros((o, i), dtype=np.float32); b = np.zeros(o, dtype=np.float32)
            _check(self.ctx, self.L.cac_mlp_get_weights(self.ctx, l, _p(w, _lib._f), _p(b, _lib._f)))
            out.append((w, b))
        return out

    def mlp_set_weights(self, weights):
        for l, (w, b) in enumerate(weights):
            w = np.ascontiguousarray(w, dtype=np.float32); b = np.ascontiguousarray(b, dtype=np.float32)
            _check(self.ctx, self.L.cac_mlp_set_weights(self.ctx, l, _p(w, _lib._f), _p(b, _lib._f)))

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.mlp_shape[3]), dtype=np.float32)
        _check(self.ctx, self.L.cac_mlp_forward(self.ctx, x.shape[0], _p(x, _lib._f), _p(y, _lib._f)))
        return y

    def mlp_grads(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32); t = np.ascontiguousarray(t, dtype=np.float32)
        dims = self._layer_dims()
        bufs = [np.zeros(o * i + o, dtype=np.float32) for o, i in dims]
        arr = (_lib._f * len(bufs))(*[_p(b, _lib._f) for b in bufs])
        loss = C.c_float()
        _check(self.ctx, self.L.cac_mlp_grads(self.ctx, x.shape[0], _p(x, _lib._f), _p(t, _lib._f), C.byref(loss),
                                              arr))
        grads = [(b[:o * i].reshape(o, i), b[o * i:]) for b, (o, i) in zip(bufs, dims)]
        return loss.value, grads

    def mlp_train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32); t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        _check(self.ctx, self.L.cac_mlp_train_step(self.ctx, x.shape[0], _p(x, _lib._f), _p(t, _lib._f), lr,
                                                   C.byref(loss)))
        return loss.value

    def mlp_set_dataset(self, x, t):
        self._ds = (np.ascontiguousarray(x, dtype=np.float32), np.ascontiguousarray(t, dtype=np.float32))
        _check(self.ctx, self.L.cac_mlp_set_dataset(self.ctx, self._ds[0].shape[0], _p(self._ds[0], _lib._f),
                                                    _p(self._ds[1], _lib._f)))

    def mlp_train(self, epochs, batch, lr, seed=0):
        loss = C.c_float()
        _check(self.ctx, self.L.cac_mlp_train(self.ctx, epochs, batch, lr, seed, C.byref(loss)))
        return loss.value

    def surrogate_setup(self, X0, nbr, in_scale, out_scale, strain_limit=0.02):
        X0 = np.ascontiguousarray(X0, dtype=np.float64); nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        a = np.ascontiguousarray(in_scale, dtype=np.float32); b = np.ascontiguousarray(out_scale, dtype=np.float32)
        _check(self.ctx, self.L.cac_surrogate_setup(self.ctx, _p(X0), _p(nbr, _lib._i), _p(a, _lib._f),
                                                    _p(b, _lib._f), strain_limit))

    def surrogate_set_fit(self, A, Q, b):
        A = np.ascontiguousarray(A, dtype=np.float32); Q = np.ascontiguousarray(Q, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        _check(self.ctx, self.L.cac_surrogate_set_fit(self.ctx, _p(A, _lib._f), _p(Q, _lib._f), _p(b, _lib._f)))

    def surrogate_enable(self, on=True):
        _check(self.ctx, self.L.cac_surrogate_enable(self.ctx, int(on)))

    def surrogate_features(self):
        f = np.zeros((self.ne, SURR_NFEAT), dtype=np.float32); c = np.zeros(self.ne, dtype=np.int32)
        _check(self.ctx, self.L.cac_surrogate_features(self.ctx, _p(f, _lib._f), _p(c, _lib._i)))
        return f, c

    def surrogate_info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(self.ctx, self.L.cac_surrogate_info(self.ctx, C.byref(a), C.byref(b), C.byref(c)))
        return dict(interface=a.value, strained=b.value, surrogate=c.value)



class Engine(_MlpMixin):
    """One device context (cac_ctx) and the run state around it."""

    def __init__(self, device: int = 0):
        self.L = _lib.lib()
        h = C.c_void_p()
        rc = self.L.cac_ctx_create(device, C.byref(h))
        if rc != _lib.CAC_OK:
            raise _lib.CacError(rc, "cac_ctx_create failed (no HIP device?)")
        self.ctx = h
        self.n = 0
        self.ne = 0
        self.nint = 0
        self.pair = None
        self.sys: Optional[CACSystem] = None
        self.box = None
        self.mass = 0.0
        self.nsite = 0
        self.neighbor_args = None

    def close(self):
        if self.ctx:
            self.L.cac_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- set-up
    def set_box(self, lo, hi):
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        per = np.ones(3, dtype=np.int32)
        _check(self.ctx, self.L.cac_set_box(self.ctx, _p(lo), _p(hi), _p(per, _lib._i)))
        self.box = (lo.copy(), hi.copy())

    def upload_atoms(self, x, v, mass):
        x = np.ascontiguousarray(x, dtype=np.float64)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        self.n = x.shape[0]
        _check(self.ctx, self.L.cac_upload_atoms(self.ctx, self.n, _p(x), _p(vv) if vv is not None else None, mass))
        self.mass = mass

    def upload_elements(self, nodex, nodev, nae, mass_e, intp_ijk, intp_w):
        nodex = np.ascontiguousarray(nodex, dtype=np.float64)
        nv = None if nodev is None else np.ascontiguousarray(nodev, dtype=np.float64)
        nae_c = np.ascontiguousarray(nae, dtype=np.int32)
        self.ne = nodex.shape[0]
        self.nsite = int(np.prod(nae_c))
        _check(self.ctx, self.L.cac_upload_elements(self.ctx, self.ne, _p(nae_c, _lib._i), _p(nodex),
                                                    _p(nv) if nv is not None else None, mass_e))
        ijk = np.ascontiguousarray(intp_ijk, dtype=np.int32)
        w = np.ascontiguousarray(intp_w, dtype=np.float64)
        self.nint = ijk.shape[0]
        _check(self.ctx, self.L.cac_set_quadrature(self.ctx, self.nint, _p(ijk, _lib._i), _p(w)))

    def load_system(self, s: CACSystem, mass: Optional[float] = None):
        self.sys = s
        self.set_box(s.box_lo, s.box_hi)
        m = s.mass if mass is None else mass
        self.mass = m
        self.upload_atoms(s.x, s.v, m)
        mass_e = float(s.nae[0] * s.nae[1] * s.nae[2]) * m
        self.mass_e = mass_e
        self.upload_elements(s.nodex, s.nodev, s.nae, mass_e, s.intp_ijk, s.intp_w)
        if getattr(s, "a0", None):
            self.set_list_order(0.5 * s.a0)                 # fcc neighbor vectors: multiples of a/2
        cl = vertex_clusters(s)
        if cl is None:                                       # no mesh vertices known: no clusters
            cl = (np.zeros(1, np.int32), np.zeros(0, np.int32))
        self.set_ip_clusters(*cl)

    def set_ip_clusters(self, start, slots):
        """Group the integration points by mesh vertex (a work grouping for
        the element passes; the forces do not change).  start[ncl+1], slots:
        node slots 8*e + k per cluster."""
        start = np.ascontiguousarray(start, dtype=np.int32)
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        _check(self.ctx, self.L.cac_set_ip_clusters(self.ctx, int(start.size - 1), _p(start, _lib._i),
                                                    _p(slots, _lib._i)))

    def deform(self, scale):
        """Affine remap of the box and its contents by scale[3] about box lo."""
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        _check(self.ctx, self.L.cac_deform(self.ctx, _p(sc)))

    def set_deform_rate(self, rate):
        """Engineering strain rate per time unit (ps for metal units) applied
        by run() before every step; zeros stop it."""
        r = np.ascontiguousarray(rate, dtype=np.float64)
        _check(self.ctx, self.L.cac_set_deform_rate(self.ctx, _p(r)))

    def cluster_info(self):
        """(clusters, cluster passes active after the last build, largest
        union of a cluster's lists)."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(self.ctx, self.L.cac_cluster_info(self.ctx, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, bool(b.value), c.value

    def set_pair(self, pair):
        self.pair = pair
        pair.install(self)

    def set_neighbor(self, cutneigh: float, skin: Optional[float] = None, every: int = 1, delay: int = 0,
                     check: bool = True, maxneigh: int = 0, maxnip: int = 0):
        """neighbor <cutneigh> bin + neigh_modify.  The rebuild trigger is half
        of the true skin (cutneigh - force cutoff); the reference takes half of
        the whole cutoff (neighbor.cpp:334, 357-358), which would let pairs
        enter the force cutoff unlisted."""
        if skin is None:
            cut = 0.0
            if isinstance(self.pair, PairEAM) and self.pair.funcfl is not None:
                cut = self.pair.init_one(1, 1)
            elif isinstance(self.pair, PairLJCut) and self.pair.cut is not None:
                cut = self.pair.cut
            skin = cutneigh - cut
            if skin <= 0.0:
                raise PairError("neighbor cutoff must exceed the force cutoff")
        triggersq = 0.25 * skin * skin
        self.neighbor_args = (cutneigh, skin, every, delay, check)
        _check(self.ctx, self.L.cac_set_neighbor(self.ctx, cutneigh, triggersq, every, delay, 1 if check else 0,
                                                 maxneigh, maxnip))

    def set_list_order(self, spacing: float):
        """Sort each atom list by displacement rounded to `spacing` (a/2 for
        fcc): a gather-locality hint, the neighbor sets do not change."""
        _check(self.ctx, self.L.cac_set_list_order(self.ctx, float(spacing)))

    def set_coupling(self, on: bool = True):
        """Atom-element coupling (build-defined, off by default = the
        reference): atom lists also hold element sites, integration-point
        lists also atoms.  Takes effect at the next build."""
        _check(self.ctx, self.L.cac_set_coupling(self.ctx, 1 if on else 0))

    # ---- actions
    def build(self):
        _check(self.ctx, self.L.cac_build_neighbors(self.ctx))

    def decide(self) -> bool:
        flag = C.c_int()
        _check(self.ctx, self.L.cac_decide(self.ctx, C.byref(flag)))
        return bool(flag.value)

    def force(self, eflag=False, vflag=False, precision=_lib.PREC_FP64):
        _check(self.ctx, self.L.cac_force(self.ctx, int(eflag), int(vflag), precision))

    def nve_initial(self, dtv, dtf):
        _check(self.ctx, self.L.cac_nve_initial(self.ctx, dtv, dtf))

    def nve_final(self, dtf):
        _check(self.ctx, self.L.cac_nve_final(self.ctx, dtf))

    def run(self, nsteps: int, dt: float = 0.001, precision=_lib.PREC_FP64, ftm2v: float = FTM2V_METAL) -> int:
        nb = C.c_int()
        dtf = 0.5 * dt * ftm2v                                  # fix_nve.cpp:43
        _check(self.ctx, self.L.cac_run(self.ctx, nsteps, dt, dtf, precision, C.byref(nb)))
        return nb.value

    def synchronize(self):
        _check(self.ctx, self.L.cac_synchronize(self.ctx))

    # ---- results
    def atoms(self):
        x = np.zeros((self.n, 3)); v = np.zeros((self.n, 3)); f = np.zeros((self.n, 3))
        _check(self.ctx, self.L.cac_download_atoms(self.ctx, _p(x), _p(v), _p(f)))
        return x, v, f

    def elements(self):
        a = np.zeros((self.ne, 8, 3)); b = np.zeros((self.ne, 8, 3)); c = np.zeros((self.ne, 8, 3))
        _check(self.ctx, self.L.cac_download_elements(self.ctx, _p(a), _p(b), _p(c)))
        return a, b, c

    def energy_virial(self):
        e = np.zeros(2); w = np.zeros(6)
        _check(self.ctx, self.L.cac_get_energy_virial(self.ctx, _p(e), _p(w)))
        return e, w

    # ---- thermodynamic output (thermo_style pe ke temp press etotal; the
    # reference prints only the step, thermo.cpp:195-209)
    def thermo(self, mvv2e: float = MVV2E_METAL, boltz: float = BOLTZ_METAL, nktv2p: float = NKTV2P_METAL):
        """pe, ke, temp, press, etotal of the current state (metal units by
        default): one FP64 force call with the tallies on the device."""
        out = np.zeros(5)
        _check(self.ctx, self.L.cac_thermo(self.ctx, mvv2e, boltz, nktv2p, _p(out)))
        return dict(zip(("pe", "ke", "temp", "press", "etotal"), (float(v) for v in out)))

    @staticmethod
    def thermo_header() -> str:
        return f"{'Step':>10} {'PotEng':>18} {'KinEng':>16} {'Temp':>12} {'Press':>14} {'TotEng':>18}"

    @staticmethod
    def thermo_line(step: int, t: dict) -> str:
        """One thermo line, the columns of thermo_header (LAMMPS' one-line
        format: step as an integer, the rest %g-like)."""
        return (f"{step:>10d} {t['pe']:>18.8f} {t['ke']:>16.8f} {t['temp']:>12.6f} {t['press']:>14.6f} "
                f"{t['etotal']:>18.8f}")

    # ---- centro-symmetry and the xyz dump (Output::centro_atom,
    # output.cpp:352-441; DumpXyz, dump_xyz.cpp:36-76, 127-141)
    def centro(self, nnn: int = 12) -> np.ndarray:
        """Centro-symmetry parameter of every atom over its current list (the
        reference computes it right after a build, verlet.cpp:93-94)."""
        out = np.zeros(self.n)
        _check(self.ctx, self.L.cac_centro_atoms(self.ctx, int(nnn), _p(out)))
        return out

    def interpolated_sites(self) -> np.ndarray:
        """Every site of every element, [ne * nsite][3] in element and
        inter_ele order (Element::interpolate_atom_vec, element.cpp:1241-1265)."""
        x = np.zeros((self.ne * self.nsite, 3))
        if x.size:
            _check(self.ctx, self.L.cac_interpolate_sites(self.ctx, _p(x)))
        return x

    def centro_of_sites(self, nnn: int = 12, device: int = 0):
        """The reference's executed path (verlet.cpp:91-94): the atoms are
        replaced by the interpolated element sites (element.cpp:1226-1233),
        the atom list is built over them and their centro-symmetry computed.
        Runs in a second context with this one's box, pair style and neighbor
        settings; returns (site positions, values)."""
        if self.box is None or self.pair is None or self.neighbor_args is None:
            raise PairError("centro_of_sites needs the box, the pair style and the neighbor settings")
        xs = self.interpolated_sites()
        with Engine(device) as e2:
            e2.set_box(*self.box)
            e2.upload_atoms(xs, None, self.mass)
            e2.set_pair(self.pair)
            cutneigh, skin, every, delay, check = self.neighbor_args
            e2.set_neighbor(cutneigh, skin, every, delay, check)
            e2.build()
            c = e2.centro(nnn)
        return xs, c

    def write_dump_xyz(self, path: str, nnn: int = 12, lo: float = 1.0, hi: float = 10.0) -> int:
        """Write the reference's xyz dump: a count line, "XYZ", then
        "Cu x y z c" (%g) for every atom with lo < c < hi (Dump::count and
        DumpXyz::pack_noscale_noimage select 1 < c < 10).  Returns the count."""
        c = self.centro(nnn)
        x, _, _ = self.atoms()
        sel = np.nonzero((c > lo) & (c < hi))[0]
        with open(path, "w") as fh:
            fh.write(f"{len(sel)}\nXYZ\n")
            for i in sel:
                fh.write("Cu %g %g %g %g\n" % (x[i, 0], x[i, 1], x[i, 2], c[i]))
        return len(sel)

    def write_dump_atom(self, path: str, step: int, append: bool = False, tag=None, nodetag=None) -> int:
        """One atom-dump snapshot of the current state (DumpAtom,
        dump_atom.cpp:35-157): atoms, then element nodes as type-1 rows."""
        from .dump import write_dump_atom
        x, v, f = self.atoms()
        nx, nv, nf = self.elements()
        return write_dump_atom(path, step, self.box[0], self.box[1], x, f, v, nx, nf, nv, tag=tag, nodetag=nodetag,
                               append=append)

    def write_dump_element(self, path: str, step: int, append: bool = False) -> int:
        """One element-dump snapshot of the current state (DumpElement,
        dump_element.cpp:72-263)."""
        from .dump import write_dump_element
        nx, nv, nf = self.elements()
        return write_dump_element(path, step, nx, nv, nf, append=append)

    def capacities(self):
        a, b = C.c_int(), C.c_int()
        _check(self.ctx, self.L.cac_get_capacities(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def atom_neighbors(self):
        """Atom lists as uint64 entries j | img << 27 | site << 32 (site ==
        ATOM_SITE for an atom j, else site `site` of element j)."""
        maxn, _ = self.capacities()
        num = np.zeros(self.n, dtype=np.int32)
        nb = np.zeros((max(self.n, 1), maxn), dtype=np.uint64)
        _check(self.ctx, self.L.cac_get_atom_neighbors(self.ctx, maxn, _p(num, _lib._i), _p(nb, _lib._u64)))
        return num, nb[:self.n]

    def ip_neighbors(self):
        _, maxn = self.capacities()
        nip = self.ne * self.nint
        num = np.zeros(nip, dtype=np.int32)
        nb = np.zeros((max(nip, 1), maxn), dtype=np.uint64)
        _check(self.ctx, self.L.cac_get_ip_neighbors(self.ctx, maxn, _p(num, _lib._i), _p(nb, _lib._u64)))
        return num, nb[:nip]

    def rho(self):
        r = np.zeros(self.n); fp = np.zeros(self.n)
        _check(self.ctx, self.L.cac_get_rho(self.ctx, _p(r), _p(fp)))
        return r, fp

    def erho(self):
        e = np.zeros((self.ne, 8))
        _check(self.ctx, self.L.cac_get_erho(self.ctx, _p(e)))
        return e

    def ip_forces(self):
        f = np.zeros((self.ne * self.nint, 3))
        _check(self.ctx, self.L.cac_get_ip_forces(self.ctx, _p(f)))
        return f

    def profile(self, on: bool = True):
        _check(self.ctx, self.L.cac_profile_enable(self.ctx, int(on)))

    def profile_avg_ms(self, name: str):
        cnt = C.c_int()
        ms = self.L.cac_profile_avg_ms(self.ctx, name.encode(), C.byref(cnt))
        return ms, cnt.value

