"""Element passes by vertex cluster (cac_set_ip_clusters) against the
per-integration-point passes, FP64 and mixed, with and without atom-element
coupling, and the clusters' own checks.  The two are separate kernels that
the compiler fuses multiplies and adds in independently (e.g. which square
of r^2 = dx*dx + dy*dy + dz*dz is fused), so they agree to rounding:
1e-12 norm-relative, against the oracle's 1e-10 bar."""
import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM, _lib
from cac_code_amd.lattice import vertex_clusters

from gpu_common import CUTNEIGH_EAM, rel_err, small_system

TOL = 1e-12

pytestmark = pytest.mark.gpu

SYSTEMS = {
    "small": lambda: small_system(seed=7, L=10, nae=5, core=4.0),
    "elements_only": lambda: small_system(seed=11, L=10, nae=5, core=0.0),
    "medium": lambda: small_system(seed=3, L=20, nae=5, core=8.0, sigma=0.08),
    "finest": lambda: small_system(seed=5, L=8, nae=2, core=3.0, sigma=0.03),
}


def eam_pair(cu_funcfl):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    return p


def forces(s, pair, clusters, coupled, precision, ev, cutneigh=CUTNEIGH_EAM):
    eng = Engine(0)
    try:
        eng.load_system(s)
        if not clusters:
            eng.set_ip_clusters(np.zeros(1, np.int32), np.zeros(0, np.int32))
        eng.set_pair(pair)
        eng.set_neighbor(cutneigh, every=1, delay=0, check=True)
        eng.set_coupling(coupled)
        eng.build()
        ncl, active, umax = eng.cluster_info()
        assert active == clusters, (ncl, active, umax)
        eng.force(eflag=ev, vflag=ev, precision=precision)
        _, _, f = eng.atoms()
        _, _, nf = eng.elements()
        out = [f, nf, eng.erho(), eng.ip_forces()]
        if ev:
            e, w = eng.energy_virial()
            out += [np.asarray(e), np.asarray(w)]
        eng.run(3, precision=precision)
        x, v, f2 = eng.atoms()
        nx, nv, nf2 = eng.elements()
        return out + [x, v, f2, nx, nv, nf2]
    finally:
        eng.close()


@pytest.mark.parametrize("precision", [_lib.PREC_FP64, _lib.PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("coupled", [False, True], ids=["plain", "coupled"])
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_clusters_equal_per_point_passes(name, coupled, precision, cu_funcfl):
    s = SYSTEMS[name]()
    print(f"[clusters] {name} coupled={coupled} prec={precision}: atoms={s.n_atoms} elements={s.n_elements}",
          flush=True)
    pair = eam_pair(cu_funcfl)
    ev = precision == _lib.PREC_FP64
    a = forces(s, pair, True, coupled, precision, ev)
    b = forces(s, pair, False, coupled, precision, ev)
    names = ["f", "nodef", "erho", "ipf", "energy", "virial", "x", "v", "f3", "nodex", "nodev", "nodef3"]
    names = names if len(a) == len(names) else names[:4] + names[6:]
    diff = {n: rel_err(qa, qb) for n, qa, qb in zip(names, a, b)}
    print("  max |cluster - per point| / max |per point|:", {k: f"{v:.1e}" for k, v in diff.items()}, flush=True)
    bad = [k for k, v in diff.items() if not v <= TOL]
    assert not bad, f"differ: {bad}"


@pytest.mark.parametrize("precision", [_lib.PREC_FP64, _lib.PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("coupled", [False, True], ids=["plain", "coupled"])
def test_clusters_long_lists_equal_per_point_passes(coupled, precision, cu_funcfl):
    """A 7.5 A list cutoff gives integration-point lists of about 150 entries,
    more than 12 steps per lane: the cluster force pass then walks every step
    with the skipped entries masked, instead of the kept entries from
    registers that the 5.95 A lists use."""
    s = SYSTEMS["small"]()
    print(f"[clusters] long lists coupled={coupled} prec={precision}: atoms={s.n_atoms} elements={s.n_elements}",
          flush=True)
    pair = eam_pair(cu_funcfl)
    ev = precision == _lib.PREC_FP64
    a = forces(s, pair, True, coupled, precision, ev, cutneigh=7.5)
    b = forces(s, pair, False, coupled, precision, ev, cutneigh=7.5)
    names = ["f", "nodef", "erho", "ipf", "energy", "virial", "x", "v", "f3", "nodex", "nodev", "nodef3"]
    names = names if len(a) == len(names) else names[:4] + names[6:]
    diff = {n: rel_err(qa, qb) for n, qa, qb in zip(names, a, b)}
    print("  max |cluster - per point| / max |per point|:", {k: f"{v:.1e}" for k, v in diff.items()}, flush=True)
    bad = [k for k, v in diff.items() if not v <= TOL]
    assert not bad, f"differ: {bad}"


def test_cluster_argument_checks():
    s = small_system(seed=7, L=10, nae=5, core=4.0)
    eng = Engine(0)
    try:
        eng.load_system(s)
        start, slots = vertex_clusters(s)
        bad = slots.copy()
        bad[1] = bad[0]                                     # a slot twice
        with pytest.raises(_lib.CacError):
            eng.set_ip_clusters(start, bad)
        with pytest.raises(_lib.CacError):
            eng.set_ip_clusters(np.array([0, 9], np.int32), np.arange(9, dtype=np.int32))  # 9 slots, wrong cover
        eng.set_ip_clusters(start, slots)
    finally:
        eng.close()
