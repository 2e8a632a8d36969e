"""The device-resident Verlet loop (cac_run) against the same step sequence
issued call by call through the C-ABI, and against the oracle's NVE.

cac_run fuses the final half-step of one step with the initial half-step of
the next, writes the rebuild check to host memory and queues the force before
reading it (recomputing after a rebuild).  None of that may change a bit:
the call-by-call sequence (nve_initial, decide, [build], force, nve_final per
step; verlet.cpp:149-219) must give identical positions, velocities and
rebuild counts, with rebuilds forced often by a small skin.
"""
import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM, _lib
from cac_code_amd.lattice import FTM2V_METAL

from gpu_common import CUTNEIGH_EAM, rel_err, small_system

pytestmark = pytest.mark.gpu

DT = 0.001
DTF = 0.5 * DT * FTM2V_METAL


def eam_pair(cu_funcfl):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    return p


def engine(s, pair, skin, every, delay, check, coupled):
    eng = Engine(0)
    eng.load_system(s)
    eng.set_pair(pair)
    eng.set_neighbor(CUTNEIGH_EAM, skin=skin, every=every, delay=delay, check=check)
    eng.set_coupling(coupled)
    eng.build()
    eng.force()
    return eng


def stepwise(eng, nsteps, precision):
    builds = 0
    for _ in range(nsteps):
        eng.nve_initial(DT, DTF)
        if eng.decide():
            eng.build()
            builds += 1
        eng.force(precision=precision)
        eng.nve_final(DTF)
    return builds


def state(eng):
    x, v, f = eng.atoms()
    nx, nv, nf = eng.elements()
    return [x, v, f, nx, nv, nf]


@pytest.mark.parametrize("precision", [_lib.PREC_FP64, _lib.PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("every,delay,check", [(1, 0, True), (2, 4, True), (5, 0, False)],
                         ids=["every1", "every2_delay4", "nocheck5"])
def test_run_equals_stepwise_api(cu_funcfl, precision, every, delay, check):
    s = small_system(seed=21, L=10, nae=5, core=4.0, sigma=0.02, temperature=600.0)
    pair = eam_pair(cu_funcfl)
    nsteps = 30
    print(f"[run] prec={precision} every={every} delay={delay} check={check}: atoms={s.n_atoms} "
          f"elements={s.n_elements}", flush=True)
    # trigger 0.04 A: rebuilds every few steps at 600 K
    a = engine(s, pair, 0.08, every, delay, check, True)
    b = engine(s, pair, 0.08, every, delay, check, True)
    try:
        na = a.run(nsteps, dt=DT, precision=precision)
        nb = stepwise(b, nsteps, precision)
        assert na == nb and na > 0, (na, nb)
        for qa, qb, name in zip(state(a), state(b), ["x", "v", "f", "nodex", "nodev", "nodef"]):
            assert np.array_equal(qa, qb), name
    finally:
        a.close()
        b.close()


def test_run_matches_oracle_nve(orc, cu_funcfl, cu_tables):
    """Five coupled EAM steps without a rebuild against the oracle's FixNVE
    and forces (1e-10 norm-relative)."""
    s = small_system(seed=23, L=10, nae=5, core=4.0, sigma=0.03, temperature=300.0)
    pair = eam_pair(cu_funcfl)
    eng = engine(s, pair, 1.0, 1, 0, True, True)
    try:
        nb = eng.run(5, dt=DT)
        assert nb == 0
        got = state(eng)
    finally:
        eng.close()
    es = orc.ElementSet(s.nodex.copy(), s.nae, s.intp_ijk, s.intp_w)
    x, v, nodev = s.x.copy(), s.v.copy(), s.nodev.copy()
    anum, anb, inum, inb = orc.coupled_neighbors(x, es, s.box_lo, s.box_hi, CUTNEIGH_EAM, 2, 200, 200)

    def force():
        r = orc.eam_coupled(x, es, s.box_lo, s.box_hi, anum, anb, inum, inb, cu_tables)
        return r["f"], r["nodef"]

    f, nodef = force()
    xc = np.zeros((s.n_elements, 3))
    for _ in range(5):
        orc.nve_atoms_initial(x, v, f, DT, DTF, s.mass)
        orc.nve_nodes_initial(es.nodex, nodev, nodef, xc, DT, DTF, s.mass_e)
        f, nodef = force()
        orc.nve_atoms_final(v, f, DTF, s.mass)
        orc.nve_nodes_final(nodev, nodef, DTF, s.mass_e)
    for g, r, name in zip(got, [x, v, f, es.nodex, nodev, nodef], ["x", "v", "f", "nodex", "nodev", "nodef"]):
        assert rel_err(g, r) < 1e-10, name


def test_deform_matches_rescaled_system(cu_funcfl):
    """cac_deform maps atoms, nodes and the box affinely: the forces equal
    those of the system built directly at the deformed coordinates."""
    s = small_system(seed=25, L=10, nae=5, core=4.0, sigma=0.03, temperature=300.0)
    scale = np.array([1.002, 0.999, 1.001])
    pair = eam_pair(cu_funcfl)
    a = engine(s, pair, 1.0, 1, 0, True, True)
    try:
        a.deform(scale)
        a.force()
        fa = state(a)
    finally:
        a.close()
    t = small_system(seed=25, L=10, nae=5, core=4.0, sigma=0.03, temperature=300.0)
    t.x = np.ascontiguousarray(t.box_lo + (t.x - t.box_lo) * scale)
    t.nodex = np.ascontiguousarray(t.box_lo + (t.nodex - t.box_lo) * scale)
    t.box_hi = t.box_lo + (t.box_hi - t.box_lo) * scale
    b = engine(t, pair, 1.0, 1, 0, True, True)
    try:
        fb = state(b)
    finally:
        b.close()
    for qa, qb, name in zip(fa, fb, ["x", "v", "f", "nodex", "nodev", "nodef"]):
        assert rel_err(qa, qb) < 1e-10, name


@pytest.mark.parametrize("core", [0.0, 100.0], ids=["elements", "atoms"])
def test_strain_rate_ramp_in_run(cu_funcfl, core):
    """A constant engineering strain rate in cac_run: a perfect, motionless
    crystal with no force on any particle (all elements, or all atoms) is
    stretched exactly to L0*(1 + rate*t).  (An atom-element interface is
    not force-free: the coupling carries ghost forces there.)"""
    from cac_code_amd.lattice import build_cac_system
    s = build_cac_system(10, 5, core, temperature=0.0)
    pair = eam_pair(cu_funcfl)
    rate = np.array([1e-2, 0.0, 0.0])               # per ps
    nsteps = 10
    eng = engine(s, pair, 1.0, 1, 0, True, True)
    try:
        eng.set_deform_rate(rate)
        eng.run(nsteps, dt=DT)
        x, v, _ = eng.atoms()
        nx, _, _ = eng.elements()
    finally:
        eng.close()
    stretch = 1.0 + rate * nsteps * DT
    assert s.n_atoms == 0 or s.n_elements == 0
    assert np.allclose(x, s.box_lo + (s.x - s.box_lo) * stretch, rtol=0, atol=1e-9)
    assert np.allclose(nx, s.box_lo + (s.nodex - s.box_lo) * stretch, rtol=0, atol=1e-9)


def test_run_strain_rate_equals_stepwise(cu_funcfl):
    """cac_run under a strain rate (the remap fused into the step kernels)
    equals cac_deform at the start of each step of the call-by-call
    sequence, bit for bit, with rebuilds."""
    s = small_system(seed=27, L=10, nae=5, core=4.0, sigma=0.02, temperature=600.0)
    pair = eam_pair(cu_funcfl)
    rate = np.array([2e-2, -1e-2, 0.5e-2])   # per ps
    nsteps = 20
    a = engine(s, pair, 0.08, 1, 0, True, True)
    b = engine(s, pair, 0.08, 1, 0, True, True)
    try:
        a.set_deform_rate(rate)
        na = a.run(nsteps, dt=DT)
        l0 = s.box_hi - s.box_lo
        prd = l0.copy()
        t = 0.0
        nb = 0
        for _ in range(nsteps):
            t += DT
            scale = np.array([l0[d] * (1.0 + rate[d] * t) / prd[d] for d in range(3)])
            prd = np.array([prd[d] * scale[d] for d in range(3)])
            b.deform(scale)
            b.nve_initial(DT, DTF)
            if b.decide():
                b.build()
                nb += 1
            b.force()
            b.nve_final(DTF)
        assert na == nb and na > 0, (na, nb)
        for qa, qb, name in zip(state(a), state(b), ["x", "v", "f", "nodex", "nodev", "nodef"]):
            assert np.array_equal(qa, qb), name
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("precision", [_lib.PREC_FP64, _lib.PREC_MIXED], ids=["fp64", "mixed"])
def test_run_inner_mask_equals_stepwise(cu_funcfl, precision):
    """Hot run with a wide rebuild trigger: atoms cross the density pass's
    inner-mask margin between rebuilds, so cac_run's density passes switch
    between inner and full gathers; the call-by-call path (full gathers
    every step) must give the same bits."""
    s = small_system(seed=29, L=10, nae=5, core=4.0, sigma=0.02, temperature=1500.0)
    pair = eam_pair(cu_funcfl)
    nsteps = 40
    a = engine(s, pair, 0.9, 1, 0, True, True)
    b = engine(s, pair, 0.9, 1, 0, True, True)
    try:
        na = a.run(nsteps, dt=DT, precision=precision)
        nb = stepwise(b, nsteps, precision)
        assert na == nb, (na, nb)
        for qa, qb, name in zip(state(a), state(b), ["x", "v", "f", "nodex", "nodev", "nodef"]):
            assert np.array_equal(qa, qb), name
    finally:
        a.close()
        b.close()


def _forces(eng):
    eng.force(eflag=True, vflag=True)
    _, _, f = eng.atoms()
    _, _, nf = eng.elements()
    return [f, nf, eng.erho(), np.asarray(eng.energy_virial()[0])]


def _fresh(s, pair, coupled=True):
    eng = Engine(0)
    try:
        eng.load_system(s)
        eng.set_pair(pair)
        eng.set_neighbor(CUTNEIGH_EAM, skin=1.0)
        eng.set_coupling(coupled)
        eng.build()
        return _forces(eng)
    finally:
        eng.close()


def test_reload_into_one_engine(cu_funcfl):
    """Systems of different sizes loaded one after another into one Engine
    give the forces of a fresh context bit for bit: the vertex clusters,
    integration-point buffers and lists of the earlier mesh are not reused
    (a smaller mesh, then a mesh without clusters)."""
    from cac_code_amd.lattice import build_cac_system
    pair = eam_pair(cu_funcfl)
    big = small_system(seed=3, L=20, nae=5, core=8.0, sigma=0.05)
    small = small_system(seed=7, L=10, nae=5, core=4.0)
    bare = small_system(seed=9, L=10, nae=5, core=6.0)
    bare.blocks = None                        # no mesh vertices known: no clusters
    eng = Engine(0)
    try:
        for s in (big, small, bare):
            print(f"[reload] atoms={s.n_atoms} elements={s.n_elements}", flush=True)
            eng.load_system(s)
            eng.set_pair(pair)
            eng.set_neighbor(CUTNEIGH_EAM, skin=1.0)
            eng.set_coupling(True)
            eng.build()
            got = _forces(eng)
            assert eng.cluster_info()[1] == (s.blocks is not None)
            for a, b in zip(got, _fresh(s, pair)):
                assert np.array_equal(a, b)
    finally:
        eng.close()


def test_failed_build_refuses_force(cu_funcfl):
    """A build that fails leaves no lists in use: after a remap makes the box
    narrower than three neighbor cutoffs, cac_build_neighbors fails and
    cac_force reports that the lists are not built instead of running on the
    old ones."""
    s = small_system(seed=7, L=10, nae=5, core=4.0)
    eng = engine(s, eam_pair(cu_funcfl), 1.0, 1, 0, True, True)
    try:
        eng.deform([0.3, 0.3, 0.3])
        with pytest.raises(_lib.CacError) as e1:
            eng.build()
        assert e1.value.code == _lib.CAC_ERR_BOX
        with pytest.raises(_lib.CacError) as e2:
            eng.force()
        assert "not built" in str(e2.value)
    finally:
        eng.close()
