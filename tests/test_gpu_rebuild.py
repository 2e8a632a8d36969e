"""Device rebuilds against the oracle: Neighbor::decide / check_distance
(neighbor.cpp:1150-1217), Domain::pbc / elepbc (domain.cpp:535-745) and the
lists built after them.

A hot (600 K) coupled system drifts along +x and -y, so atoms and element
centers cross box faces, with a small skin so lists are rebuilt every few
steps.  Step by step through the C-ABI (nve_initial, decide, [build], force,
nve_final): at every step the oracle decides on the downloaded positions
with its own held positions (those of the last build); the device's decision
must be the same.  At every rebuild the oracle wraps the downloaded positions
(orc_pbc_atoms, orc_pbc_elements on the node-mean centers); the device's
wrapped atoms and nodes must equal them bit for bit, its lists must equal the
oracle's lists over them (sorted sets), and its forces the oracle's to
1e-10.  Then cac_run over the same steps must make the same number of
rebuilds and end in the same state as the step-by-step sequence.
"""
import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM
from cac_code_amd.lattice import FTM2V_METAL, build_cac_system

from gpu_common import CUTNEIGH_EAM, assert_same_sets, jitter, oracle_state, rel_err

pytestmark = pytest.mark.gpu

DT = 0.001
DTF = 0.5 * DT * FTM2V_METAL
SKIN = 0.12                      # rebuild trigger 0.06 A


def eam_pair(cu_funcfl):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    return p


def drifting_system():
    """nae=5 box with a core; shifted so the lowest atom plane along x and
    the highest along y sit 0.02 A inside the faces, plus a drift of
    (+30, -30, 0) A/ps on every atom and node so those planes and the
    element centers next to the faces cross them within a few steps."""
    s = jitter(build_cac_system(10, 5, 4.0, temperature=600.0, seed=91), 0.02, 92)
    L = s.box_hi - s.box_lo
    shift = np.zeros(3)
    xc = s.nodex.mean(axis=1)
    # an element center 0.04 A below the x-hi face, the lowest atom y 0.02 A above y-lo
    shift[0] = (s.box_hi[0] - 0.04) - xc[:, 0].max()
    shift[1] = (s.box_lo[1] + 0.02) - s.x[:, 1].min()
    s.x = np.ascontiguousarray(np.mod(s.x + shift - s.box_lo, L) + s.box_lo)
    s.nodex = np.ascontiguousarray(s.nodex + shift)
    drift = np.array([30.0, -30.0, 0.0])
    s.v = np.ascontiguousarray(s.v + drift)
    s.nodev = np.ascontiguousarray(s.nodev + drift)
    return s


def engine(s, cu_funcfl):
    eng = Engine(0)
    eng.load_system(s)
    eng.set_pair(eam_pair(cu_funcfl))
    eng.set_neighbor(CUTNEIGH_EAM, skin=SKIN, every=1, delay=0, check=True)
    eng.set_coupling(True)
    eng.build()
    eng.force()
    return eng


def test_rebuilds_match_oracle(orc, cu_funcfl, cu_tables):
    s = drifting_system()
    print(f"[rebuild] atoms={s.n_atoms} elements={s.n_elements}", flush=True)
    trig = 0.25 * SKIN * SKIN
    lo, hi = s.box_lo, s.box_hi
    eng = engine(s, cu_funcfl)
    try:
        x, _, _ = eng.atoms()
        nodex, _, _ = eng.elements()
        xh, nh = x.copy(), nodex.copy()          # the oracle's held positions (at the build)
        nbuild = 0
        wraps = np.zeros(2, dtype=int)
        nsteps = 30
        for step in range(nsteps):
            eng.nve_initial(DT, DTF)
            x, _, _ = eng.atoms()
            nodex, _, _ = eng.elements()
            want = (orc.displacement_exceeds(x, xh, trig) or
                    orc.displacement_exceeds(nodex.reshape(-1, 3), nh.reshape(-1, 3), trig))
            got = eng.decide()
            assert got == want, f"step {step}: device decides {got}, oracle {want}"
            if got:
                nbuild += 1
                xw = x.copy()
                orc.pbc_atoms(xw, lo, hi)
                nw = nodex.copy()
                xc = orc.element_centers(nw)
                orc.pbc_elements(xc, nw, lo, hi)
                wraps += [int(np.any(xw != x)), int(np.any(nw != nodex))]
                eng.build()
                xd, _, _ = eng.atoms()
                nd, _, _ = eng.elements()
                assert np.array_equal(xd, xw), f"step {step}: wrapped atoms differ"
                assert np.array_equal(nd, nw), f"step {step}: wrapped nodes differ"
                r = s.__class__(**{**s.__dict__, "x": xd, "nodex": nd})
                ref = oracle_state(orc, r, CUTNEIGH_EAM, tables=cu_tables, coupled=True)
                num, nb = eng.atom_neighbors()
                assert_same_sets(num, nb, ref["anum"], ref["anb"], f"atoms after rebuild {nbuild}")
                num, nb = eng.ip_neighbors()
                assert_same_sets(num, nb, ref["inum"], ref["inb"], f"integration points after rebuild {nbuild}")
                eng.force()
                _, _, f = eng.atoms()
                _, _, nf = eng.elements()
                assert rel_err(f, ref["atoms"]["f"]) < 1e-10
                assert rel_err(nf, ref["elems"]["nodef"]) < 1e-10
                xh, nh = xd.copy(), nd.copy()
            else:
                eng.force()
            eng.nve_final(DTF)
        print(f"  {nbuild} rebuilds in {nsteps} steps; rebuilds that wrapped atoms / nodes: {wraps}", flush=True)
        assert nbuild >= 3
        assert wraps[0] > 0 and wraps[1] > 0          # both wraps were exercised
        stepwise = [*eng.atoms(), *eng.elements()]
    finally:
        eng.close()
    eng = engine(s, cu_funcfl)
    try:
        nb_run = eng.run(nsteps, dt=DT)
        assert nb_run == nbuild
        for qa, qb, name in zip([*eng.atoms(), *eng.elements()], stepwise, ["x", "v", "f", "nodex", "nodev", "nodef"]):
            assert np.array_equal(qa, qb), name
    finally:
        eng.close()


def test_compressive_remap_rebuilds(orc, cu_funcfl, cu_tables):
    """A compressive remap brings pairs from beyond cutneigh inside the force
    cutoff without any motion: the rebuild check must account for it.  Scale
    0.82 takes pairs just beyond the 5.95 A list cutoff to 4.88 A, inside the
    4.95 A force cutoff, more than the whole skin of 1 A allows: decide must
    ask for a rebuild, and the forces then equal the oracle's at the
    compressed coordinates.  A mild compression (0.995) only shrinks the
    allowance (0.5 A - 0.005*5.95/2 A) and keeps the lists."""
    s = jitter(build_cac_system(10, 5, 4.0, temperature=0.0), 0.02, 93)
    eng = Engine(0)
    try:
        eng.load_system(s)
        eng.set_pair(eam_pair(cu_funcfl))
        eng.set_neighbor(CUTNEIGH_EAM, skin=1.0, every=1, delay=0, check=True)
        eng.set_coupling(True)
        eng.build()
        eng.deform([0.82, 0.82, 0.82])
        assert eng.decide()                      # the remap alone used up the skin
        eng.build()
        eng.force()
        _, _, f = eng.atoms()
        _, _, nf = eng.elements()
        x, _, _ = eng.atoms()
        nodex, _, _ = eng.elements()
        # a mild compression keeps the lists: the allowance shrinks instead
        eng.deform([0.995, 1.0, 1.0])
        assert not eng.decide()
    finally:
        eng.close()
    r = s.__class__(**{**s.__dict__, "x": x, "nodex": nodex, "box_hi": s.box_lo + 0.82 * (s.box_hi - s.box_lo)})
    ref = oracle_state(orc, r, CUTNEIGH_EAM, tables=cu_tables, coupled=True)
    assert rel_err(f, ref["atoms"]["f"]) < 1e-10
    assert rel_err(nf, ref["elems"]["nodef"]) < 1e-10


def test_compressive_rate_in_run(cu_funcfl):
    """cac_run under a compressive strain rate: with a 0.3 A skin a 5%
    compression since the build uses up the 0.15 A allowance, so rebuilds
    come from the remap as well as from motion; the call-by-call path
    (cac_deform, nve_initial, decide, build, force, nve_final per step)
    makes the same ones and ends in the same state bit for bit."""
    s = jitter(build_cac_system(10, 5, 4.0, temperature=300.0, seed=95), 0.02, 96)
    rate = np.array([-5.0, 0.0, 0.0])             # per ps: 0.5% per step along x
    nsteps = 20
    engs = []
    try:
        for _ in range(2):
            e = Engine(0)
            engs.append(e)
            e.load_system(s)
            e.set_pair(eam_pair(cu_funcfl))
            e.set_neighbor(CUTNEIGH_EAM, skin=0.3, every=1, delay=0, check=True)
            e.set_coupling(True)
            e.build()
            e.force()
        a, b = engs
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
        print(f"[compressive rate] rebuilds: run {na}, call by call {nb}", flush=True)
        assert na == nb and na >= 2
        for qa, qb, name in zip([*a.atoms(), *a.elements()], [*b.atoms(), *b.elements()],
                                ["x", "v", "f", "nodex", "nodev", "nodef"]):
            assert np.array_equal(qa, qb), name
    finally:
        for e in engs:
            e.close()
