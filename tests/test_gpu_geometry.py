"""GPU parity at the element sizes of the EAM workloads, against the oracle.

* nae = 11 (the 500k-atom + 20k-element EAM system's elements): a 33-cell box
  with an atomistic core, plain and coupled.  The coupled interface clusters
  have list unions above the 256 LDS rows of the cluster passes, so the
  branch that keeps a cluster's rows in global scratch runs (asserted through
  cac_cluster_info).  Neighbor sets bit-exact; FP64 rho, F', erho,
  integration-point forces, atom and node forces, energy and virial within
  1e-10 norm-relative.  The coupled forces also obey Newton's third law
  across the interface on the device: atom forces plus nodal reactions sum to
  zero.
* The same system in mixed precision (FP32 pair and embedding arithmetic,
  FP64 sums) against the FP64 oracle: force errors stated and bounded.
* nae = 21 (the 8M-atom + 200k-element system's elements) with the
  edge-dislocation dipole and the 1e8/s strain ramp: steps of cac_run, then
  lists and forces at the reached positions against the oracle.
"""
import dataclasses

import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM, _lib
from cac_code_amd.lattice import build_cac_system, edge_dipole

from gpu_common import CUTNEIGH_EAM, assert_same_sets, jitter, oracle_state, rel_err, rms_rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
US_SMALL = 256      # LDS rows of the cluster passes (larger unions use global scratch)


def eam_pair(cu_funcfl):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    return p


def nae11_system():
    return jitter(build_cac_system(33, 11, 8.0, temperature=300.0, seed=4928459), 0.05, 12)


def nae21_system():
    s = build_cac_system(42, 21, 10.0, temperature=300.0, seed=4928459)
    edge_dipole(s)
    return s


def engine(s, cu_funcfl, coupled, skin=1.0):
    eng = Engine(0)
    eng.load_system(s)
    eng.set_pair(eam_pair(cu_funcfl))
    eng.set_neighbor(CUTNEIGH_EAM, skin=skin, every=1, delay=0, check=True)
    eng.set_coupling(coupled)
    eng.build()
    return eng


def check_sets(eng, ref, s):
    num, nb = eng.atom_neighbors()
    assert_same_sets(num, nb, ref["anum"], ref["anb"], "atoms")
    num, nb = eng.ip_neighbors()
    assert_same_sets(num, nb, ref["inum"], ref["inb"], "integration points")


def check_fp64(eng, ref):
    eng.force(eflag=True, vflag=True)
    e, w = eng.energy_virial()
    rho, fp = eng.rho()
    _, _, f = eng.atoms()
    _, _, nodef = eng.elements()
    a, el = ref["atoms"], ref["elems"]
    errs = dict(rho=rel_err(rho, a["rho"]), fp=rel_err(fp, a["fp"]), f=rel_err(f, a["f"]),
                erho=rel_err(eng.erho(), el["erho"]), ipf=rel_err(eng.ip_forces(), el["ipf"]),
                nodef=rel_err(nodef, el["nodef"]))
    print("  fp64 errors:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < TOL, k
    assert abs(e[0] - a["eatom"].sum()) <= TOL * abs(a["eatom"].sum())
    assert abs(e[1] - el["eip"].sum()) <= TOL * abs(el["eip"].sum())
    assert rel_err(w, a["vatom"].sum(axis=0) + el["vip"].sum(axis=0)) < TOL
    return f, nodef


@pytest.mark.parametrize("coupled", [False, True], ids=["plain", "coupled"])
def test_nae11_fp64_parity(orc, cu_funcfl, cu_tables, coupled):
    s = nae11_system()
    print(f"[nae11 fp64] coupled={coupled}: atoms={s.n_atoms} elements={s.n_elements}", flush=True)
    ref = oracle_state(orc, s, CUTNEIGH_EAM, tables=cu_tables, coupled=coupled, maxneigh=256, maxnip=512)
    eng = engine(s, cu_funcfl, coupled)
    try:
        ncl, active, umax = eng.cluster_info()
        print(f"  clusters {ncl}, active {active}, largest union {umax}", flush=True)
        assert active
        if coupled:
            assert umax > US_SMALL          # the global-scratch rows of the cluster passes ran
        check_sets(eng, ref, s)
        f, nodef = check_fp64(eng, ref)
        if coupled:
            # Newton's third law across the interface, on the device's forces:
            # nodal reactions = node force - integration-point scatter
            ipf = eng.ip_forces().reshape(s.n_elements, -1, 3)
            es = ref["es"]
            scat = np.einsum("gk,egd->ekd", es.weight[:, None] * es.spa_int, ipf)
            total = f.sum(axis=0) + (nodef - scat).reshape(-1, 3).sum(axis=0)
            assert np.linalg.norm(f.sum(axis=0)) > 1e-3 * np.abs(f).max()
            assert np.abs(total).max() < 1e-11 * np.abs(f).max() * s.n_atoms
    finally:
        eng.close()


def test_nae11_mixed_against_oracle(orc, cu_funcfl, cu_tables):
    """Mixed precision (FP32 pair / embedding math, FP64 sums) against the
    FP64 oracle on the coupled nae=11 system."""
    s = nae11_system()
    ref = oracle_state(orc, s, CUTNEIGH_EAM, tables=cu_tables, coupled=True, maxneigh=256, maxnip=512)
    eng = engine(s, cu_funcfl, True)
    try:
        eng.force(eflag=True, vflag=True, precision=_lib.PREC_MIXED)
        e, w = eng.energy_virial()
        _, _, f = eng.atoms()
        _, _, nodef = eng.elements()
    finally:
        eng.close()
    a, el = ref["atoms"], ref["elems"]
    out = dict(atom_rms=rms_rel(f, a["f"]), atom_max=rel_err(f, a["f"]),
               node_rms=rms_rel(nodef, el["nodef"]), node_max=rel_err(nodef, el["nodef"]),
               energy=abs(e.sum() - (a["eatom"].sum() + el["eip"].sum())) / abs(a["eatom"].sum() + el["eip"].sum()))
    print("  mixed vs FP64 oracle:", {k: f"{v:.2e}" for k, v in out.items()}, flush=True)
    # FP32 pair arithmetic: each pair force to ~1e-7; net forces are small
    # differences of ~40 pair terms, so their relative error is larger
    assert out["atom_rms"] < 2e-5 and out["atom_max"] < 1e-4
    assert out["node_rms"] < 2e-4 and out["node_max"] < 1e-3
    assert out["energy"] < 1e-6


@pytest.mark.parametrize("coupled", [False, True], ids=["plain", "coupled"])
def test_nae21_dipole_strain_parity(orc, cu_funcfl, cu_tables, coupled):
    """nae=21 elements, the Volterra dipole and the 1e8/s uniaxial strain ramp:
    10 steps of cac_run (mixed, as the large workload runs), then the device's
    lists and FP64 forces at the reached positions against the oracle's."""
    s = nae21_system()
    print(f"[nae21] coupled={coupled}: atoms={s.n_atoms} elements={s.n_elements}", flush=True)
    eng = engine(s, cu_funcfl, coupled, skin=0.2)
    try:
        eng.set_deform_rate([1e8 * 1e-12, 0.0, 0.0])            # per ps
        nb = eng.run(10, precision=_lib.PREC_MIXED)
        print(f"  rebuilds in 10 steps: {nb}", flush=True)
        eng.build()
        x, _, _ = eng.atoms()
        nodex, _, _ = eng.elements()
        # the box the device has reached: cac_run's own accumulation of the
        # elapsed time and of the per-step scales (double arithmetic, as here)
        rate = [1e8 * 1e-12, 0.0, 0.0]
        L0 = s.box_hi - s.box_lo
        prd = L0.copy()
        t = 0.0
        for _ in range(10):
            t += 0.001
            prd = np.array([prd[d] * (L0[d] * (1.0 + rate[d] * t) / prd[d]) for d in range(3)])
        r = dataclasses.replace(s, x=x, nodex=nodex, box_hi=s.box_lo + prd)
        ref = oracle_state(orc, r, CUTNEIGH_EAM, tables=cu_tables, coupled=coupled, maxneigh=256, maxnip=512)
        check_sets(eng, ref, r)
        check_fp64(eng, ref)
    finally:
        eng.close()
