"""A system read from a CAC data file runs on the device exactly as the
generated one: same forces bit for bit after a force call and steps, and the
same integration-point clusters (from the node tags)."""
import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM
from cac_code_amd import datafile as D

from gpu_common import CUTNEIGH_EAM, small_system

pytestmark = pytest.mark.gpu


def _run(s, cu_funcfl):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    eng = Engine(0)
    try:
        eng.load_system(s)
        eng.set_pair(p)
        eng.set_neighbor(CUTNEIGH_EAM, every=1, delay=0, check=True)
        eng.set_coupling(True)
        eng.build()
        eng.force()
        info = eng.cluster_info()
        nb = eng.run(10, dt=0.002)
        x, v, f = eng.atoms()
        nx, nv, nf = eng.elements()
    finally:
        eng.close()
    return info, nb, (x, v, f, nx, nv, nf)


def test_datafile_system_runs_identically(tmp_path, cu_funcfl):
    s = small_system(seed=21, L=10, nae=5, core=4.0, temperature=300.0)
    path = str(tmp_path / "small.cac")
    D.write_data(path, D.from_system(s))
    s2 = D.to_system(D.read_data(path), a0=s.a0)
    s2.v = s.v.copy()           # Atoms lines carry no velocities (atom_vec_atomic.cpp:49-71); nodes do
    print(f"[datafile] atoms={s.n_atoms} elements={s.n_elements}", flush=True)
    i1, nb1, out1 = _run(s, cu_funcfl)
    i2, nb2, out2 = _run(s2, cu_funcfl)
    assert i1 == i2 and i1[1], (i1, i2)
    assert nb1 == nb2
    for a, b in zip(out1, out2):
        assert np.array_equal(a, b)
