"""The on-disk formats through the device path.

* pair_style eam/alloy from a setfl file: the device forces of a coupled
  nae=5 system equal the oracle's, computed with the oracle's own tables of
  the same setfl (array2spline restated, pair_eam.cpp:527-580), to 1e-10;
  the same functions as a funcfl (resampled onto nr-1 knots,
  pair_eam.cpp:311-523) give the same energy to 1e-9.
* Engine.write_dump_atom / write_dump_element write the device state: the
  files equal the reference's printf formats (dump_atom.cpp:158-182,
  dump_element.cpp:156-172) applied to the downloaded arrays.
"""
import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM, PairEAMAlloy
from cac_code_amd.potential import johnson_cu_funcfl, setfl_from_funcfl

from gpu_common import CUTNEIGH_EAM, oracle_state, rel_err, small_system
from test_io_native import _atom_dump_text, _element_dump_text

pytestmark = pytest.mark.gpu


def engine(s, pair):
    e = Engine(0)
    e.load_system(s)
    e.set_pair(pair)
    e.set_neighbor(CUTNEIGH_EAM, skin=1.0)
    e.set_coupling(True)
    e.build()
    e.force(eflag=1)
    return e


def test_setfl_forces_match_oracle(orc):
    s = small_system(seed=21)
    sf = setfl_from_funcfl(johnson_cu_funcfl())
    p = PairEAMAlloy()
    p.settings([])
    p.coeff(["*", "*", "Cu.setfl", "Cu"], setfl=sf)
    print(f"[setfl] atoms={s.n_atoms} elements={s.n_elements}", flush=True)
    with engine(s, p) as e:
        _, _, f = e.atoms()
        _, _, nf = e.elements()
        en_setfl = e.energy_virial()[0].sum()
    ref = oracle_state(orc, s, CUTNEIGH_EAM, tables=orc.eam_tables_setfl(sf), coupled=True)
    assert rel_err(f, ref["atoms"]["f"]) < 1e-10
    assert rel_err(nf, ref["elems"]["nodef"]) < 1e-10
    q = PairEAM()
    q.settings([])
    q.coeff(["1", "1", "Cu_johnson.eam"], funcfl=johnson_cu_funcfl())
    with engine(s, q) as e:
        en_funcfl = e.energy_virial()[0].sum()
    print(f"  energy setfl {en_setfl:.12f}, funcfl {en_funcfl:.12f}", flush=True)
    assert abs(en_setfl - en_funcfl) < 1e-9 * abs(en_funcfl)


def test_engine_dumps(tmp_path, cu_funcfl):
    s = small_system(seed=22)
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    with engine(s, p) as e:
        e.run(3, dt=0.001)
        pa, pe = str(tmp_path / "atoms.dump"), str(tmp_path / "elements.dump")
        assert e.write_dump_atom(pa, 3) == s.n_atoms + 8 * s.n_elements
        assert e.write_dump_element(pe, 3) == s.n_elements
        x, v, f = e.atoms()
        nx, nv, nf = e.elements()
    n, ne = s.n_atoms, s.n_elements
    want = _atom_dump_text(3, s.box_lo, s.box_hi, x, f, v, nx, nf, nv, np.arange(1, n + 1), np.ones(n, int),
                           np.arange(1, 8 * ne + 1))
    assert open(pa).read() == want
    assert open(pe).read() == _element_dump_text(3, nx, nv, nf)
