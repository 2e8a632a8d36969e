"""GPU centro-symmetry (cac_centro_atoms) against the CPU oracle and the fcc
closed forms (Output::centro_atom, output.cpp:352-441).

Bar: within 1e-10 relative of the oracle on the same positions (absolute
floor 1e-12 A^2 for values that vanish on a perfect lattice); the closed
forms as in tests/test_oracle_centro.py.
"""
import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM
from cac_code_amd._lib import CacError
from cac_code_amd.lattice import build_cac_system

from gpu_common import CUTNEIGH_EAM, oracle_state, small_system

pytestmark = pytest.mark.gpu

A = 3.615
RTOL, ATOL = 1e-10, 1e-12


def eam_pair(cu_funcfl):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    return p


def engine_for(s, cu_funcfl, coupled=False):
    eng = Engine(0)
    eng.load_system(s)
    eng.set_pair(eam_pair(cu_funcfl))
    eng.set_neighbor(CUTNEIGH_EAM, every=1, delay=0, check=True)
    eng.set_coupling(coupled)
    eng.build()
    return eng


def oracle_centro(orc, s, coupled, nnn=12):
    ref = oracle_state(orc, s, CUTNEIGH_EAM, coupled=coupled)
    return orc.centro_atoms(s.x, s.box_lo, s.box_hi, ref["anum"], ref["anb"], CUTNEIGH_EAM, nnn,
                            ref["es"] if coupled else None)


SYSTEMS = {
    "small": lambda: small_system(seed=7, L=10, nae=5, core=4.0),
    "medium": lambda: small_system(seed=3, L=20, nae=5, core=8.0, sigma=0.08),
    "finest": lambda: small_system(seed=5, L=8, nae=2, core=3.0, sigma=0.03),
}


@pytest.mark.parametrize("coupled", [False, True], ids=["plain", "coupled"])
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_centro_parity(name, coupled, orc, cu_funcfl):
    s = SYSTEMS[name]()
    print(f"[centro] {name} coupled={coupled}: atoms={s.n_atoms} elements={s.n_elements}", flush=True)
    ref = oracle_centro(orc, s, coupled)
    eng = engine_for(s, cu_funcfl, coupled)
    try:
        c = eng.centro()
    finally:
        eng.close()
    np.testing.assert_allclose(c, ref, rtol=RTOL, atol=ATOL)
    assert np.count_nonzero(ref > 0.01) > 0


def test_centro_closed_forms_on_device(cu_funcfl):
    s = build_cac_system(6, 3, 6.0)                     # all atoms, perfect crystal
    i = 211
    u = np.array([0.03, -0.04, 0.0])
    s.x = s.x.copy()
    s.x[i] += u
    print(f"[centro] closed forms: atoms={s.n_atoms}", flush=True)
    eng = engine_for(s, cu_funcfl)
    try:
        c = eng.centro()
    finally:
        eng.close()
    prd = s.box_hi - s.box_lo
    d = s.x - s.x[i]
    r = np.linalg.norm(d - prd * np.round(d / prd), axis=1)
    first = np.nonzero((r > 0.1) & (r < 0.8 * A))[0]
    assert len(first) == 12
    u2 = float(u @ u)
    np.testing.assert_allclose(c[i], 24.0 * u2, rtol=1e-9)
    np.testing.assert_allclose(c[first], u2, rtol=1e-9)
    rest = np.ones(s.n_atoms, bool)
    rest[first] = False
    rest[i] = False
    assert np.all(np.abs(c[rest]) < 1e-20)


def test_centro_vacancy_and_dump(tmp_path, cu_funcfl):
    s = build_cac_system(6, 3, 6.0)
    v = 100
    xv = s.x[v].copy()
    keep = np.arange(s.n_atoms) != v
    s.x = np.ascontiguousarray(s.x[keep])
    s.v = np.ascontiguousarray(s.v[keep])
    print(f"[centro] vacancy: atoms={s.n_atoms}", flush=True)
    eng = engine_for(s, cu_funcfl)
    try:
        c = eng.centro()
        path = tmp_path / "defects.xyz"
        count = eng.write_dump_xyz(str(path))
    finally:
        eng.close()
    prd = s.box_hi - s.box_lo
    d = s.x - xv
    r = np.linalg.norm(d - prd * np.round(d / prd), axis=1)
    first = r < 0.8 * A
    assert np.count_nonzero(first) == 12
    np.testing.assert_allclose(c[first], 0.5 * A * A, rtol=1e-12)
    assert np.all(np.abs(c[~first]) < 1e-20)
    lines = path.read_text().splitlines()
    assert count == 12 and lines[0] == "12" and lines[1] == "XYZ" and len(lines) == 14
    assert all(ln.startswith("Cu ") and len(ln.split()) == 5 for ln in lines[2:])


def test_centro_after_run_uses_current_positions(orc, cu_funcfl):
    # after steps the lists are Verlet lists of an older build; the 12
    # nearest are inside them, so the value equals a fresh oracle build's
    s = small_system(seed=9, L=10, nae=5, core=4.0, temperature=600.0)
    print(f"[centro] after run: atoms={s.n_atoms} elements={s.n_elements}", flush=True)
    eng = engine_for(s, cu_funcfl, coupled=True)
    try:
        eng.force()
        eng.run(30, dt=0.002)
        c = eng.centro()
        x, _, _ = eng.atoms()
        nodex, _, _ = eng.elements()
    finally:
        eng.close()
    # the device wraps positions at rebuilds only; the oracle's cell search
    # wants them in the box (a whole-period shift moves the value by rounding)
    orc.pbc_atoms(x, s.box_lo, s.box_hi)
    xc = orc.element_centers(nodex)
    orc.pbc_elements(xc, nodex, s.box_lo, s.box_hi)
    s.x, s.nodex = x, nodex
    ref = oracle_centro(orc, s, True)
    np.testing.assert_allclose(c, ref, rtol=RTOL, atol=ATOL)


def test_centro_bcc_nnn8(orc, cu_funcfl):
    a, L = 2.87, 7                                      # box >= 3 neighbor cutoffs
    base = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    g = np.array([[i, j, k] for k in range(L) for j in range(L) for i in range(L)], float)
    x = ((g[:, None, :] + base[None]) * a).reshape(-1, 3) + 0.1
    rng = np.random.default_rng(4)
    x += 0.04 * rng.standard_normal(x.shape)
    lo, hi = np.zeros(3), np.full(3, L * a)
    print(f"[centro] bcc nnn=8: atoms={len(x)}", flush=True)
    eng = Engine(0)
    try:
        eng.set_box(lo, hi)
        eng.upload_atoms(x, None, 55.85)
        eng.set_pair(eam_pair(cu_funcfl))
        eng.set_neighbor(CUTNEIGH_EAM, every=1, delay=0, check=True)
        eng.build()
        c = eng.centro(nnn=8)
    finally:
        eng.close()
    num, nb = orc.atom_neighbors(x, lo, hi, CUTNEIGH_EAM, 200)
    ref = orc.centro_atoms(x, lo, hi, num, nb, CUTNEIGH_EAM, 8)
    np.testing.assert_allclose(c, ref, rtol=RTOL, atol=ATOL)


def test_centro_argument_errors(cu_funcfl):
    s = build_cac_system(6, 3, 6.0)
    eng = Engine(0)
    try:
        eng.load_system(s)
        eng.set_pair(eam_pair(cu_funcfl))
        eng.set_neighbor(CUTNEIGH_EAM)
        with pytest.raises(CacError):
            eng.centro()                                 # lists not built
        eng.build()
        with pytest.raises(CacError):
            eng.centro(nnn=10)
    finally:
        eng.close()


def test_centro_of_interpolated_sites(orc, cu_funcfl):
    # the reference's executed path (verlet.cpp:91-94): atoms replaced by the
    # interpolated element sites, list built over them, their centro-symmetry
    s = small_system(seed=13, L=10, nae=5, core=4.0, sigma=0.03)
    print(f"[centro] interpolated sites: elements={s.n_elements} nae={list(s.nae)}", flush=True)
    eng = engine_for(s, cu_funcfl)
    try:
        xs = eng.interpolated_sites()
        xs2, c = eng.centro_of_sites()
    finally:
        eng.close()
    spa = orc.site_shapes(s.nae)                        # [nsite][8], inter_ele order
    ref = np.zeros((s.n_elements, spa.shape[0], 3))
    for k in range(8):                                  # node order from 0.0, unfused
        ref = ref + spa[None, :, k, None] * s.nodex[:, None, k, :]
    np.testing.assert_allclose(xs, ref.reshape(-1, 3), rtol=1e-14, atol=1e-12)
    assert np.array_equal(xs, xs2)
    w = np.ascontiguousarray(xs.copy())
    orc.pbc_atoms(w, s.box_lo, s.box_hi)
    num, nb = orc.atom_neighbors(w, s.box_lo, s.box_hi, CUTNEIGH_EAM, 200)
    cref = orc.centro_atoms(w, s.box_lo, s.box_hi, num, nb, CUTNEIGH_EAM, 12)
    np.testing.assert_allclose(c, cref, rtol=RTOL, atol=ATOL)
    # the atomistic core is gone: its surface is in the dump window
    assert np.count_nonzero((cref > 1) & (cref < 10)) > 0
