"""LAMMPS setfl files (pair_style eam/alloy; the reference reads funcfl
only, pair_eam.cpp:197-268) and their spline tables.

* A hand-written two-element file: every array lands where the format puts
  it (elements in file order, F then rho per element, then r*phi for the
  pairs (0,0), (1,0), (1,1)), values split over lines of any length.
* write_setfl -> read_setfl is exact.
* The native table builder (cac_eam_tables_from_setfl) equals the oracle's
  array2spline restatement (pair_eam.cpp:527-580) over the same arrays bit
  for bit, and reproduces a cubic: the 5-point knot slope is exact for
  polynomials of degree <= 4, so with f(m) = cubic in the knot index the
  interior slopes are f'(m) and the spline equals f between interior knots
  (a wrong stencil weight or a p^2/p^3 coefficient swap fails).
* pair_coeff errors of eam/alloy.
"""
import numpy as np
import pytest

from cac_code_amd import PairEAMAlloy, PairError, eam_tables_setfl, read_setfl, write_setfl
from cac_code_amd.potential import Setfl, johnson_cu_funcfl, setfl_from_funcfl

HAND = """line one
line two
line three
2 Cu Ni
5 0.1 5 0.5 2.0
29 63.546 3.615 fcc
0 -1 -2
-3 -4
1 2 3 4 5
28 58.69 3.52 fcc
0 -2 -4 -6 -8
10 20
30 40 50
7 7 7 7 7
8 8 8 8 8
9 9 9 9 9
"""


def test_hand_written_setfl(tmp_path):
    p = tmp_path / "two.setfl"
    p.write_text(HAND)
    sf = read_setfl(str(p))
    assert sf.comments == ["line one", "line two", "line three"]
    assert sf.elements == ["Cu", "Ni"] and (sf.nrho, sf.drho, sf.nr, sf.dr, sf.cut) == (5, 0.1, 5, 0.5, 2.0)
    assert sf.number == [29, 28] and sf.mass == [63.546, 58.69] and sf.lattice == ["fcc", "fcc"]
    np.testing.assert_array_equal(sf.frho[0], [0, -1, -2, -3, -4])
    np.testing.assert_array_equal(sf.rhor[0], [1, 2, 3, 4, 5])
    np.testing.assert_array_equal(sf.frho[1], [0, -2, -4, -6, -8])
    np.testing.assert_array_equal(sf.rhor[1], [10, 20, 30, 40, 50])
    assert sorted(sf.z2r) == [(0, 0), (1, 0), (1, 1)]
    np.testing.assert_array_equal(sf.z2r[(0, 0)], [7] * 5)
    np.testing.assert_array_equal(sf.z2r[(1, 0)], [8] * 5)
    np.testing.assert_array_equal(sf.z2r[(1, 1)], [9] * 5)
    assert sf.element_index("Ni") == 1
    with pytest.raises(ValueError, match="No matching element"):
        sf.element_index("Al")
    q = tmp_path / "short.setfl"
    q.write_text(HAND[:HAND.index("9 9 9")])
    with pytest.raises(ValueError, match="premature end"):
        read_setfl(str(q))


def test_round_trip_exact(tmp_path):
    sf = setfl_from_funcfl(johnson_cu_funcfl())
    p = str(tmp_path / "cu.setfl")
    write_setfl(p, sf)
    sf2 = read_setfl(p)
    assert sf2.elements == ["Cu"] and sf2.nrho == sf.nrho and sf2.drho == sf.drho and sf2.cut == sf.cut
    for a, b in zip(sf.frho + sf.rhor + [sf.z2r[(0, 0)]], sf2.frho + sf2.rhor + [sf2.z2r[(0, 0)]]):
        assert np.array_equal(a, b)


def test_tables_equal_oracle(orc):
    sf = setfl_from_funcfl(johnson_cu_funcfl())
    t, o = eam_tables_setfl(sf), orc.eam_tables_setfl(sf)
    for k in ("frho_spline", "rhor_spline", "z2r_spline"):
        assert np.array_equal(t[k], o[k]), k
    # the knot values are the file's values, unresampled
    assert np.array_equal(t["rhor_spline"][1:, 6], sf.rhor[0])
    assert np.array_equal(t["frho_spline"][1:, 6], sf.frho[0])


def test_tables_reproduce_a_cubic():
    n, h = 40, 0.25
    m = np.arange(1, n + 1, dtype=np.float64)
    f = lambda q: 0.5 * q ** 3 - 2.0 * q ** 2 + 3.0 * q - 7.0          # noqa: E731
    df = lambda q: 1.5 * q ** 2 - 4.0 * q + 3.0                         # noqa: E731
    sf = Setfl(["", "", ""], ["X"], n, h, n, h, 1.0, [1], [1.0], [0.0], [""], [f(m)], [f(m)], {(0, 0): f(m)})
    sp = eam_tables_setfl(sf)["rhor_spline"]
    interior = np.arange(3, n - 1)
    np.testing.assert_allclose(sp[interior, 5], df(interior.astype(float)), rtol=1e-13)
    for q in (5, 17, 33):
        for p in (0.0, 0.25, 0.5, 0.9):
            c = sp[q]
            val = ((c[3] * p + c[4]) * p + c[5]) * p + c[6]               # pair_eam.cpp:655
            dval = ((c[0] * p + c[1]) * p + c[2])                         # derivative / h, :720
            assert abs(val - f(q + p)) < 1e-10 * abs(f(q + p))
            assert abs(dval - df(q + p) / h) < 1e-10 * abs(df(q + p) / h)


def test_pair_coeff_errors(tmp_path):
    sf = setfl_from_funcfl(johnson_cu_funcfl())
    p = PairEAMAlloy()
    p.settings([])
    with pytest.raises(PairError, match="Incorrect args"):
        p.coeff(["1", "1", "f", "Cu"], setfl=sf)
    with pytest.raises(PairError, match="Incorrect args"):
        p.coeff(["*", "*", "f"], setfl=sf)
    with pytest.raises(ValueError, match="No matching element"):
        p.coeff(["*", "*", "f", "Ni"], setfl=sf)
    with pytest.raises(PairError, match="All pair coeffs"):
        p.init_one(1, 1)
    p.coeff(["*", "*", "f", "Cu"], setfl=sf)
    assert p.mass == sf.mass[0] and p.init_one(1, 1) == sf.cut
