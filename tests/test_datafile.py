"""CAC data file reader and writer (read_data.cpp:277-437, element.cpp:195-540,
1086-1166, atom_vec_atomic.cpp:49-71): a hand-written file in the reference's
layout, its error texts, and write -> read round trips bit for bit."""
import numpy as np
import pytest

from cac_code_amd import datafile as D
from cac_code_amd.lattice import build_cac_system, lj_c0_system, vertex_clusters

HAND = """# hand-written CAC data file

2 atoms   # two atoms
1 elements
1 atom types
1 element types
0.0 20.0 xlo xhi
0.0 20.0 ylo yhi
-5.0 15.0 zlo zhi

Masses

1 63.546

NAAE

1 3 3 3

Numint

1 8

Intp

1 1 0 0 0 3.375
1 2 2 0 0 3.375
1 3 2 2 0 3.375
1 4 0 2 0 3.375
1 5 0 0 2 3.375
1 6 2 0 2 3.375
1 7 2 2 2 3.375
1 8 0 2 2 3.375

Elements

7 1 5.0 5.0 5.0 0 0 1

Nodes

7 1 4.0 4.0 4.0 11
7 2 6.0 4.0 4.0 12
7 3 6.0 6.0 4.0 13
7 4 4.0 6.0 4.0 14
7 5 4.0 4.0 6.0 15
7 6 6.0 4.0 6.0 16
7 7 6.0 6.0 6.0 17
7 8 4.0 6.0 6.0 18

Atoms

3 1 12.0 12.0 12.0
4 1 14.5 12.0 12.0
"""


def _write(tmp_path, text, name="data.cac"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_hand_written_file(tmp_path):
    d = D.read_data(_write(tmp_path, HAND))
    assert d.comment.startswith("# hand-written")
    np.testing.assert_array_equal(d.box_lo, [0.0, 0.0, -5.0])
    np.testing.assert_array_equal(d.box_hi, [20.0, 20.0, 15.0])
    assert d.masses == {1: 63.546} and d.nae == {1: (3, 3, 3)} and d.numint == {1: 8}
    assert d.intp[1].shape == (8, 4) and np.all(d.intp[1][:, 3] == 3.375)
    np.testing.assert_array_equal(d.intp[1][2, :3], [2, 2, 0])
    assert list(d.elem_tag) == [7] and list(d.elem_image[0]) == [0, 0, 1]
    np.testing.assert_array_equal(d.nodex[0, 6], [6.0, 6.0, 6.0])
    assert np.all(d.nodev == 0.0) and list(d.nodetag[0]) == list(range(11, 19))
    assert list(d.atom_tag) == [3, 4] and d.atom_image is None
    np.testing.assert_array_equal(d.x[1], [14.5, 12.0, 12.0])
    s = D.to_system(d)
    assert s.n_atoms == 2 and s.n_elements == 1 and s.nae == (3, 3, 3)
    assert s.intp_ijk.dtype == np.int32 and s.intp_w.sum() == 27.0


@pytest.mark.parametrize("edit,msg", [
    (("Masses", "Bonds"), "Unknown identifier in data file: Bonds"),
    (("1 3 3 3", "1 3 1 3"), "Invalid naae values"),
    (("1 3 3 3", "2 3 3 3"), "Invalid type for nae set"),
    (("7 1 5.0 5.0 5.0 0 0 1", "7 2 5.0 5.0 5.0 0 0 1"), "Invalid atom type in Elements section"),
    (("7 1 5.0 5.0 5.0 0 0 1", "7 1 5.0 5.0 5.0 0 1"), "Incorrect element format"),
    (("7 1 4.0 4.0 4.0 11", "7 1 4.0 4.0 11"), "Incorrect node format"),
    (("4 1 14.5 12.0 12.0", "4 1 14.5 12.0"), "Incorrect atom format"),
])
def test_errors_follow_the_reference(tmp_path, edit, msg):
    with pytest.raises(D.DataFileError, match=msg):
        D.read_data(_write(tmp_path, HAND.replace(*edit, 1)))


def test_nodes_before_elements_and_truncation(tmp_path):
    i, j = HAND.index("Elements\n"), HAND.index("Nodes\n")
    k = HAND.index("Atoms\n")
    swapped = HAND[:i] + HAND[j:k] + HAND[i:j] + HAND[k:]
    with pytest.raises(D.DataFileError, match="Nodes section before Elements"):
        D.read_data(_write(tmp_path, swapped))
    with pytest.raises(D.DataFileError, match="Unexpected end of data file"):
        D.read_data(_write(tmp_path, HAND[:HAND.index("7 8 4.0")]))


def test_center_must_match_node_centroid(tmp_path):
    d = D.read_data(_write(tmp_path, HAND.replace("7 1 5.0 5.0 5.0", "7 1 5.0 5.0 5.2")))
    with pytest.raises(D.DataFileError, match="center"):
        D.to_system(d)
    D.to_system(D.read_data(_write(tmp_path, HAND.replace("7 1 5.0 5.0 5.0", "7 1 5.0 5.0 5.05"))))


@pytest.mark.parametrize("make", [lambda: build_cac_system(10, 5, 4.0, temperature=300.0, seed=3),
                                  lambda: lj_c0_system()], ids=["eam_small", "lj_c0"])
def test_round_trip_bit_exact(tmp_path, make):
    s = make()
    p = str(tmp_path / "rt.cac")
    D.write_data(p, D.from_system(s))
    s2 = D.to_system(D.read_data(p), a0=s.a0)
    for a, b in ((s.box_lo, s2.box_lo), (s.box_hi, s2.box_hi), (s.x, s2.x), (s.nodex, s2.nodex),
                 (s.nodev, s2.nodev), (s.intp_ijk, s2.intp_ijk), (s.intp_w, s2.intp_w)):
        assert np.array_equal(a, b)
    assert s2.mass == s.mass and tuple(s2.nae) == tuple(s.nae)
    # node tags carry the mesh: the same integration-point clusters
    c1, c2 = vertex_clusters(s), vertex_clusters(s2)
    assert np.array_equal(c1[0], c2[0]) and np.array_equal(c1[1], c2[1])
    # and a second write is the same text
    p2 = str(tmp_path / "rt2.cac")
    D.write_data(p2, D.from_system(s2))
    assert open(p).read() == open(p2).read()
