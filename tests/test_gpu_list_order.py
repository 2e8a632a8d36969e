"""Order of the atom lists (k_atom_neigh): every row is sorted by the key
(lattice-rounded displacement k = rint((x_j' - x_i) / spacing), components
clamped to [-128, 127], ordered by min(|k|^2, 255), then k_z, k_y, k_x; then
the entry j | img << 27).  The key is recomputed on the host from the
downloaded positions and each row must be strictly increasing in it.  On an
fcc lattice at a/2 the rows therefore run shell by shell: the test also
checks that the first three shells (the entries within the EAM cutoff)
fill the front of every row of a perfect lattice.

Rows whose rounded displacements are all distinct (an fcc lattice at a/2) are
ranked through a presence bitmap of the codes; rows where two neighbors share
a code (a coarse spacing, or atoms near a dislocation core) are ranked by a
sweep over the row.  Both must give the key order, so the test covers a
jittered lattice and an edge-dislocation dipole at a/2, a coarse spacing, and
spacing 0 (rows sorted by entry), and asserts that the neighbor sets agree
across all of them.
"""
import numpy as np
import pytest

from cac_code_amd import Engine, PairEAM
from cac_code_amd.lattice import build_cac_system, edge_dipole

from gpu_common import CUTNEIGH_EAM, jitter, sorted_rows

pytestmark = pytest.mark.gpu


def host_keys(x, lo, hi, num, nb, spacing):
    """The device's list key of every stored entry, from the same positions
    and the same operations (x + s*prd for an image, rint of d * (1/spacing)
    clamped to [-128, 127], |k|^2 capped at 255 in front)."""
    prd = hi - lo
    keys = []
    qinv = 1.0 / spacing if spacing > 0.0 else 0.0
    for i in range(len(num)):
        e = (nb[i, :num[i]] & np.uint64(0xFFFFFFFF)).astype(np.uint64)
        j = (e & np.uint64(0x7FFFFFF)).astype(np.int64)
        img = (e >> np.uint64(27)).astype(np.int64)
        if qinv == 0.0:
            keys.append(e)
            continue
        s = np.stack([img % 3 - 1, (img // 3) % 3 - 1, img // 9 - 1], axis=1)
        xj = x[j].copy()
        for d in range(3):
            m = s[:, d] != 0
            xj[m, d] = xj[m, d] + s[m, d].astype(np.float64) * prd[d]
        k = np.clip(np.rint((xj - x[i]) * qinv), -128, 127).astype(np.int64)
        k2 = np.minimum((k * k).sum(axis=1), 255)
        q = k + 128
        hiw = (k2 << 24) | (q[:, 2] << 16) | (q[:, 1] << 8) | q[:, 0]
        keys.append((hiw.astype(np.uint64) << np.uint64(32)) | e)
    return keys


def lists(s, cu_funcfl, spacing):
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu_johnson.eam"], funcfl=cu_funcfl)
    with Engine(0) as eng:
        eng.load_system(s)
        eng.set_pair(p)
        eng.set_neighbor(CUTNEIGH_EAM, skin=1.0, every=1, delay=0, check=True)
        eng.set_list_order(spacing)
        eng.build()
        num, nb = eng.atom_neighbors()
        x, _, _ = eng.atoms()
    return x, num, nb


def dipole_system():
    s = build_cac_system(20, 5, 12.0, temperature=300.0, seed=31)
    edge_dipole(s)
    return s


SYSTEMS = {
    "jittered": lambda: jitter(build_cac_system(16, 4, 1.0e6, temperature=300.0, seed=29), 0.05, 30),
    "dipole": dipole_system,
}


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_rows_sorted_by_key(name, cu_funcfl):
    s = SYSTEMS[name]()
    half = 0.5 * s.a0
    ref = None
    for spacing in (half, 2.0 * s.a0, 0.0):
        print(f"[list order] {name}: atoms={s.n_atoms} spacing={spacing}", flush=True)
        x, num, nb = lists(s, cu_funcfl, spacing)
        keys = host_keys(x, s.box_lo, s.box_hi, num, nb, spacing)
        bad = [i for i, k in enumerate(keys) if k.size > 1 and not np.all(k[1:] > k[:-1])]
        assert not bad, f"spacing {spacing}: rows not in key order, e.g. atoms {bad[:10]}"
        if spacing == half and name == "jittered":
            # fcc shells at a/2: |k|^2 = 2, 4, 6 (12 + 6 + 24 entries, within
            # the 4.95 A cutoff), then 8 and 10 (12 + 24, the skin)
            k2 = [(k >> np.uint64(56)).astype(np.int64) for k in keys]
            full = [r for r in k2 if r.size == 78]
            assert len(full) > len(k2) // 2
            assert all(np.all(r[:42] <= 6) and np.all(r[42:] > 6) for r in full)
        rows = sorted_rows(num, nb)
        if ref is None:
            ref = (num.copy(), rows)
            if spacing == half:
                # both rankings run: rows with distinct rounded displacements
                # and rows where some share one
                clash = sum(1 for k in keys if np.unique(k >> np.uint64(32)).size < k.size)
                print(f"[list order] {name}: {clash} of {len(keys)} rows share a rounded displacement", flush=True)
                if name == "dipole":
                    assert 0 < clash < len(keys)
        else:
            assert np.array_equal(num, ref[0])
            assert all(np.array_equal(a, b) for a, b in zip(rows, ref[1]))
