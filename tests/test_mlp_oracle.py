"""CPU pins for the MLP oracle (oracle/mlp.py) and the surrogate mesh
adjacency: the oracle's gradients equal finite differences of its own loss
(fp64 mode), BF16 rounding is round-to-nearest-even, and the block
adjacency is the periodic 26-neighborhood."""
import numpy as np

from oracle import mlp as M
from cac_code_amd.lattice import build_cac_system
from cac_code_amd.surrogate import mesh_neighbors


def _net(rng, dims):
    return [(rng.normal(size=(dims[l + 1], dims[l])) / np.sqrt(dims[l]), rng.normal(size=dims[l + 1]) * 0.1)
            for l in range(len(dims) - 1)]


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [7, 16, 16, 5]
    w = _net(rng, dims)
    x = rng.normal(size=(11, 7))
    t = rng.normal(size=(11, 5))
    loss, grads = M.loss_grads(w, x, t, mode="fp64")
    h = 1e-6
    for l in range(len(w)):
        W, b = w[l]
        for (i, j) in [(0, 0), (W.shape[0] - 1, W.shape[1] - 1), (1, 2)]:
            wp = [(a.copy(), c.copy()) for a, c in w]
            wm = [(a.copy(), c.copy()) for a, c in w]
            wp[l][0][i, j] += h
            wm[l][0][i, j] -= h
            fd = (M.loss_grads(wp, x, t, "fp64")[0] - M.loss_grads(wm, x, t, "fp64")[0]) / (2 * h)
            assert abs(fd - grads[l][0][i, j]) < 1e-6 * max(1.0, abs(fd)), (l, i, j)
        bp = [(a.copy(), c.copy()) for a, c in w]
        bm = [(a.copy(), c.copy()) for a, c in w]
        bp[l][1][1] += h
        bm[l][1][1] -= h
        fd = (M.loss_grads(bp, x, t, "fp64")[0] - M.loss_grads(bm, x, t, "fp64")[0]) / (2 * h)
        assert abs(fd - grads[l][1][1]) < 1e-6 * max(1.0, abs(fd))


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.0e38], dtype=np.float32)
    r = M.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0  # exactly representable? 2^-8 is below bf16 spacing
    assert r[3] == 1.0                         # tie to even
    assert r[2] == 1.0 + 2 ** -7               # above the tie rounds up
    assert r[4] == -2.5
    # bf16 keeps 8 significant bits: relative error below 2^-8
    x = np.random.default_rng(1).normal(size=1000).astype(np.float32)
    assert np.max(np.abs(M.bf16(x) - x) / np.abs(x)) <= 2 ** -8


def test_mesh_neighbors_periodic_26():
    s = build_cac_system(33, 11, 0.0)          # 108 elements, all interior
    nbr = mesh_neighbors(s)
    assert nbr.shape == (108, 26) and np.all(nbr >= 0)
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]
    opp = [offs.index((-a, -b, -c)) for a, b, c in offs]
    for e in range(108):
        for m in range(26):
            assert nbr[nbr[e, m], opp[m]] == e
    # neighbor centroids sit one block step away
    cen = s.nodex.mean(axis=1)
    prd = s.box_hi - s.box_lo
    d = cen[nbr[0]] - cen[0]
    d -= prd * np.rint(d / prd)
    P = 0.5 * 3.615 * np.array([[1.0, 1.0, 0.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0]])
    expect = 11 * np.array(offs, dtype=float) @ P
    assert np.allclose(d, expect, atol=1e-9)


def test_mesh_neighbors_mark_interface():
    s = build_cac_system(33, 11, 8.0)          # a core of atoms
    nbr = mesh_neighbors(s)
    assert (nbr < 0).any(axis=1).sum() > 0
    assert (nbr >= 0).all(axis=1).sum() > 0
