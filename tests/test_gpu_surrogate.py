"""GPU tests of the learned element surrogate: the BF16 MFMA MLP against the
CPU oracle (forward and gradients), device training, surrogate features in
closed form, bit-exact fallback elements, and what the trained network gains
over the path without it."""
import time

import numpy as np
import pytest

from oracle import mlp as M
from cac_code_amd import Engine, PairEAM, johnson_cu_funcfl
from cac_code_amd.lattice import build_cac_system
from cac_code_amd import surrogate as S

from gpu_common import jitter

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair():
    p = PairEAM()
    p.settings([])
    p.coeff(["1", "1", "Cu.eam"], funcfl=johnson_cu_funcfl())
    return p


NIN = S.NIN


def _rand_net(eng, seed):
    eng.mlp_create(NIN, 256, 3, 24, seed)
    return [(w.astype(np.float64), b.astype(np.float64)) for w, b in eng.mlp_get_weights()]


@pytest.mark.parametrize("n", [1, 37, 300])
def test_mlp_forward_matches_oracle(n):
    print(f"[mlp forward] n={n}", flush=True)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, NIN)).astype(np.float32)
    with Engine(0) as eng:
        w = _rand_net(eng, 7)
        y = eng.mlp_forward(x)
    yr, _ = M.forward(w, x, "bf16")
    # FP32 accumulation on the device vs float64 here, on identical BF16
    # operands: agreement to FP32 accumulation error
    assert np.max(np.abs(y - yr)) <= 2e-3 * np.max(np.abs(yr)), np.max(np.abs(y - yr))
    assert np.sqrt(np.mean((y - yr) ** 2)) <= 2e-4 * np.sqrt(np.mean(yr ** 2))


def test_mlp_gradients_match_oracle():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(96, NIN)).astype(np.float32)
    t = rng.normal(size=(96, 24)).astype(np.float32)
    with Engine(0) as eng:
        w = _rand_net(eng, 11)
        loss, grads = eng.mlp_grads(x, t)
    lr, gr = M.loss_grads(w, x, t, "bf16")
    assert abs(loss - lr) <= 1e-4 * lr
    for l, ((gw, gb), (rw, rb)) in enumerate(zip(grads, gr)):
        ew = np.linalg.norm(gw - rw) / np.linalg.norm(rw)
        eb = np.linalg.norm(gb - rb) / np.linalg.norm(rb)
        assert ew < 5e-3 and eb < 5e-3, (l, ew, eb)


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(NIN, 24)) / 10
    x = rng.normal(size=(8192, NIN)).astype(np.float32)
    t = np.tanh(x @ A).astype(np.float32)
    with Engine(0) as eng:
        eng.mlp_create(NIN, 256, 3, 24, 1)
        eng.mlp_set_dataset(x, t)
        l0 = eng.mlp_train(1, 256, 1e-3, 0)
        l1 = eng.mlp_train(20, 256, 1e-3, 1)
        # one host-batch step keeps working next to the resident set
        l2 = eng.mlp_train_step(x[:512], t[:512], 1e-4)
    assert l1 < 0.1 * l0, (l0, l1)
    assert np.isfinite(l2)


def test_features_closed_form_under_strain(pair):
    s = build_cac_system(33, 11, 0.0)
    X0 = s.nodex.copy()
    nbr = S.mesh_neighbors(s)
    eps = np.array([0.01, -0.015, 0.005])
    with Engine(0) as eng:
        eng.load_system(s)
        eng.set_pair(pair)
        eng.surrogate_setup(X0, nbr, np.ones(NIN), np.ones(24))
        f, cls = eng.surrogate_features()
        assert np.max(np.abs(f)) < 1e-9 and np.all(cls == 0)
        eng.set_box(s.box_lo, s.box_lo + (s.box_hi - s.box_lo) * (1 + eps))
        eng.upload_elements(X0 * (1 + eps), None, s.nae, s.mass_e, s.intp_ijk, s.intp_w)
        f, cls = eng.surrogate_features()
    r0 = X0 - X0.mean(axis=1, keepdims=True)
    assert np.allclose(f[:, :24], (r0 * eps).reshape(-1, 24), atol=1e-5)
    cen = X0.mean(axis=1)
    prd = s.box_hi - s.box_lo
    d = cen[nbr] - cen[:, None, :]
    d -= prd * np.rint(d / prd)
    assert np.allclose(f[:, 24:102], (d * eps).reshape(-1, 78), atol=1e-5)
    # corner inputs: for own corner k and each of the 7 neighbor blocks that
    # touch it (offset components 0 or toward the corner), the neighbor's
    # node nearest to corner k, as an offset from the own centroid
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]
    expect = []
    for k in range(8):
        toward = np.sign(r0[0, k] @ np.linalg.inv(_block_axes(s)))
        ms = [m for m, o in enumerate(offs) if all(oi == 0 or oi == ti for oi, ti in zip(o, toward))]
        assert len(ms) == 7
        for m in ms:
            j = nbr[:, m]
            v = X0[j] - X0[np.arange(len(j)), k][:, None, :]
            v -= prd * np.rint(v / prd)
            dist = np.linalg.norm(v, axis=2)
            node = np.argmin(dist, axis=1)
            # a block touching the corner has a node within a few lattice
            # vectors of it (one primitive vector per shifted axis, <= a*sqrt(3))
            assert np.all(dist.min(axis=1) < 3.615 * np.sqrt(3) + 1e-6)
            c = X0[j, node] - cen
            c -= prd * np.rint(c / prd)
            expect.append(c * eps)
    expect = np.stack(expect, axis=1).reshape(-1, 168)
    assert np.allclose(f[:, 102:], expect, atol=1e-5)
    assert np.all(cls == 0)                     # |eps| < 2%
    with Engine(0) as eng:
        eng.load_system(s)
        eng.set_pair(pair)
        eng.surrogate_setup(X0, nbr, np.ones(NIN), np.ones(24), strain_limit=0.012)
        eng.set_box(s.box_lo, s.box_lo + (s.box_hi - s.box_lo) * (1 + eps))
        eng.upload_elements(X0 * (1 + eps), None, s.nae, s.mass_e, s.intp_ijk, s.intp_w)
        _, cls = eng.surrogate_features()
    assert np.all(cls == 2)                     # 1.5% > 1.2% limit: strained


def _block_axes(s):
    """Rows: the element's three edge vectors (node 0 to nodes 1, 3, 4)."""
    X = s.nodex[0]
    return np.stack([X[1] - X[0], X[3] - X[0], X[4] - X[0]])


@pytest.fixture(scope="module")
def small_model(pair):
    ds = S.generate_dataset(pair, 12000, L=33, seed=21, log=print)
    model = S.train(ds, epochs=40, batch=512, lr=1e-3, log=print)
    return ds, model


def test_trained_surrogate_gain(pair, small_model):
    """On held-out plain-path nodal forces, the linear part alone explains
    most of the force and the trained network removes much of what is left."""
    ds, model = small_model
    held = S.generate_dataset(pair, 3000, L=33, seed=99)
    err = S.rel_rms(S.predict(model, held.x), held.t)
    err_lin = S.rel_rms(S.predict(model, held.x, linear_only=True), held.t)
    print(f"[surrogate] held-out rel RMS nodal-force error {err:.4f} (linear part alone {err_lin:.4f}, zero 1.0)")
    assert err_lin < 0.2
    assert err < 0.7 * err_lin


def test_fallback_elements_keep_plain_forces(pair, small_model):
    """With the surrogate on, interface elements use the plain path on the
    subsets F, R, R2 and get exactly the plain nodal forces; the others get
    the surrogate's (linear part on the device in FP32, MLP in BF16)."""
    _, model = small_model
    s = build_cac_system(33, 11, 8.0, temperature=300.0, seed=3)
    X0 = s.nodex.copy()
    s = jitter(s, 0.04, 4)
    with Engine(0) as eng:
        eng.load_system(s)
        eng.set_pair(pair)
        eng.set_neighbor(5.95, skin=1.0)
        eng.build()
        eng.force()
        _, _, plain = eng.elements()
        _, _, fplain = eng.atoms()
        s.X0 = X0
        info = S.attach(eng, model, s)
        eng.force()
        _, _, surr = eng.elements()
        _, _, fsurr = eng.atoms()
        f, cls = eng.surrogate_features()
    print(f"[surrogate] {info}")
    fb = cls != 0
    assert fb.sum() == info["interface"] + info["strained"] and (~fb).sum() == info["surrogate"]
    assert fb.any() and (~fb).any()
    assert np.array_equal(surr[fb], plain[fb])
    assert np.array_equal(fsurr, fplain)
    pred = S.predict(model, f[~fb])
    assert np.allclose(surr[~fb].reshape(-1, 24), pred, rtol=2e-2, atol=2e-3 * np.abs(pred).max())
