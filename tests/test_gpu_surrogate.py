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
    # operands.  Hidden activations are rounded to BF16 between layers: where
    # the two sums fall on either side of a BF16 rounding boundary (a few in
    # 1e4 activations) the rounded activations differ by one BF16 ulp
    # (2^-8 relative), which sets the RMS bound; the max bound is a few ulps.
    assert np.max(np.abs(y - yr)) <= 2e-3 * np.max(np.abs(yr)), np.max(np.abs(y - yr))
    assert np.sqrt(np.mean((y - yr) ** 2)) <= 5e-4 * np.sqrt(np.mean(yr ** 2))


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


def _features_numpy(s, X0, X, box, nbr):
    """The 279 features restated in numpy: the affine map by least squares,
    corner nodes found by nearest-node search around each corner."""
    prd0 = s.box_hi - s.box_lo
    ne = len(X)
    c0, c = X0.mean(axis=1), X.mean(axis=1)
    r0 = X0 - c0[:, None]
    G0 = np.einsum("eki,ekj->eij", r0, r0)
    Fe = np.einsum("eki,ekj->eij", X - c[:, None], r0) @ np.linalg.inv(G0)
    dev = lambda v, v0: v - np.einsum("eij,e...j->e...i", Fe, v0)
    mi = lambda v, p: v - p * np.rint(v / p)
    own = dev(X - c[:, None], r0).reshape(ne, 24)
    cent = dev(mi(c[nbr] - c[:, None, :], box), mi(c0[nbr] - c0[:, None, :], prd0)).reshape(ne, 78)
    offs = [(a, b, cc) for a in (-1, 0, 1) for b in (-1, 0, 1) for cc in (-1, 0, 1) if (a, b, cc) != (0, 0, 0)]
    axes = np.stack([X0[0, 1] - X0[0, 0], X0[0, 3] - X0[0, 0], X0[0, 4] - X0[0, 0]])
    corner = []
    for k in range(8):
        toward = np.sign(r0[0, k] @ np.linalg.inv(axes))
        ms = [m for m, o in enumerate(offs) if all(oi == 0 or oi == ti for oi, ti in zip(o, toward))]
        assert len(ms) == 7
        for m in ms:
            j = nbr[:, m]
            v = mi(X0[j] - X0[np.arange(ne), k][:, None, :], prd0)
            dist = np.linalg.norm(v, axis=2)
            n = np.argmin(dist, axis=1)
            # a block touching the corner has a node within a few lattice
            # vectors of it (one primitive vector per shifted axis)
            assert np.all(dist.min(axis=1) < 3.615 * np.sqrt(3) + 1e-6)
            corner.append(dev(mi(X[j, n] - c, box)[:, None, :], mi(X0[j, n] - c0, prd0)[:, None, :])[:, 0])
    corner = np.stack(corner, axis=1).reshape(ne, 168)
    strain = (Fe - np.eye(3)).reshape(ne, 9)
    return np.concatenate([own, cent, corner, strain], axis=1)


def test_features_match_numpy_restatement(pair):
    s = build_cac_system(33, 11, 0.0)
    X0 = s.nodex.copy()
    nbr = S.mesh_neighbors(s)
    prd = s.box_hi - s.box_lo
    eps = np.array([0.01, -0.015, 0.005])
    rng = np.random.default_rng(8)
    with Engine(0) as eng:
        eng.load_system(s)
        eng.set_pair(pair)
        eng.surrogate_setup(X0, nbr, np.ones(NIN), np.ones(24))
        f, cls = eng.surrogate_features()
        assert np.max(np.abs(f)) < 1e-9 and np.all(cls == 0)
        # homogeneous strain: zero deviations, strain features diag(eps)
        box = prd * (1 + eps)
        eng.set_box(s.box_lo, s.box_lo + box)
        eng.upload_elements(X0 * (1 + eps), None, s.nae, s.mass_e, s.intp_ijk, s.intp_w)
        f, cls = eng.surrogate_features()
        assert np.max(np.abs(f[:, :270])) < 1e-9
        assert np.allclose(f[:, 270:], np.diag(eps).reshape(1, 9), atol=1e-12)
        assert np.all(cls == 0)                     # |eps| < 2%
        # strain plus independent nodal displacements
        X = X0 * (1 + eps) + 0.02 * rng.standard_normal(X0.shape)
        eng.upload_elements(X, None, s.nae, s.mass_e, s.intp_ijk, s.intp_w)
        f, _ = eng.surrogate_features()
    ref = _features_numpy(s, X0, X, box, nbr)
    assert np.max(np.abs(f - ref)) < 1e-6, np.max(np.abs(f - ref))
    assert np.max(np.abs(ref[:, 102:270])) > 1e-2  # the corner deviations are exercised
    with Engine(0) as eng:
        eng.load_system(s)
        eng.set_pair(pair)
        eng.surrogate_setup(X0, nbr, np.ones(NIN), np.ones(24), strain_limit=0.012)
        eng.set_box(s.box_lo, s.box_lo + box)
        eng.upload_elements(X0 * (1 + eps), None, s.nae, s.mass_e, s.intp_ijk, s.intp_w)
        _, cls = eng.surrogate_features()
    assert np.all(cls == 2)                     # 1.5% > 1.2% limit: strained


@pytest.fixture(scope="module")
def small_model(pair):
    ds = S.generate_dataset(pair, 24000, L=33, seed=21, log=print)
    model = S.train(ds, epochs=40, batch=512, lr=1e-3, log=print)
    return ds, model


def test_trained_surrogate_gain(pair, small_model):
    """On held-out plain-path nodal forces, the least-squares part explains
    all but a few percent and the trained network removes a good part of
    what is left."""
    ds, model = small_model
    held = S.generate_dataset(pair, 4000, L=33, seed=99)
    err = S.rel_rms(S.predict(model, held.x), held.t)
    err_fit = S.rel_rms(S.predict(model, held.x, structured_only=True), held.t)
    print(f"[surrogate] held-out rel RMS nodal-force error {err:.5f} (least-squares part alone {err_fit:.5f})")
    assert err_fit < 0.05
    assert err < 0.8 * err_fit


def test_fallback_elements_keep_plain_forces(pair, small_model):
    """With the surrogate on, interface elements use the plain path on the
    subsets F, R, R2 and get the plain nodal forces (to rounding); the others get
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
    # the plain forces: per-point kernels for the fallback subsets against the
    # cluster kernels of the full plain call, equal to rounding (independently
    # fused multiply-adds)
    assert np.max(np.abs(surr[fb] - plain[fb])) <= 1e-12 * np.max(np.abs(plain[fb]))
    assert np.max(np.abs(fsurr - fplain)) <= 1e-12 * max(np.max(np.abs(fplain)), 1e-30)
    pred = S.predict(model, f[~fb])
    assert np.allclose(surr[~fb].reshape(-1, 24), pred, rtol=2e-2, atol=2e-3 * np.abs(pred).max())


@pytest.fixture(scope="module")
def config_model(small_model):
    ds, _ = small_model
    return S.train_config(ds, epochs=40, batch=512, lr=1e-3, log=print)


def test_configured_network(pair, small_model, config_model):
    """The configured 102-input network (own nodes and neighbor centroids,
    MLP only) runs on the same device path: fallback elements keep the
    plain forces, surrogate elements get the network's forces (BF16
    inference on the device as in predict).  On held-out thermal data it
    misses by far more than the structured surrogate: a node's force
    depends on the nodes of the elements meeting at it, which the centroids
    do not carry."""
    _, model = small_model
    held = S.generate_dataset(pair, 4000, L=33, seed=99)
    err = S.rel_rms(S.predict(model, held.x), held.t)
    err_cfg = S.rel_rms(S.predict(config_model, held.x), held.t)
    print(f"[surrogate] held-out rel RMS: configured 102-input network {err_cfg:.5f}, structured {err:.5f}")
    assert err < 0.5 * err_cfg
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
        s.X0 = X0
        info = S.attach(eng, config_model, s)
        eng.force()
        _, _, surr = eng.elements()
        f, cls = eng.surrogate_features()
    fb = cls != 0
    assert fb.any() and (~fb).any() and (~fb).sum() == info["surrogate"]
    assert np.max(np.abs(surr[fb] - plain[fb])) <= 1e-12 * np.max(np.abs(plain[fb]))
    pred = S.predict(config_model, f[~fb])
    assert np.allclose(surr[~fb].reshape(-1, 24), pred, rtol=2e-2, atol=2e-3 * np.abs(pred).max())
