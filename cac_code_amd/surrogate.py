"""Learned element surrogate: data from the plain integration-point path,
fitting, and deployment on the force path.

Inputs (279, computed on the device by k_surrogate.hip): deviations from the
element's best-fit affine map of its 8 nodes (24), of its 26 neighbor
centroids (78) and of the 7 neighbor nodes that meet each corner (168), then
the strain F - I of that map (9).  Outputs: the element's 24 nodal forces.
Prediction of force c = 3k + d (normalized units):

    b_c + A_c . x + sum_{i<=j} Q_c[ij] z_i z_j + MLP(x)_c

with z the 33 local inputs of corner k (own node k, its 7 corner neighbors,
the strain).  A, Q, b come from FP64 least squares; the MLP
(279 -> 256 -> 256 -> 256 -> 24, SiLU, BF16 weights and activations, FP32
accumulation, k_mlp.hip) is trained with the device Adam on what they leave.

Why these inputs and this structure (measured on the oracle, 300 K nodal
noise, |strain| <= 2%): a node's force depends on the nodes of the 7
neighbor elements that meet it, so the 24 + 78 element/centroid inputs
alone leave ~70% of the nodal force unexplained under thermal motion; with
the corner nodes a linear map leaves ~8%, nearly all of it the anharmonic
(quadratic) response, which the corner-local quadratic terms capture down
to ~0.8%; the MLP takes the remainder to ~0.2% held-out.

The configured network (``train_config``): the first 102 inputs only (own
nodes and neighbor centroids), MLP 102 -> 256 -> 256 -> 256 -> 24 (SiLU,
BF16, FP32 accumulation) trained with the same Adam on the forces
themselves, no least-squares part.  It is deployed by the same device path
(an MLP of 102 inputs switches the feature and output kernels) and reported
beside the structured one.

Training data: periodic all-element fcc boxes of the same element type under
random normal strains up to `strain_max` (an orthogonal periodic box cannot
carry shear), periodic long-wave displacement modes (local strain variation,
including shear), and per-node Gaussian noise of the equipartition amplitude
sqrt(kB T / k_node) at `temperature`, with k_node the diagonal stiffness of
one node measured on the plain path; targets are the FP64 nodal forces of the
plain path on the same device.  Elements outside the envelope (interface
elements, strain above the limit) fall back to the plain path at run time.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .engine import SURR_NFEAT, SURR_NFEAT_CONFIG, SURR_NOUT, SURR_NQUAD, Engine, PairEAM
from .lattice import CACSystem, build_cac_system

NIN, NHID, NLAYERS, NOUT = SURR_NFEAT, 256, 3, SURR_NOUT
NIN_CONFIG = SURR_NFEAT_CONFIG
NLOC = 33
_IU = np.triu_indices(NLOC)
assert len(_IU[0]) == SURR_NQUAD


def local_features(k: int) -> np.ndarray:
    """Feature indices of corner k's 33 local inputs: own node k (3), the 7
    neighbor nodes meeting corner k (21), the strain (9)."""
    return np.concatenate([3 * k + np.arange(3), 102 + 21 * k + np.arange(21), 270 + np.arange(9)])


def quad_terms(xn: np.ndarray, k: int) -> np.ndarray:
    """[n][561] products z_i z_j (i <= j, row-major) of corner k's inputs."""
    z = xn[:, local_features(k)].astype(np.float64)
    return z[:, _IU[0]] * z[:, _IU[1]]
BOLTZ_METAL = 8.617343e-5


def mesh_neighbors(s: CACSystem) -> np.ndarray:
    """[ne][26] element indices of the 26 neighboring blocks (di, dj, dk in
    {-1,0,1}^3 minus 0, i outermost), -1 where the block is not an element.
    Blocks are matched through their centroids wrapped into the box (doubled
    a/2 integer units, as lattice.fcc_box_blocks enumerates them)."""
    if s.blocks is None:
        raise ValueError("system has no block indices (not built by the mesher)")
    nae, L = s.nae[0], s.cells
    lim = 4 * L

    def key(B):
        cx = (2 * nae * (B[..., 0] + B[..., 1]) + 2 * (nae - 1) + 1) % lim
        cy = (2 * nae * (B[..., 0] + B[..., 2]) + 2 * (nae - 1) + 1) % lim
        cz = (2 * nae * (B[..., 1] + B[..., 2]) + 2 * (nae - 1) + 1) % lim
        return (cx.astype(np.int64) * lim + cy) * lim + cz

    k0 = key(s.blocks)
    order = np.argsort(k0)
    ks = k0[order]
    offs = np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
                     if (a, b, c) != (0, 0, 0)], dtype=np.int64)
    kn = key(s.blocks[:, None, :] + offs[None, :, :])
    pos = np.clip(np.searchsorted(ks, kn), 0, len(ks) - 1)
    nbr = np.where(ks[pos] == kn, order[pos], -1).astype(np.int32)
    return nbr


def _periodic_modes(rng, box, nmodes, strain):
    """Displacement field sum_m A_m sin(k_m . x + phi_m), |k| = 2pi/L along
    random lattice directions, amplitude giving local strain ~ strain."""
    L = box
    modes = []
    for _ in range(nmodes):
        n = rng.integers(-1, 2, 3)
        if not n.any():
            n[rng.integers(0, 3)] = 1
        k = 2 * np.pi * n / L
        amp = rng.normal(size=3)
        amp *= strain * rng.uniform(0, 1) / (np.linalg.norm(amp) * np.linalg.norm(k))
        modes.append((k, amp, rng.uniform(0, 2 * np.pi)))
    return modes


def _apply(x, modes):
    u = np.zeros_like(x)
    for k, amp, phi in modes:
        u += np.sin(x @ k + phi)[..., None] * amp
    return u


@dataclass
class Dataset:
    x: np.ndarray           # [n][279] raw features
    t: np.ndarray           # [n][24] nodal forces (eV/A)
    meta: dict = field(default_factory=dict)


def _plain_nodef(eng, base, box_hi, nodex, cutneigh):
    eng.set_box(base.box_lo, box_hi)
    eng.upload_elements(nodex, None, base.nae, base.mass_e, base.intp_ijk, base.intp_w)
    eng.upload_atoms(np.zeros((0, 3)), None, base.mass)
    eng.set_neighbor(cutneigh, skin=1.0)
    eng.build()
    eng.force()
    return eng.elements()[2]


def node_stiffness(eng, base, cutneigh: float = 5.95, h: float = 0.01) -> float:
    """Mean diagonal stiffness (eV/A^2) of one node of the unstrained box:
    central differences of its own force under +-h moves along x, y, z."""
    X0 = base.nodex
    k = []
    for d in range(3):
        fs = []
        for sgn in (1.0, -1.0):
            X = X0.copy()
            X[0, 0, d] += sgn * h
            fs.append(_plain_nodef(eng, base, base.box_hi, X, cutneigh)[0, 0, d])
        k.append(-(fs[0] - fs[1]) / (2 * h))
    return float(np.mean(k))


def generate_dataset(pair: PairEAM, n_samples: int, L: int = 33, nae: int = 11, seed: int = 1234,
                     strain_max: float = 0.02, mode_strain: float = 0.004, temperature: float = 300.0,
                     cutneigh: float = 5.95, device: int = 0, log=None) -> Dataset:
    """Plain-path (FP64, GPU) nodal forces of strained, thermally displaced
    all-element boxes; features from the device feature kernel.  Each
    configuration draws its strain magnitude uniformly in [0, strain_max] and
    its noise amplitude uniformly in [0.25, 1.5] x the equipartition one."""
    rng = np.random.Generator(np.random.PCG64(seed))
    base = build_cac_system(L, nae, 0.0)
    ne = base.n_elements
    nbr = mesh_neighbors(base)
    X0 = base.nodex.copy()
    prd = base.box_hi - base.box_lo
    xs, ts = [], []
    t0 = time.perf_counter()
    ncfg = 0
    with Engine(device) as eng:
        eng.load_system(base)
        eng.set_pair(pair)
        knode = node_stiffness(eng, base, cutneigh)
        sigma = float(np.sqrt(BOLTZ_METAL * temperature / knode))
        ones = np.ones(NIN, dtype=np.float32)
        eng.surrogate_setup(X0, nbr, ones, np.ones(NOUT, dtype=np.float32), strain_limit=10.0)
        total = 0
        while total < n_samples:
            eps = rng.uniform(-1, 1, 3)
            eps *= strain_max * rng.uniform(0, 1) / max(np.abs(eps).max(), 1e-12)
            F = 1.0 + eps
            modes = _periodic_modes(rng, prd, 3, mode_strain)
            noise = sigma * rng.uniform(0.25, 1.5)
            nodex = X0 * F + _apply(X0, modes) * F + noise * rng.standard_normal(X0.shape)
            nodef = _plain_nodef(eng, base, base.box_lo + prd * F, nodex, cutneigh)
            f, _ = eng.surrogate_features()
            xs.append(f)
            ts.append(nodef.reshape(ne, 24).astype(np.float32))
            total += ne
            ncfg += 1
    x = np.concatenate(xs)[:n_samples]
    t = np.concatenate(ts)[:n_samples]
    if log:
        log(f"[surrogate] dataset: {x.shape[0]} samples from {ncfg} configurations in {time.perf_counter() - t0:.1f}s"
            f" (k_node {knode:.1f} eV/A^2, sigma {sigma:.4f} A at {temperature} K)")
    return Dataset(x, t, dict(L=L, nae=nae, seed=seed, strain_max=strain_max, mode_strain=mode_strain,
                              temperature=temperature, k_node=knode, sigma=sigma, configurations=ncfg))


def generate_dataset_trajectory(pair: PairEAM, system_fn, n_samples: int, seed: int = 1234,
                                strain_max: float = 0.02, first: int = 100, every: int = 100,
                                cutneigh: float = 5.95, coupled: bool = True, device: int = 0,
                                log=None) -> Dataset:
    """Plain-path (FP64, GPU) nodal forces of the deployment system itself:
    snapshots of its NVE trajectory (velocities drawn with `seed`, so the
    thermal state differs from any run evaluated with another seed), taken
    after `first` steps and then every `every` steps, each labelled under a
    random normal strain of magnitude uniform in [0, strain_max] (box, atoms
    and nodes mapped affinely).  Only elements the surrogate would cover
    (class 0: no atom in their lists, all mesh neighbors present, strain
    within the envelope) become samples; features are taken against the
    unstrained reference nodes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = system_fn(seed)
    X0 = s.nodex.copy()
    nbr = mesh_neighbors(s)
    lo = s.box_lo.copy()
    prd = s.box_hi - s.box_lo
    ne = s.n_elements
    xs, ts = [], []
    total = 0
    nsnap = 0
    t0 = time.perf_counter()
    with Engine(device) as run, Engine(device) as lab:
        for e in (run, lab):
            e.load_system(s)
            e.set_pair(pair)
            e.set_neighbor(cutneigh, skin=1.0, every=1, delay=10)
            e.set_coupling(coupled)
        run.build()
        run.force()
        ones = np.ones(NIN, dtype=np.float32)
        lab.surrogate_setup(X0, nbr, ones, np.ones(NOUT, dtype=np.float32), strain_limit=10.0)
        steps = 0
        while total < n_samples:
            k = first if steps == 0 else every
            run.run(k)
            steps += k
            x, _, _ = run.atoms()
            nodex, _, _ = run.elements()
            eps = rng.uniform(-1, 1, 3)
            eps *= strain_max * rng.uniform(0, 1) / max(np.abs(eps).max(), 1e-12)
            F = 1.0 + eps
            lab.set_box(lo, lo + prd * F)
            lab.upload_atoms(lo + (x - lo) * F, None, s.mass)
            lab.upload_elements(lo + (nodex - lo) * F, None, s.nae, s.mass_e, s.intp_ijk, s.intp_w)
            lab.build()
            lab.force()
            f, cls = lab.surrogate_features()
            keep = cls == 0
            xs.append(f[keep])
            ts.append(lab.elements()[2].reshape(ne, 24)[keep].astype(np.float32))
            total += int(keep.sum())
            nsnap += 1
    x = np.concatenate(xs)[:n_samples]
    t = np.concatenate(ts)[:n_samples]
    if log:
        log(f"[surrogate] trajectory dataset: {x.shape[0]} samples from {nsnap} snapshots (steps {first}.."
            f"{steps}, seed {seed}) in {time.perf_counter() - t0:.1f}s")
    return Dataset(x, t, dict(kind="trajectory", seed=seed, strain_max=strain_max, first=first, every=every,
                              snapshots=nsnap, steps=steps, coupled=coupled))


@dataclass
class Model:
    weights: list           # MLP layers [(W [out][in], b [out])], FP32 masters
    in_scale: np.ndarray    # [nin]
    out_scale: np.ndarray   # [24]
    fit_A: np.ndarray       # [24][279] least-squares linear part, normalized units (None: MLP only)
    fit_Q: np.ndarray       # [24][561] least-squares corner-local quadratic part
    fit_b: np.ndarray       # [24]
    history: list = field(default_factory=list)
    nin: int = NIN          # 279, or 102 for the configured network


def fit_structured(xn: np.ndarray, tn: np.ndarray, ridge: float = 1e-10):
    """FP64 least squares of the linear + corner-local quadratic part, one
    normal-equation solve per corner (its three outputs share the design
    [x, quadratic terms of corner k, 1])."""
    n = len(xn)
    A = np.zeros((NOUT, NIN))
    Q = np.zeros((NOUT, SURR_NQUAD))
    b = np.zeros(NOUT)
    for k in range(8):
        m = NIN + SURR_NQUAD + 1
        G = np.zeros((m, m))
        R = np.zeros((m, 3))
        for i in range(0, n, 32768):
            xb = xn[i:i + 32768].astype(np.float64)
            Z = np.concatenate([xb, quad_terms(xb, k), np.ones((len(xb), 1))], axis=1)
            G += Z.T @ Z
            R += Z.T @ tn[i:i + 32768, 3 * k:3 * k + 3].astype(np.float64)
        G[np.diag_indices(m)] += ridge * np.trace(G) / m
        W = np.linalg.solve(G, R)
        A[3 * k:3 * k + 3] = W[:NIN].T
        Q[3 * k:3 * k + 3] = W[NIN:NIN + SURR_NQUAD].T
        b[3 * k:3 * k + 3] = W[-1]
    return A.astype(np.float32), Q.astype(np.float32), b.astype(np.float32)


def structured_predict(A, Q, b, xn: np.ndarray) -> np.ndarray:
    """Normalized prediction of the least-squares part (FP64 on the host)."""
    xn = xn.astype(np.float64)
    y = xn @ A.astype(np.float64).T + b
    for k in range(8):
        y[:, 3 * k:3 * k + 3] += quad_terms(xn, k) @ Q[3 * k:3 * k + 3].astype(np.float64).T
    return y


def train(ds: Dataset, epochs: int = 60, batch: int = 1024, lr: float = 1e-3, seed: int = 0, device: int = 0,
          log=None, lr_decay_to: float = 0.05) -> Model:
    """Least squares of the structured part, then Adam on the device MLP over
    its residual (scaled to unit RMS per output while training, the scale
    folded into the last layer after); the rate falls geometrically from lr
    to lr*lr_decay_to over the epochs."""
    in_scale = np.sqrt(np.mean(ds.x.astype(np.float64) ** 2, axis=0)).astype(np.float32)
    in_scale = np.maximum(in_scale, 1e-6 * max(float(in_scale.max()), 1e-12)).astype(np.float32)
    out_scale = np.sqrt(np.mean(ds.t.astype(np.float64) ** 2, axis=0)).astype(np.float32)
    out_scale = np.maximum(out_scale, 1e-6 * max(float(out_scale.max()), 1e-12)).astype(np.float32)
    xn = (ds.x * (1.0 / in_scale)).astype(np.float32)
    tn = ds.t / out_scale
    t0 = time.perf_counter()
    A, Q, b = fit_structured(xn, tn)
    res = tn - structured_predict(A, Q, b, xn)
    r_scale = np.sqrt(np.mean(res ** 2, axis=0))
    r_scale = np.maximum(r_scale, 1e-6 * max(float(r_scale.max()), 1e-12))
    if log:
        log(f"[surrogate] least-squares part: train rel RMS residual {float(np.sqrt(np.mean(res ** 2))):.5f}"
            f" ({time.perf_counter() - t0:.1f}s)")
    tn = (res / r_scale).astype(np.float32)
    hist = []
    t0 = time.perf_counter()
    with Engine(device) as eng:
        eng.mlp_create(NIN, NHID, NLAYERS, NOUT, seed)
        eng.mlp_set_dataset(xn, tn)
        chunk = max(1, epochs // 10)
        done = 0
        while done < epochs:
            e = min(chunk, epochs - done)
            rate = lr * (lr_decay_to ** (done / max(epochs - 1, 1)))
            loss = eng.mlp_train(e, batch, rate, seed + done)
            done += e
            hist.append((done, rate, loss))
            if log:
                log(f"[surrogate] epoch {done}/{epochs} lr {rate:.2e} loss {loss:.3e}")
        w = eng.mlp_get_weights()
        W, bl = w[-1]
        w[-1] = ((W * r_scale[:, None]).astype(np.float32), (bl * r_scale).astype(np.float32))
        # a perfect lattice (every deviation zero) carries no nodal force:
        # shift the constant term so the whole prediction vanishes at x = 0
        eng.mlp_set_weights(w)
        y0 = eng.mlp_forward(np.zeros((1, NIN), dtype=np.float32))[0].astype(np.float64)
    b = (b.astype(np.float64) - (b.astype(np.float64) + y0)).astype(np.float32)
    if log:
        log(f"[surrogate] MLP trained in {time.perf_counter() - t0:.1f}s (constant term set for zero force at "
            f"zero deviation, shift {float(np.sqrt(np.mean(y0 ** 2))):.2e})")
    return Model(w, in_scale, out_scale, A, Q, b, hist)


def train_config(ds: Dataset, epochs: int = 60, batch: int = 1024, lr: float = 1e-3, seed: int = 0, device: int = 0,
                 log=None, lr_decay_to: float = 0.05) -> Model:
    """The configured network: MLP 102 -> 256 -> 256 -> 256 -> 24 on the
    first 102 features, trained with the device Adam on the normalized forces
    (same schedule as train); the output bias is then shifted so a perfect
    lattice (zero deviation) gets zero force."""
    x = np.ascontiguousarray(ds.x[:, :NIN_CONFIG])
    in_scale = np.sqrt(np.mean(x.astype(np.float64) ** 2, axis=0)).astype(np.float32)
    in_scale = np.maximum(in_scale, 1e-6 * max(float(in_scale.max()), 1e-12)).astype(np.float32)
    out_scale = np.sqrt(np.mean(ds.t.astype(np.float64) ** 2, axis=0)).astype(np.float32)
    out_scale = np.maximum(out_scale, 1e-6 * max(float(out_scale.max()), 1e-12)).astype(np.float32)
    xn = (x * (1.0 / in_scale)).astype(np.float32)
    tn = (ds.t / out_scale).astype(np.float32)
    hist = []
    t0 = time.perf_counter()
    with Engine(device) as eng:
        eng.mlp_create(NIN_CONFIG, NHID, NLAYERS, NOUT, seed)
        eng.mlp_set_dataset(xn, tn)
        chunk = max(1, epochs // 10)
        done = 0
        while done < epochs:
            e = min(chunk, epochs - done)
            rate = lr * (lr_decay_to ** (done / max(epochs - 1, 1)))
            loss = eng.mlp_train(e, batch, rate, seed + done)
            done += e
            hist.append((done, rate, loss))
            if log:
                log(f"[surrogate/config] epoch {done}/{epochs} lr {rate:.2e} loss {loss:.3e}")
        w = eng.mlp_get_weights()
        y0 = eng.mlp_forward(np.zeros((1, NIN_CONFIG), dtype=np.float32))[0]
        W, bl = w[-1]
        w[-1] = (W, (bl - y0).astype(np.float32))
    if log:
        log(f"[surrogate/config] MLP trained in {time.perf_counter() - t0:.1f}s")
    return Model(w, in_scale, out_scale, None, None, None, hist, nin=NIN_CONFIG)


def predict(model: Model, x: np.ndarray, device: int = 0, structured_only: bool = False) -> np.ndarray:
    """Nodal forces [n][24] for raw features x [n][279]: the least-squares
    part on the host (FP64), the MLP on the device (the configured network:
    the MLP of the first 102 features alone)."""
    xn = (x[:, :model.nin] * (1.0 / model.in_scale)).astype(np.float32)
    y = (structured_predict(model.fit_A, model.fit_Q, model.fit_b, xn) if model.fit_A is not None
         else np.zeros((len(x), NOUT)))
    if not structured_only:
        with Engine(device) as eng:
            eng.mlp_create(model.nin, NHID, NLAYERS, NOUT, 0)
            eng.mlp_set_weights(model.weights)
            y = y + eng.mlp_forward(xn)
    return y * model.out_scale


def rel_rms(pred, ref):
    pred = np.asarray(pred, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.sqrt(np.mean((pred - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))


def attach(eng: Engine, model: Model, s: CACSystem, strain_limit: float = 0.02):
    """Load the model into eng's context and switch its element forces to
    the surrogate (fallback elements keep the plain path)."""
    eng.mlp_create(model.nin, NHID, NLAYERS, NOUT, 0)
    eng.mlp_set_weights(model.weights)
    in_scale = np.ones(NIN, dtype=np.float32)             # the device scales all 279 features
    in_scale[:model.nin] = model.in_scale
    eng.surrogate_setup(s.nodex if getattr(s, "X0", None) is None else s.X0, mesh_neighbors(s), in_scale,
                        model.out_scale, strain_limit)
    if model.fit_A is not None:
        eng.surrogate_set_fit(model.fit_A, model.fit_Q, model.fit_b)
    eng.surrogate_enable(True)
    return eng.surrogate_info()
