"""Synthetic CAC systems: an fcc lattice in a periodic cubic box, meshed into
8-node rhombohedral elements with an atomistic core.

Geometry (units: Angstrom).  The fcc primitive vectors
    p1 = a/2 (1,1,0),  p2 = a/2 (1,0,1),  p3 = a/2 (0,1,1)
span the lattice.  An element is a block of nae^3 sites
    n = nae*B + (ii, jj, kk),  0 <= ii, jj, kk < nae,
whose parent coordinate xi runs along p1 (ii), eta along p2 (jj), zeta along
p3 (kk), matching the site order and node order of the reference
(element.cpp:1394-1423, element.cpp:1375-1382):
    node 0 (-,-,-) 1 (+,-,-) 2 (+,+,-) 3 (-,+,-)
    node 4 (-,-,+) 5 (+,-,+) 6 (+,+,+) 7 (-,+,+).
Blocks tile the lattice without sharing sites, and a cubic box of L cells is
periodic for the block tiling when nae divides L (L*a*x = L*(p1+p2-p3) etc.),
giving 4*L^3/nae^3 blocks.  A block is kept as an element iff none of its
sites lies in the central core cube; the sites of dropped blocks are atoms.

Integration points are the 8 corner sites with weight nae^3/8 each (the
reference maps Intp lattice indices to parent coordinates, element.cpp:
1170-1191, and needs one at every node, element.cpp:1296-1319).

Velocities: Gaussian from numpy's PCG64 stream with a fixed seed, zero total
momentum, rescaled to the exact temperature over atoms and nodes (node mass
m_e/8, fix_nve.cpp:124).  The reference has no velocity create
(velocity.cpp:119-297), so this is the build's own definition.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

BOLTZ_METAL = 8.617343e-5       # update.cpp:108
MVV2E_METAL = 1.0364269e-4      # update.cpp:110
FTM2V_METAL = 1.0 / 1.0364269e-4  # update.cpp:111
NKTV2P_METAL = 1.6021765e6      # update.cpp:113

# node k -> (ii, jj, kk) corner flags (0 = low, 1 = high)
NODE_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                         [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)


@dataclass
class CACSystem:
    box_lo: np.ndarray
    box_hi: np.ndarray
    mass: float                 # atom mass (g/mol)
    x: np.ndarray               # [n,3]
    v: np.ndarray               # [n,3]
    nae: tuple                  # sites per element edge
    nodex: np.ndarray           # [ne,8,3]
    nodev: np.ndarray           # [ne,8,3]
    intp_ijk: np.ndarray        # [nint,3] lattice indices of integration points
    intp_w: np.ndarray          # [nint]
    blocks: np.ndarray = field(default=None)  # [ne,3] block indices (mesh adjacency)
    a0: float = 3.615
    cells: int = 0
    node_tags: np.ndarray = field(default=None)  # [ne,8] mesh vertex ids (data file), when blocks are unknown

    @property
    def n_atoms(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.nodex.shape[0])

    @property
    def mass_e(self) -> float:
        """Element mass nae0*nae1*nae2*m (element.cpp:1104)."""
        return float(self.nae[0] * self.nae[1] * self.nae[2]) * self.mass

    @property
    def n_integrated(self) -> int:
        return self.n_atoms + 8 * self.n_elements

    @property
    def n_represented(self) -> int:
        return self.n_atoms + int(np.prod(self.nae)) * self.n_elements


def corner_quadrature(nae):
    """8 corner integration points, weight nae^3/8 each, in node order."""
    n = np.asarray(nae, dtype=np.int64)
    ijk = NODE_CORNERS * (n - 1)[None, :]
    w = np.full(8, float(np.prod(n)) / 8.0)
    return ijk.astype(np.int32), w


def _prim(a):
    return 0.5 * a * np.array([[1.0, 1.0, 0.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0]])


def fcc_box_blocks(L: int, nae: int):
    """Distinct blocks of the periodic tiling: block indices B whose centroid
    lies in [0, L*a)^3.  Works in units of a/2 with integer arithmetic, the
    lattice shifted by a/4 along each axis so no centroid sits on a face."""
    if L % nae:
        raise ValueError("box edge L must be a multiple of nae for a periodic tiling")
    M = L // nae
    h2 = nae - 1                       # 2*h, h = (nae-1)/2 the centroid offset
    rng = np.arange(-2 * M - 1, 3 * M + 2)
    B1, B2, B3 = np.meshgrid(rng, rng, rng, indexing="ij")
    B1, B2, B3 = B1.ravel(), B2.ravel(), B3.ravel()
    # 2*centroid in units of a/2, plus 1 for the a/4 shift (doubled units)
    cx = 2 * nae * (B1 + B2) + 2 * h2 + 1
    cy = 2 * nae * (B1 + B3) + 2 * h2 + 1
    cz = 2 * nae * (B2 + B3) + 2 * h2 + 1
    lim = 4 * L
    keep = (cx >= 0) & (cx < lim) & (cy >= 0) & (cy < lim) & (cz >= 0) & (cz < lim)
    B = np.stack([B1[keep], B2[keep], B3[keep]], axis=1)
    if B.shape[0] != 4 * M ** 3:
        raise RuntimeError("block enumeration did not find 4*M^3 blocks")
    return B


def build_cac_system(L: int, nae: int, core_cells: float, a: float = 3.615, mass: float = 63.546,
                     temperature: float = 0.0, seed: int = 4928459, core_center=None,
                     min_atoms: int = 0) -> CACSystem:
    """fcc box of L^3 cubic cells; blocks with any site inside the central
    core cube of edge core_cells*a become atoms, the rest become elements.
    With min_atoms, further blocks become atoms, nearest first (periodic
    distance of the block centroid to the core center, then block order),
    until the atoms number at least min_atoms: the cube's atom count grows
    in steps of whole block layers."""
    P = _prim(a)
    origin = np.full(3, a / 4.0)
    box_lo = np.zeros(3)
    box_hi = np.full(3, L * a)
    B = fcc_box_blocks(L, nae)
    loc = np.stack(np.meshgrid(np.arange(nae), np.arange(nae), np.arange(nae), indexing="ij"),
                   axis=-1).reshape(-1, 3)                     # site order ii, jj, kk
    center = box_hi / 2.0 if core_center is None else np.asarray(core_center, dtype=np.float64)
    half = 0.5 * core_cells * a
    # A block is dropped iff one of its sites (wrapped into the box) lies in
    # the core cube.  Only blocks whose corner-spanned box, taken over the
    # periodic images of their centroid, meets the core can qualify (a block
    # is a parallelepiped, the hull of its corner sites), so the sites are
    # generated for those candidates alone.
    corners0 = NODE_CORNERS * (nae - 1)
    cpos = origin + (nae * B[:, None, :] + corners0[None, :, :]) @ P          # [nb, 8, 3] unwrapped
    cen = cpos.mean(axis=1)
    shift = np.mod(cen - box_lo, box_hi - box_lo) + box_lo - cen               # wrap by the centroid
    cmin = cpos.min(axis=1) + shift
    cmax = cpos.max(axis=1) + shift
    L3 = box_hi - box_lo
    cand = np.zeros(B.shape[0], dtype=bool)
    for img in range(27):
        sft = (np.array([img % 3, (img // 3) % 3, img // 9]) - 1) * L3
        cand |= np.all((cmin + sft < center + half) & (cmax + sft > center - half), axis=1)
    drop = np.zeros(B.shape[0], dtype=bool)
    idx = np.flatnonzero(cand)
    chunk = max(1, 2_000_000 // (nae ** 3))
    for s in range(0, idx.size, chunk):
        ii = idx[s:s + chunk]
        n = nae * B[ii, None, :] + loc[None, :, :]
        pos = origin + n @ P
        pos = np.mod(pos - box_lo, box_hi - box_lo) + box_lo
        inside = np.all(np.abs(pos - center) < half, axis=-1)
        drop[ii] = inside.any(axis=1)
    nsite = nae ** 3
    need = -(-(min_atoms - nsite * int(drop.sum())) // nsite)
    if need > 0:
        cw = np.mod(cen - box_lo, L3) + box_lo
        d = cw - center
        d -= L3 * np.round(d / L3)
        dist = np.sqrt((d * d).sum(axis=1))
        order = np.lexsort((np.arange(B.shape[0]), dist))
        extra = order[~drop[order]][:need]
        drop[extra] = True
    keepB = B[~drop]
    dropB = B[drop]
    # atoms: all sites of dropped blocks, wrapped into the box
    if dropB.shape[0]:
        n = nae * dropB[:, None, :] + loc[None, :, :]
        x = (origin + n.reshape(-1, 3) @ P)
        x = np.mod(x - box_lo, box_hi - box_lo) + box_lo
        x = np.where(x >= box_hi, box_lo, x)
    else:
        x = np.zeros((0, 3))
    # elements: node = corner site, contiguous block (centroid in the box)
    corners = NODE_CORNERS * (nae - 1)
    nodex = origin + (nae * keepB[:, None, :] + corners[None, :, :]) @ P
    nodex = np.ascontiguousarray(nodex, dtype=np.float64)
    ijk, w = corner_quadrature((nae, nae, nae))
    sysm = CACSystem(box_lo, box_hi, mass, np.ascontiguousarray(x, dtype=np.float64),
                     np.zeros_like(x), (nae, nae, nae), nodex, np.zeros_like(nodex), ijk, w,
                     blocks=keepB, a0=a, cells=L)
    if temperature > 0.0:
        assign_velocities(sysm, temperature, seed)
    return sysm


def vertex_clusters(s: CACSystem):
    """Node slots 8*e + k grouped by mesh vertex, as (start[ncl+1], slots).
    Node k of the element of block B (corner flags c_k) sits next to the
    block-grid vertex V = B + c_k: the 8 elements meeting at V put their
    nodes on the lattice points nae*V - c' around it.  Vertices are matched
    by the lattice point nae*V wrapped into the periodic box (integer
    arithmetic in units of a/2); clusters are ordered by their first slot.
    Without block indices the node tags of a data file name the vertices."""
    if s.n_elements == 0:
        return None
    if s.blocks is None:
        if s.node_tags is None:
            return None
        key = np.asarray(s.node_tags, dtype=np.int64).reshape(-1)
    else:
        nae = int(s.nae[0])
        V = s.blocks[:, None, :].astype(np.int64) + NODE_CORNERS[None, :, :]          # [ne, 8, 3]
        n = nae * V.reshape(-1, 3)
        c2 = np.stack([n[:, 0] + n[:, 1], n[:, 0] + n[:, 2], n[:, 1] + n[:, 2]], axis=1)  # a/2 units
        c2 = np.mod(c2, 2 * s.cells)
        key = (c2[:, 2] * (2 * s.cells) + c2[:, 1]) * (2 * s.cells) + c2[:, 0]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    cut = np.flatnonzero(np.diff(ks)) + 1
    groups = np.split(order, cut)
    groups.sort(key=lambda g: int(g.min()))
    start = np.zeros(len(groups) + 1, dtype=np.int32)
    start[1:] = np.cumsum([len(g) for g in groups])
    slots = np.concatenate([np.sort(g) for g in groups]).astype(np.int32)
    return start, slots


def assign_velocities(s: CACSystem, temperature: float, seed: int) -> None:
    """Gaussian velocities at `temperature` over atoms and nodes, zero total
    momentum, exact temperature after rescale (build-defined)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n, ne = s.n_atoms, s.n_elements
    m_node = s.mass_e / 8.0
    masses = np.concatenate([np.full(n, s.mass), np.full(8 * ne, m_node)])
    v = rng.standard_normal((n + 8 * ne, 3)) * np.sqrt(BOLTZ_METAL * temperature / (masses * MVV2E_METAL))[:, None]
    p = (masses[:, None] * v).sum(axis=0)
    v -= p / masses.sum()
    ke = 0.5 * MVV2E_METAL * float((masses[:, None] * v * v).sum())
    dof = 3 * (n + 8 * ne) - 3
    t_now = 2.0 * ke / (dof * BOLTZ_METAL)
    if t_now > 0:
        v *= np.sqrt(temperature / t_now)
    s.v = np.ascontiguousarray(v[:n])
    s.nodev = np.ascontiguousarray(v[n:].reshape(ne, 8, 3))


def kinetic_energy(s: CACSystem, v=None, nodev=None) -> float:
    v = s.v if v is None else v
    nodev = s.nodev if nodev is None else nodev
    return 0.5 * MVV2E_METAL * (s.mass * float((v * v).sum()) + s.mass_e / 8.0 * float((nodev * nodev).sum()))


def temperature_of(s: CACSystem, ke: float) -> float:
    dof = 3 * (s.n_atoms + 8 * s.n_elements) - 3
    return 2.0 * ke / (dof * BOLTZ_METAL)


# named workloads -----------------------------------------------------------

def lj_c0_system(seed: int = 4928459) -> CACSystem:
    """LJ Cu: 40^3-cell periodic box, central 10^3-cell core, nae=5, 10 K."""
    return build_cac_system(40, 5, 10, temperature=10.0, seed=seed)


def edge_dipole(s: CACSystem, nu: float = 0.34, separation: Optional[float] = None, center=None) -> None:
    """Insert an edge-dislocation dipole (isotropic Volterra field) into the
    atom and node positions, in place.  Burgers vector b = (a/2)[110]
    (|b| = a/sqrt(2)), line along [001], the two cores `separation` apart
    along the glide-plane normal [-110]/sqrt(2), centered on `center` (default:
    the box center, separation half the atomistic core's width so both cores
    lie among the atoms).  Each dislocation displaces a point at (x', y')
    from its core, in the frame x' = b/|b|, y' = normal, by
        u_x' =  b/2pi [theta + x'y' / (2(1-nu) r^2)]
        u_y' = -b/2pi [(1-2nu)/(4(1-nu)) ln r^2 + (x'^2 - y'^2) / (4(1-nu) r^2)]
    (theta = atan2(y', x')); the second core carries -b.  Between the two
    branch cuts the material is displaced by the lattice vector b, so the
    periodic box stays consistent.  The positions are wrapped back into the
    box."""
    a = s.a0
    bmag = a / np.sqrt(2.0)
    ex = np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
    ey = np.array([-1.0, 1.0, 0.0]) / np.sqrt(2.0)
    c = (s.box_lo + s.box_hi) / 2.0 if center is None else np.asarray(center, dtype=np.float64)
    if separation is None:
        span = np.ptp(s.x, axis=0).min() if s.n_atoms else 0.25 * (s.box_hi - s.box_lo).min()
        separation = 0.5 * span
    cores = [(c - 0.5 * separation * ey, 1.0), (c + 0.5 * separation * ey, -1.0)]

    def disp(p):
        u = np.zeros_like(p)
        for c0, sign in cores:
            d = p - c0
            xp, yp = d @ ex, d @ ey
            r2 = np.maximum(xp * xp + yp * yp, 1e-12)
            ux = bmag / (2 * np.pi) * (np.arctan2(yp, xp) + xp * yp / (2 * (1 - nu) * r2))
            uy = -bmag / (2 * np.pi) * ((1 - 2 * nu) / (4 * (1 - nu)) * np.log(r2)
                                        + (xp * xp - yp * yp) / (4 * (1 - nu) * r2))
            u += sign * (ux[:, None] * ex[None, :] + uy[:, None] * ey[None, :])
        return u

    L3 = s.box_hi - s.box_lo
    if s.n_atoms:
        x = s.x + disp(s.x)
        s.x = np.ascontiguousarray(np.mod(x - s.box_lo, L3) + s.box_lo)
    if s.n_elements:
        X = s.nodex.reshape(-1, 3)
        X = (X + disp(X)).reshape(s.nodex.shape)
        # keep each element's nodes together: wrap by the element center
        cen = X.mean(axis=1)
        X = X + (np.mod(cen - s.box_lo, L3) + s.box_lo - cen)[:, None, :]
        s.nodex = np.ascontiguousarray(X)


def eam_c3_system(seed: int = 4928459, L: int = 777, nae: int = 21, core_cells: float = 106.0,
                  temperature: float = 300.0) -> CACSystem:
    """Large EAM CAC: periodic box of L^3 fcc cells (L a multiple of nae),
    elements of nae sites per edge with 8 corner integration points, an
    atomistic core (~8M atoms at the defaults, 202,612 blocks of 21^3 sites),
    an edge-dislocation dipole in the core, seeded velocities."""
    s = build_cac_system(L, nae, core_cells, temperature=0.0, seed=seed)
    edge_dipole(s)
    if temperature > 0.0:
        assign_velocities(s, temperature, seed)
    return s


def eam_c1_system(seed: int = 4928459, L: int = 198, core_cells: float = 42.0) -> CACSystem:
    """EAM Cu: L=198 cells (715.8 A), nae=11, an atomistic core of at least
    500,000 atoms (the 42-cell cube's 484,484 plus the 12 blocks nearest to
    it: 500,456 atoms) and 22,952 elements, 300 K."""
    return build_cac_system(L, 11, core_cells, temperature=300.0, seed=seed, min_atoms=500_000)
