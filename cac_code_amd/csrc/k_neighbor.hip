// k_neighbor.hip -- periodic wrap, cell binning and the three neighbor builds
// of the CAC path on gfx950: atoms (full list), elements (candidate element
// images), and integration points (interpolated sites of neighbor elements).
//
// Bit-exactness: this translation unit is compiled with -ffp-contract=off.
// Image coordinates are x + s*prd (one add per component), the value the
// reference's ghost copies carry (atom_vec_atomic.cpp:202-221,
// element.cpp:933-953); site positions are sum_k N_k*X'_k in node order
// (full_node_neighbor.cpp:141-159); rsq is summed x, y, z left to right and
// compared with <= (neighbor.cpp:609, full_node_neighbor.cpp:167,216).
// Search pruning (cells, element radii, element hulls) is conservative by
// construction, so the sets equal the brute-force sets.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "cac_internal.h"

namespace cac {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void img_shift(int img, int s[3]) {
  s[0] = img % 3 - 1;
  s[1] = (img / 3) % 3 - 1;
  s[2] = img / 9 - 1;
}

__device__ __forceinline__ double shifted(double x, int s, double prd) {
  return s ? x + (double)s * prd : x;
}

struct Box3 {
  double lo[3], hi[3], prd[3];
};

Box3 box_of(const cac_ctx *ctx) {
  Box3 b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = ctx->lo[d];
    b.hi[d] = ctx->hi[d];
    b.prd[d] = ctx->prd[d];
  }
  return b;
}

struct Grid3 {
  int n[3];
  double inv[3];
};

// ---------------------------------------------------------------- pbc
// Domain::pbc, domain.cpp:535-633.
__global__ void k_pbc_atoms(int n, double *__restrict__ x4, Box3 b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) {
    double x = x4[4 * i + d];
    if (x < b.lo[d]) x += b.prd[d];
    if (x >= b.hi[d]) {
      x -= b.prd[d];
      x = x > b.lo[d] ? x : b.lo[d];
    }
    x4[4 * i + d] = x;
  }
}

// Element centers (node mean, fix_nve.cpp:139-153) and Domain::elepbc
// (domain.cpp:647-745; the nodes move with the center, fixing the x-hi sign
// slip at domain.cpp:682).
__global__ void k_pbc_elements(int ne, double *__restrict__ nodex, double *__restrict__ xc, Box3 b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  double *X = nodex + 24 * (size_t)i;
  for (int d = 0; d < 3; ++d) {
    double c = 0.0;
    for (int k = 0; k < 8; ++k) c += X[3 * k + d];
    c = c / 8;
    if (c < b.lo[d]) {
      c += b.prd[d];
      for (int k = 0; k < 8; ++k) X[3 * k + d] += b.prd[d];
    }
    if (c >= b.hi[d]) {
      c -= b.prd[d];
      for (int k = 0; k < 8; ++k) X[3 * k + d] -= b.prd[d];
      c = c > b.lo[d] ? c : b.lo[d];
    }
    xc[3 * i + d] = c;
  }
}

// ---------------------------------------------------------------- cells
__device__ __forceinline__ int cell_coord(double x, double lo, double inv, int n) {
  int k = (int)floor((x - lo) * inv);
  return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

__global__ void k_cell_keys(int n, const double *__restrict__ p, int stride, Box3 b, Grid3 g,
                            uint32_t *__restrict__ key, uint32_t *__restrict__ val) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c0 = cell_coord(p[stride * i + 0], b.lo[0], g.inv[0], g.n[0]);
  int c1 = cell_coord(p[stride * i + 1], b.lo[1], g.inv[1], g.n[1]);
  int c2 = cell_coord(p[stride * i + 2], b.lo[2], g.inv[2], g.n[2]);
  key[i] = (uint32_t)((c2 * g.n[1] + c1) * g.n[0] + c0);
  val[i] = (uint32_t)i;
}

__global__ void k_fill_int(int n, int *__restrict__ a, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_cell_bounds(int n, const uint32_t *__restrict__ key, int *__restrict__ cs,
                              int *__restrict__ ce) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t k = key[p];
  if (p == 0 || key[p - 1] != k) cs[k] = p;
  if (p == n - 1 || key[p + 1] != k) ce[k] = p + 1;
}

// ---------------------------------------------------------------- atoms
// Full list (Neighbor::half_bin_no_newton, neighbor.cpp:546-625): every
// (j, image) != (i, self) with rsq <= cutneighsq, from the 27 cells around
// i's cell (cell edge >= cutneigh, >= 3 cells per dim).
//
// One workgroup of four wavefronts per non-empty cell.  The candidates (the
// atoms of the 27 surrounding cells, image shift applied) are staged in LDS
// once and shared by the cell's atoms; each wavefront then takes the cell's
// atoms in turn, tests the staged candidates 64 at a time and compacts the
// accepted entries into LDS.  Each entry is written straight to its rank in
// the row (rank = number of smaller keys), i.e. the row sorted by a key that
// puts the lattice-rounded displacement k = rint((x_j' - x_i) / spacing)
// before the entry value, in the atom_slot layout: ordered by |k|^2 first,
// then k_z, k_y, k_x.  On a lattice of that spacing every atom lists its
// neighbors in the same relative order, so at step k the 64 atoms of a
// force-pass wavefront gather neighboring rows of x4 instead of 64
// unrelated ones; and the rows run shell by shell, so the entries within
// the force cutoff fill the front of every row and the pair passes stop at
// the last one.  spacing = 0 sorts by entry value.  Keys are unique (they
// end in the entry), so the row is independent of the candidate order.  A
// cell with more candidates than the LDS holds reads them from global memory
// instead.  High word: min(|k|^2, 255) << 24 | (k_z + 128) << 16 |
// (k_y + 128) << 8 | (k_x + 128), components clamped to [-128, 127].
__host__ __device__ __forceinline__ uint32_t list_key_hi(int kx, int ky, int kz) {
  const int k2 = kx * kx + ky * ky + kz * kz;
  return ((uint32_t)(k2 < 255 ? k2 : 255) << 24) | ((uint32_t)(kz + 128) << 16) | ((uint32_t)(ky + 128) << 8) |
         (uint32_t)(kx + 128);
}
__device__ __forceinline__ uint64_t list_key(double dx, double dy, double dz, double qinv, uint32_t entry) {
  if (qinv <= 0.0) return entry;
  auto q = [&](double d) {
    const int k = (int)rint(d * qinv);
    return k < -128 ? -128 : (k > 127 ? 127 : k);
  };
  return ((uint64_t)list_key_hi(q(dx), q(dy), q(dz)) << 32) | entry;
}

// LDS written by some lanes of a wavefront is read by others after this.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void cell_split(int c, const Grid3 &g, int cc[3]) {
  cc[0] = c % g.n[0];
  cc[1] = (c / g.n[0]) % g.n[1];
  cc[2] = c / (g.n[0] * g.n[1]);
}

// Neighbor cell t (0..26, x fastest) of cell ci: its id and image.
__device__ __forceinline__ int cell_around(const int ci[3], int t, const Grid3 &g, int *img) {
  int cc[3] = {ci[0] + t % 3 - 1, ci[1] + (t / 3) % 3 - 1, ci[2] + t / 9 - 1}, s[3];
  for (int d = 0; d < 3; ++d) {
    s[d] = 0;
    if (cc[d] < 0) { cc[d] += g.n[d]; s[d] = -1; }
    else if (cc[d] >= g.n[d]) { cc[d] -= g.n[d]; s[d] = 1; }
  }
  *img = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
  return (cc[2] * g.n[1] + cc[1]) * g.n[0] + cc[0];
}

// Sorted position p starts a cell: hflag[p] = 1 (hflag[n] = 0 for the scan).
__global__ void k_cell_heads(int n, const uint32_t *__restrict__ key, int *__restrict__ hflag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) hflag[p] = (p == 0 || key[p - 1] != key[p]) ? 1 : 0;
  if (p == n) hflag[p] = 0;
}

// Non-empty cells in cell order.
__global__ void k_cell_list(int n, const uint32_t *__restrict__ key, const int *__restrict__ hflag,
                            const int *__restrict__ hpos, int *__restrict__ cells) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n && hflag[p]) cells[hpos[p]] = (int)key[p];
}

// The largest candidate count (atoms in the 27 cells around) of any
// non-empty cell: one thread per cell of the list, grid-stride over the
// count the scan left on the device, then a block maximum and one atomic
// per block.  (Over all atoms, only the one thread in ~30 that starts a cell
// had work, so each wavefront carried two lanes of scattered reads.)
constexpr int kMaxcandBlocks = 2048;
__global__ __launch_bounds__(kBlock) void k_cell_maxcand(const int *__restrict__ ncells_p,
                                                         const int *__restrict__ cells, Grid3 g,
                                                         const int *__restrict__ cs, const int *__restrict__ ce,
                                                         int *__restrict__ maxcand) {
  __shared__ int smax;
  if (threadIdx.x == 0) smax = 0;
  __syncthreads();
  const int ncells = *ncells_p;
  int best = 0;
  for (int t = (int)(blockIdx.x * blockDim.x + threadIdx.x); t < ncells; t += (int)(gridDim.x * blockDim.x)) {
    int ci[3], tot = 0, img;
    cell_split(cells[t], g, ci);
    for (int u = 0; u < 27; ++u) {
      const int cn = cell_around(ci, u, g, &img);
      tot += ce[cn] - cs[cn];
    }
    best = tot > best ? tot : best;
  }
  atomicMax(&smax, best);
  __syncthreads();
  if (threadIdx.x == 0 && smax > 0) atomicMax(maxcand, smax);
}

constexpr int kNeighWaves = 4;
constexpr int kCodeWords = 64;  // presence bitmap of the lattice-rounded displacements, per wavefront

// Dense code of a list key's lattice-rounded displacement (k_x, k_y, k_z),
// each in [-qk, qk]: the rank crank[((k_z + qk) * D + k_y + qk) * D + k_x +
// qk] (D = 2 qk + 1) of its high word among the box's high words, so codes
// order as the keys' high words order them; -1 outside that box.
__device__ __forceinline__ int list_code(uint64_t key, int qk, int D, const int16_t *__restrict__ crank) {
  const uint32_t hi = (uint32_t)(key >> 32);
  const int kx = (int)(hi & 255u) - 128, ky = (int)((hi >> 8) & 255u) - 128, kz = (int)((hi >> 16) & 255u) - 128;
  if (kx < -qk || kx > qk || ky < -qk || ky > qk || kz < -qk || kz > qk) return -1;
  return crank[((kz + qk) * D + (ky + qk)) * D + (kx + qk)];
}

__global__ __launch_bounds__(64 * kNeighWaves) void k_atom_neigh(
    int ncells, const int *__restrict__ cells, const double *__restrict__ x4, Box3 b, Grid3 g,
    const uint32_t *__restrict__ sorted, const int *__restrict__ cs, const int *__restrict__ ce, double cutsq,
    int maxneigh, int cap2, int ccap, double qinv, int qk, const int16_t *__restrict__ crank, int n,
    uint32_t *__restrict__ neigh,
    int *__restrict__ num, double *__restrict__ xhold, int *__restrict__ flags) {
  // LDS: keys[kNeighWaves][cap2] (u64), then sx, sy, sz [ccap] (f64), se [ccap] (u32)
  extern __shared__ uint64_t smem[];
  __shared__ int pre[28], rstart[27], rimg[27];
  __shared__ uint32_t cbits_all[kNeighWaves][kCodeWords];
  __shared__ int cpre_all[kNeighWaves][kCodeWords];
  const int w = xcd_block();
  if (w >= ncells) return;  // whole workgroup
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = cells[w];
  if (tid < 27) {
    int ci[3], img;
    cell_split(c, g, ci);
    const int cn = cell_around(ci, tid, g, &img);
    rstart[tid] = cs[cn];
    rimg[tid] = img;
    pre[tid + 1] = ce[cn] - cs[cn];
  }
  __syncthreads();
  if (tid == 0) {
    pre[0] = 0;
    for (int t = 0; t < 27; ++t) pre[t + 1] += pre[t];
  }
  __syncthreads();
  const int T = pre[27];
  // candidate f (0 <= f < T): range r = max{r : pre[r] <= f}
  auto fetch = [&](int f, double &xj, double &yj, double &zj, uint32_t &e) {
    int lo = 0, hi = 27;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (pre[mid] <= f) lo = mid;
      else hi = mid;
    }
    const int j = (int)sorted[rstart[lo] + f - pre[lo]];
    const int img = rimg[lo];
    int s[3];
    img_shift(img, s);
    xj = shifted(x4[4 * (size_t)j], s[0], b.prd[0]);
    yj = shifted(x4[4 * (size_t)j + 1], s[1], b.prd[1]);
    zj = shifted(x4[4 * (size_t)j + 2], s[2], b.prd[2]);
    e = (uint32_t)j | ((uint32_t)img << 27);
  };
  double *sx = reinterpret_cast<double *>(smem + kNeighWaves * cap2);
  double *sy = sx + ccap, *sz = sy + ccap;
  uint32_t *se = reinterpret_cast<uint32_t *>(sz + ccap);
  const bool staged = T <= ccap;
  if (staged)
    for (int f = tid; f < T; f += 64 * kNeighWaves) fetch(f, sx[f], sy[f], sz[f], se[f]);
  __syncthreads();
  uint64_t *keys = smem + wave * cap2;
  uint32_t *cbits = cbits_all[wave];
  int *cpre = cpre_all[wave];
  const int D = 2 * qk + 1, nwords = qk > 0 ? (D * D * D + 31) >> 5 : 0;
  for (int h = cs[c] + wave; h < ce[c]; h += kNeighWaves) {
    const int i = (int)sorted[h];
    const double xi = x4[4 * (size_t)i], yi = x4[4 * (size_t)i + 1], zi = x4[4 * (size_t)i + 2];
    const uint32_t self = (uint32_t)i | ((uint32_t)kImgSelf << 27);
    int cnt = 0;
    if (staged) {
      // accepted candidates are compacted as their index f; keys are formed
      // afterwards for those only
      for (int f0 = 0; f0 < T; f0 += 64) {
        const int f = f0 + lane;
        bool acc = false;
        if (f < T && se[f] != self) {
          const double delx = xi - sx[f], dely = yi - sy[f], delz = zi - sz[f];
          acc = delx * delx + dely * dely + delz * delz <= cutsq;
        }
        const unsigned long long am = __ballot(acc);
        if (acc) {
          const int pos = cnt + __popcll(am & ((1ull << lane) - 1ull));
          if (pos < cap2) keys[pos] = (uint64_t)f;
        }
        cnt += __popcll(am);
      }
      wave_sync();
      const int kept = cnt < cap2 ? cnt : cap2;
      for (int a = lane; a < kept; a += 64) {
        const int f = (int)keys[a];
        keys[a] = list_key(sx[f] - xi, sy[f] - yi, sz[f] - zi, qinv, se[f]);
      }
    } else {
      for (int f0 = 0; f0 < T; f0 += 64) {
        const int f = f0 + lane;
        bool acc = false;
        uint64_t key = 0;
        if (f < T) {
          double xj, yj, zj;
          uint32_t e;
          fetch(f, xj, yj, zj, e);
          if (e != self) {
            const double delx = xi - xj, dely = yi - yj, delz = zi - zj;
            const double rsq = delx * delx + dely * dely + delz * delz;
            acc = rsq <= cutsq;
            key = list_key(xj - xi, yj - yi, zj - zi, qinv, e);
          }
        }
        const unsigned long long am = __ballot(acc);
        if (acc) {
          const int pos = cnt + __popcll(am & ((1ull << lane) - 1ull));
          if (pos < cap2) keys[pos] = key;
        }
        cnt += __popcll(am);
      }
    }
    wave_sync();
    const int kept = cnt < cap2 ? cnt : cap2;
    // On a lattice the rounded displacements of one row are distinct: mark
    // each in a presence bitmap, and a key's rank is the number of marked
    // codes below its own (a prefix popcount).  A row where two keys share a
    // code (or one lies outside the code box) is ranked by the sweep below,
    // which also breaks such ties by entry; both give the same row.
    bool fast = nwords > 0;
    if (fast) {
      if (lane < nwords) cbits[lane] = 0u;
      wave_sync();
      bool clash = false;
      for (int a = lane; a < kept; a += 64) {
        const int code = list_code(keys[a], qk, D, crank);
        if (code < 0) {
          clash = true;
        } else {
          const uint32_t bit = 1u << (code & 31);
          clash |= (atomicOr(&cbits[code >> 5], bit) & bit) != 0u;
        }
      }
      fast = __ballot(clash) == 0ull;
      wave_sync();
    }
    if (fast) {
      const uint32_t wv = lane < nwords ? cbits[lane] : 0u;
      int incl = __popc(wv);
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        incl += lane >= o ? t : 0;
      }
      cpre[lane] = incl - __popc(wv);
      wave_sync();
      for (int a = lane; a < kept; a += 64) {
        const uint64_t ka = keys[a];
        const int code = list_code(ka, qk, D, crank);
        const int r = cpre[code >> 5] + __popc(cbits[code >> 5] & ((1u << (code & 31)) - 1u));
        if (r < maxneigh) neigh[atom_slot(i, r, maxneigh)] = (uint32_t)ka;
      }
    } else {
    // rank of keys a and a + 64 in one sweep over the row (broadcast reads,
    // unrolled so that several are in flight)
    for (int a = lane; a < ((kept + 127) & ~127); a += 128) {
      const uint64_t ka = a < kept ? keys[a] : 0, kb = a + 64 < kept ? keys[a + 64] : 0;
      int ra = 0, rb = 0;
#pragma unroll 8
      for (int q = 0; q < kept; ++q) {
        const uint64_t t = keys[q];
        ra += t < ka ? 1 : 0;
        rb += t < kb ? 1 : 0;
      }
      if (a < kept && ra < maxneigh) neigh[atom_slot(i, ra, maxneigh)] = (uint32_t)ka;
      if (a + 64 < kept && rb < maxneigh) neigh[atom_slot(i, rb, maxneigh)] = (uint32_t)kb;
    }
    }
    if (lane == 0) {
      num[i] = cnt;
      if (cnt > maxneigh) {
        atomicOr(&flags[0], 1);
        atomicMax(&flags[1], cnt);
      }
      xhold[i] = xi;
      xhold[n + i] = yi;
      xhold[2 * (size_t)n + i] = zi;
    }
    wave_sync();  // every lane has read keys before the next atom refills them
  }
}

// ---------------------------------------------------------------- elements
// Geometry at rebuild: element radius R = max_k |X_k - c| (every site is a
// convex combination of the nodes, so it lies within R of the center).
//
// Also a parallelepiped hull: with corner signs s_k (node k at parent
// corner (s_k + 1)/2), J = (1/8) sum_k (X_k - c) s_k^T and residuals
// r_k = X_k - c - J s_k, |r_k| <= delta.  A site is sum_k N_k X_k with
// convex weights N_k, i.e. c + J xi + sum_k N_k r_k with xi = sum_k N_k s_k in
// [-1,1]^3, so it lies within delta of P = c + J [-1,1]^3.  A point q within
// cutneigh of a site is within rho = cutneigh + delta of P, and then
// |(J^-1 (q - c))_d| <= 1 + rho |row_d(J^-1)| for every d: hull[16 i ..]
// holds J^-1 (row-major), these three bounds and the row norms |row_d(J^-1)|
// (slack for rounding added; a degenerate J gets J^-1 = 0 so the test always
// passes).
__device__ __forceinline__ bool hull_hit(const double *__restrict__ hull, int j, double dx, double dy, double dz) {
  const double *h = hull + 16 * (size_t)j;
  for (int d = 0; d < 3; ++d) {
    const double u = h[3 * d] * dx + h[3 * d + 1] * dy + h[3 * d + 2] * dz;
    if (!(fabs(u) <= h[9 + d])) return false;
  }
  return true;
}

__global__ void k_elem_radius(int ne, const double *__restrict__ nodex, const double *__restrict__ xc,
                              double cutneigh, double *__restrict__ erad, double *__restrict__ hull,
                              unsigned long long *__restrict__ rmax_bits) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const double *X = nodex + 24 * (size_t)i;
  double c[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) c[d] += X[3 * k + d];
  for (int d = 0; d < 3; ++d) c[d] = c[d] / 8;
  double r2 = 0.0;
  for (int k = 0; k < 8; ++k) {
    double dx = X[3 * k] - c[0], dy = X[3 * k + 1] - c[1], dz = X[3 * k + 2] - c[2];
    double q = dx * dx + dy * dy + dz * dz;
    r2 = q > r2 ? q : r2;
  }
  double R = sqrt(r2) * (1.0 + 1e-12) + 1e-9;
  erad[4 * i] = c[0];
  erad[4 * i + 1] = c[1];
  erad[4 * i + 2] = c[2];
  erad[4 * i + 3] = R;
  atomicMax(rmax_bits, (unsigned long long)__double_as_longlong(R));
  // corner signs of nodes 0..7: (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1)
  constexpr int kSx = 0x66, kSy = 0xCC, kSz = 0xF0;
  double J[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int k = 0; k < 8; ++k) {
    const double sg[3] = {(kSx >> k) & 1 ? 1.0 : -1.0, (kSy >> k) & 1 ? 1.0 : -1.0, (kSz >> k) & 1 ? 1.0 : -1.0};
    for (int d = 0; d < 3; ++d)
      for (int e = 0; e < 3; ++e) J[d][e] += sg[e] * (X[3 * k + d] - c[d]);
  }
  for (int d = 0; d < 3; ++d)
    for (int e = 0; e < 3; ++e) J[d][e] *= 0.125;
  double del2 = 0.0;
  for (int k = 0; k < 8; ++k) {
    const double sg[3] = {(kSx >> k) & 1 ? 1.0 : -1.0, (kSy >> k) & 1 ? 1.0 : -1.0, (kSz >> k) & 1 ? 1.0 : -1.0};
    double q = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double r = X[3 * k + d] - c[d] - (J[d][0] * sg[0] + J[d][1] * sg[1] + J[d][2] * sg[2]);
      q += r * r;
    }
    del2 = q > del2 ? q : del2;
  }
  double A[3][3];  // adjugate: J^-1 = A / det
  A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  A[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  A[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  A[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  A[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double det = J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
  double cn[3];
  for (int e = 0; e < 3; ++e) cn[e] = sqrt(J[0][e] * J[0][e] + J[1][e] * J[1][e] + J[2][e] * J[2][e]);
  const double rho = (cutneigh + sqrt(del2)) * (1.0 + 1e-9) + 1e-6;
  double *h = hull + 16 * (size_t)i;
  const bool ok = isfinite(det) && fabs(det) > 1e-6 * cn[0] * cn[1] * cn[2];
  for (int d = 0; d < 3; ++d) {
    double rn = 0.0;
    for (int e = 0; e < 3; ++e) {
      const double v = ok ? A[d][e] / det : 0.0;
      h[3 * d + e] = v;
      rn += v * v;
    }
    h[9 + d] = ok ? (1.0 + rho * sqrt(rn)) * (1.0 + 1e-9) + 1e-9 : 1e300;
    h[12 + d] = sqrt(rn) * (1.0 + 1e-9);
  }
  h[15] = 0.0;
}

// Position sum_k N_k X_k of a site with shape functions N, summed in node
// order (full_node_neighbor.cpp:141-159).
__device__ __forceinline__ void interp_site(const double *__restrict__ N, const double X[24], double out[3]) {
  double a = 0.0, b = 0.0, c = 0.0;
  for (int k = 0; k < 8; ++k) {
    a += N[k] * X[3 * k];
    b += N[k] * X[3 * k + 1];
    c += N[k] * X[3 * k + 2];
  }
  out[0] = a;
  out[1] = b;
  out[2] = c;
}

// Candidate element images of element i: every (j, image) != (i, self)
// with |c_i - c_j'| <= cutneigh + R_i + R_j.  Any site within cutneigh of
// an integration point of i lies in such an element (triangle inequality),
// so this list is a superset of what element_neighbor.cpp:37-66 needs.
__device__ __forceinline__ bool elem_pair(const double *__restrict__ erad, int i, int j, int img,
                                          const Box3 &b, double cutneigh) {
  if (j == i && img == kImgSelf) return false;
  int s[3];
  img_shift(img, s);
  double dx = erad[4 * i] - shifted(erad[4 * j], s[0], b.prd[0]);
  double dy = erad[4 * i + 1] - shifted(erad[4 * j + 1], s[1], b.prd[1]);
  double dz = erad[4 * i + 2] - shifted(erad[4 * j + 2], s[2], b.prd[2]);
  double reach = cutneigh + erad[4 * i + 3] + erad[4 * j + 3];
  return dx * dx + dy * dy + dz * dz <= reach * reach;
}

__global__ void k_elem_neigh_cells(int ne, const double *__restrict__ erad, Box3 b, Grid3 g,
                                   const uint32_t *__restrict__ sorted, const int *__restrict__ cs,
                                   const int *__restrict__ ce, double cutneigh, int maxelist,
                                   uint32_t *__restrict__ eneigh, int *__restrict__ enumb,
                                   int *__restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const int ci[3] = {cell_coord(erad[4 * i], b.lo[0], g.inv[0], g.n[0]),
                     cell_coord(erad[4 * i + 1], b.lo[1], g.inv[1], g.n[1]),
                     cell_coord(erad[4 * i + 2], b.lo[2], g.inv[2], g.n[2])};
  int cnt = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz}, s[3];
        for (int d = 0; d < 3; ++d) {
          s[d] = 0;
          if (cc[d] < 0) { cc[d] += g.n[d]; s[d] = -1; }
          else if (cc[d] >= g.n[d]) { cc[d] -= g.n[d]; s[d] = 1; }
        }
        const int img = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
        const int c = (cc[2] * g.n[1] + cc[1]) * g.n[0] + cc[0];
        for (int q = cs[c]; q < ce[c]; ++q) {
          const int j = (int)sorted[q];
          if (elem_pair(erad, i, j, img, b, cutneigh)) {
            if (cnt < maxelist) eneigh[(size_t)cnt * ne + i] = (uint32_t)j | ((uint32_t)img << 27);
            ++cnt;
          }
        }
      }
  enumb[i] = cnt;
  if (cnt > maxelist) {
    atomicOr(&flags[2], 1);
    atomicMax(&flags[3], cnt);
  }
}

// Small boxes (fewer than 3 element cells per dim): all j, all 27 images.
__global__ void k_elem_neigh_brute(int ne, const double *__restrict__ erad, Box3 b, double cutneigh,
                                   int maxelist, uint32_t *__restrict__ eneigh, int *__restrict__ enumb,
                                   int *__restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  int cnt = 0;
  for (int img = 0; img < 27; ++img)
    for (int j = 0; j < ne; ++j)
      if (elem_pair(erad, i, j, img, b, cutneigh)) {
        if (cnt < maxelist) eneigh[(size_t)cnt * ne + i] = (uint32_t)j | ((uint32_t)img << 27);
        ++cnt;
      }
  enumb[i] = cnt;
  if (cnt > maxelist) {
    atomicOr(&flags[2], 1);
    atomicMax(&flags[3], cnt);
  }
}

// Element geometry for the site searches.
struct IpGeom {
  int ne, nint, nae[3], nwords;
  double pstep[3];  // parent-coordinate step 2/(nae-1), as the host computes it
};

// Raw entry store as two 32-bit halves (lo = j | img << 27, hi = site).
// Written this way on purpose: hipcc (ROCm 7.2, gfx950) folds
// `(uint64_t)lo | ((uint64_t)0xFFFFFFFF << 32)` into a store whose high
// dword is the sign extension of j instead of all ones, which turned atom
// entries into site entries.
__device__ __forceinline__ void put_raw(uint64_t *__restrict__ out, int pos, uint32_t lo, uint32_t hi) {
  reinterpret_cast<uint2 *>(out)[pos] = make_uint2(lo, hi);
}

// Atom cells for the coupled searches.
struct AtomBins {
  int n;
  const double *x4;
  Grid3 g;
  const uint32_t *bin;
  const int *bs, *be;
};

// The 27 cell ranges around a point, one per lane 0..26 (length 0 on the
// other lanes), walked as one sequence: lane_prefix32 gives each lane the
// inclusive prefix of the lengths, range_of the range holding position f
// (all lanes must call it; f beyond the total gives a valid lane index).
__device__ __forceinline__ int lane_prefix32(int len, int lane) {
  int incl = len;
  for (int o = 1; o < 32; o <<= 1) {
    const int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  return incl;
}

__device__ __forceinline__ int range_of(int incl, int f) {
  int r = 0;  // number of ranges ending at or before f
  for (int step = 16; step > 0; step >>= 1) {
    const int idx = r + step - 1;
    const int v = __shfl(incl, idx < 26 ? idx : 26);
    if (idx < 27 && v <= f) r += step;
  }
  return r < 26 ? r : 26;
}

// Atoms within cutneigh of point p (coupled integration-point lists).  p may
// lie outside the box (element nodes follow the element center): the cells
// around its wrapped position q = p - t*prd are searched, and an atom found
// with cell image s is image s + t of p.  Lanes 0..26 read the ranges of the
// 27 cells at once (x fastest, cells beyond one period skipped); the ranges
// are then walked as one sequence, 64 atoms at a time, each lane locating
// its range by binary search over the prefix sums.  Accepted atoms are
// compacted in lane order (cell order, then bin order) as raw entries
// j | img << 27 | kAtomSite << 32.  Wave-uniform; returns the new count.
__device__ int scan_atoms(const AtomBins &ab, const Box3 &b, const double p[3], double cutneigh, int cnt, int cap,
                          uint64_t *__restrict__ out, int lane) {
  const double cutsq = cutneigh * cutneigh;
  int t[3], ci[3];
  for (int d = 0; d < 3; ++d) {
    t[d] = p[d] < b.lo[d] ? -1 : (p[d] >= b.hi[d] ? 1 : 0);
    const double q = p[d] - (double)t[d] * b.prd[d];
    ci[d] = cell_coord(q, b.lo[d], ab.g.inv[d], ab.g.n[d]);
  }
  int qs = 0, len = 0, img = 0;
  if (lane < 27) {
    int cc[3] = {ci[0] + lane % 3 - 1, ci[1] + (lane / 3) % 3 - 1, ci[2] + lane / 9 - 1}, s[3];
    bool far = false;
    for (int d = 0; d < 3; ++d) {
      s[d] = t[d];
      if (cc[d] < 0) { cc[d] += ab.g.n[d]; s[d] -= 1; }
      else if (cc[d] >= ab.g.n[d]) { cc[d] -= ab.g.n[d]; s[d] += 1; }
      far |= s[d] < -1 || s[d] > 1;
    }
    if (!far) {  // beyond one period: out of reach when prd > 2*(cutneigh + R)
      img = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
      const int c = (cc[2] * ab.g.n[1] + cc[1]) * ab.g.n[0] + cc[0];
      qs = ab.bs[c];
      len = ab.be[c] - qs;
    }
  }
  const int incl = lane_prefix32(len, lane), excl = incl - len;
  const int total = __shfl(incl, 26);
  for (int f0 = 0; f0 < total; f0 += 64) {
    const int f = f0 + lane;
    const int r = range_of(incl, f);
    const int rs = __shfl(qs, r), rx = __shfl(excl, r), rimg = __shfl(img, r);
    bool acc = false;
    int j = 0;
    if (f < total) {
      int s[3];
      img_shift(rimg, s);
      j = (int)ab.bin[rs + f - rx];
      const double delx = shifted(ab.x4[4 * (size_t)j], s[0], b.prd[0]) - p[0];
      const double dely = shifted(ab.x4[4 * (size_t)j + 1], s[1], b.prd[1]) - p[1];
      const double delz = shifted(ab.x4[4 * (size_t)j + 2], s[2], b.prd[2]) - p[2];
      acc = delx * delx + dely * dely + delz * delz <= cutsq;
    }
    const unsigned long long am = __ballot(acc);
    if (acc) {
      const int pos = cnt + __popcll(am & ((1ull << lane) - 1ull));
      if (pos < cap) put_raw(out, pos, (uint32_t)j | ((uint32_t)rimg << 27), kAtomSite);
    }
    cnt += __popcll(am);
  }
  return cnt;
}

// Sites of element image (cj, cimg) within cutneigh of each point q[g]
// (g < kIpGroup) whose bit is set in gmask -- a group of integration points of
// one element, or of nearby atoms.  Candidate sites come from the hull: a site
// with parent coordinates xi within cutneigh of q has |xi_d - u_d| <=
// rho |row_d(J^-1)| < bound_d - 1, u = J^-1 (q - c') (hull_hit), and site
// index k along d has xi_d = k * 2/(nae_d - 1) - 1, so each point reaches a
// box of site indices.  Boxes are walked densely, 64 sites at a time in index
// order: one box per point, or, when the points are close together (their
// union box is no larger than the boxes' total), the union once, each site
// interpolated once and tested against every point.  Point g's accepted
// sites are compacted in lane order into out + g*cap from cnt[g] on (raw
// entries cj | cimg << 27 | site << 32, capacity cap); a site accepted by any
// point is marked in cj's referenced-site bitmap.  `self` applies the
// rsq > 1e-4 exclusion of the own element (full_node_neighbor.cpp:216).  X
// (24 doubles) and box (6 ints per point) are the wavefront's LDS slots.
// Wave-uniform.
constexpr int kIpGroup = 8;

__device__ void scan_sites_group(const IpGeom &ig, const double *__restrict__ nodex,
                                 const double *__restrict__ erad, const double *__restrict__ hull, const Box3 &b,
                                 int cj, int cimg, bool self, const double (*q)[3], unsigned gmask,
                                 double cutneigh, int (&cnt)[kIpGroup], int cap, uint64_t *__restrict__ out,
                                 uint32_t *__restrict__ bitmap, double *X, int *box, int lane) {
  const double cutsq = cutneigh * cutneigh;
  int s[3];
  img_shift(cimg, s);
  wave_sync();  // earlier readers of X and box are done
  if (lane < 24) X[lane] = shifted(nodex[24 * (size_t)cj + lane], s[lane % 3], b.prd[lane % 3]);
  if (lane < kIpGroup) {  // lane g: the index box of point g (lo > hi: none)
    int lo[3] = {1, 1, 1}, hi[3] = {0, 0, 0};
    if ((gmask >> lane) & 1u) {
      const double *h = hull + 16 * (size_t)cj;
      const double dx = q[lane][0] - shifted(erad[4 * cj], s[0], b.prd[0]);
      const double dy = q[lane][1] - shifted(erad[4 * cj + 1], s[1], b.prd[1]);
      const double dz = q[lane][2] - shifted(erad[4 * cj + 2], s[2], b.prd[2]);
      bool reach = true;
      int l[3], u[3];
      for (int d = 0; d < 3; ++d) {
        const double c = h[3 * d] * dx + h[3 * d + 1] * dy + h[3 * d + 2] * dz;
        const double w = h[9 + d] - 1.0, half = 0.5 * (double)(ig.nae[d] - 1);
        const double flo = fmax((c - w + 1.0) * half - 1e-6, 0.0);
        const double fhi = fmin((c + w + 1.0) * half + 1e-6, (double)(ig.nae[d] - 1));
        l[d] = (int)ceil(flo);
        u[d] = (int)floor(fhi);
        reach = reach && l[d] <= u[d];
      }
      if (reach)
        for (int d = 0; d < 3; ++d) {
          lo[d] = l[d];
          hi[d] = u[d];
        }
    }
    for (int d = 0; d < 3; ++d) {
      box[6 * lane + d] = lo[d];
      box[6 * lane + 3 + d] = hi[d];
    }
  }
  wave_sync();
  unsigned live = 0;
  int ulo[3] = {1 << 30, 1 << 30, 1 << 30}, uhi[3] = {-1, -1, -1};
  long vsum = 0;
  for (int g = 0; g < kIpGroup; ++g) {
    const int *bg = box + 6 * g;
    if (!((gmask >> g) & 1u) || bg[0] > bg[3]) continue;
    live |= 1u << g;
    vsum += (long)(bg[3] - bg[0] + 1) * (bg[4] - bg[1] + 1) * (bg[5] - bg[2] + 1);
    for (int d = 0; d < 3; ++d) {
      ulo[d] = min(ulo[d], bg[d]);
      uhi[d] = max(uhi[d], bg[3 + d]);
    }
  }
  if (!live) return;
  const uint32_t lo32 = (uint32_t)cj | ((uint32_t)cimg << 27);
  // sites of box [lo, hi] (index order) against the points in gm
  auto walk = [&](int lo0, int lo1, int lo2, int hi0, int hi1, int hi2, unsigned gm) {
    const int nz = hi2 - lo2 + 1, nyz = (hi1 - lo1 + 1) * nz, tot = (hi0 - lo0 + 1) * nyz;
    const float rnyz = 1.0f / (float)nyz, rnz = 1.0f / (float)nz;
    for (int t0 = 0; t0 < tot; t0 += 64) {
      const int t = t0 + lane;
      // t -> (ia, ib, ic), ic fastest: float quotient, then exact correction
      int ia = (int)((float)t * rnyz);
      ia -= ia * nyz > t ? 1 : 0;
      ia += (ia + 1) * nyz <= t ? 1 : 0;
      const int rem = t - ia * nyz;
      int ib = (int)((float)rem * rnz);
      ib -= ib * nz > rem ? 1 : 0;
      ib += (ib + 1) * nz <= rem ? 1 : 0;
      const int a = lo0 + ia, bq = lo1 + ib, cq = lo2 + rem - ib * nz;
      const bool valid = t < tot;
      int site = 0;
      double xs[3] = {0.0, 0.0, 0.0};
      if (valid) {
        site = (a * ig.nae[1] + bq) * ig.nae[2] + cq;
        // shape functions of the site, with the operations (and so the bits)
        // of the host's site table (cac_upload_elements)
        const double px = ig.pstep[0] * (double)a - 1, py = ig.pstep[1] * (double)bq - 1,
                     pz = ig.pstep[2] * (double)cq - 1;
        double N[8];
        N[0] = 0.125 * (1 - px) * (1 - py) * (1 - pz);
        N[1] = 0.125 * (1 + px) * (1 - py) * (1 - pz);
        N[2] = 0.125 * (1 + px) * (1 + py) * (1 - pz);
        N[3] = 0.125 * (1 - px) * (1 + py) * (1 - pz);
        N[4] = 0.125 * (1 - px) * (1 - py) * (1 + pz);
        N[5] = 0.125 * (1 + px) * (1 - py) * (1 + pz);
        N[6] = 0.125 * (1 + px) * (1 + py) * (1 + pz);
        N[7] = 0.125 * (1 - px) * (1 + py) * (1 + pz);
        interp_site(N, X, xs);
      }
      bool any = false;
#pragma unroll
      for (int g = 0; g < kIpGroup; ++g)
        if ((gm >> g) & 1u) {
          const double delx = xs[0] - q[g][0], dely = xs[1] - q[g][1], delz = xs[2] - q[g][2];
          const double rsq = delx * delx + dely * dely + delz * delz;
          const bool acc = valid && rsq <= cutsq && (!self || rsq > 1.0e-4);
          const unsigned long long am = __ballot(acc);
          if (acc) {
            const int pos = cnt[g] + __popcll(am & ((1ull << lane) - 1ull));
            if (pos < cap) put_raw(out + (size_t)g * cap, pos, lo32, (uint32_t)site);
          }
          cnt[g] += __popcll(am);
          any |= acc;
        }
      if (any) atomicOr(&bitmap[(size_t)cj * ig.nwords + (site >> 5)], 1u << (site & 31));
    }
  };
  const long vu = (long)(uhi[0] - ulo[0] + 1) * (uhi[1] - ulo[1] + 1) * (uhi[2] - ulo[2] + 1);
  if (vu <= vsum) {
    walk(ulo[0], ulo[1], ulo[2], uhi[0], uhi[1], uhi[2], live);
  } else {
    for (int g = 0; g < kIpGroup; ++g)
      if ((live >> g) & 1u) {
        const int *bg = box + 6 * g;
        walk(__builtin_amdgcn_readfirstlane(bg[0]), __builtin_amdgcn_readfirstlane(bg[1]),
             __builtin_amdgcn_readfirstlane(bg[2]), __builtin_amdgcn_readfirstlane(bg[3]),
             __builtin_amdgcn_readfirstlane(bg[4]), __builtin_amdgcn_readfirstlane(bg[5]), 1u << g);
      }
  }
}

// Integration-point lists (Neighbor::node_neighbor,
// full_node_neighbor.cpp:19-239): for integration point g of element i, every
// site of every candidate element image with rsq <= cutneighsq, then the
// sites of i itself with rsq > 1e-4 (full_node_neighbor.cpp:216), which drops
// the integration point's own site; with coupling, then the atoms within
// cutneigh.
//
// One wavefront per group of up to kIpGroup integration points of one
// element (they share the element's candidate list).  Lanes test 64
// candidate elements at a time against the element radius, for every point
// of the group; each survivor is scanned once for the points that reach it
// (scan_sites_group).  Per point, the list order is candidate order, then
// site index order: deterministic.
__global__ __launch_bounds__(256, 4) void k_ip_neigh(IpGeom ig, const double *__restrict__ nodex,
                                                  const double *__restrict__ spa_int,
                                                  const double *__restrict__ erad, const double *__restrict__ hull,
                                                  const uint32_t *__restrict__ eneigh, const int *__restrict__ enumb,
                                                  Box3 b, double cutneigh, int maxnip, uint64_t *__restrict__ tmp,
                                                  int *__restrict__ ipnum, uint32_t *__restrict__ bitmap,
                                                  int *__restrict__ flags, bool coupled, AtomBins ab,
                                                  int *__restrict__ etouch) {
  const int ngrp = (ig.nint + kIpGroup - 1) / kIpGroup;
  const int lane = threadIdx.x & 63;
  const int wv = (xcd_block() * (int)blockDim.x + (int)threadIdx.x) >> 6;
  if (wv >= ig.ne * ngrp) return;  // whole wavefront
  const int i = wv / ngrp, g0 = (wv % ngrp) * kIpGroup;
  const int ng = ig.nint - g0 < kIpGroup ? ig.nint - g0 : kIpGroup;
  // the group's points in LDS (lane g interpolates point g)
  __shared__ double sip[4][kIpGroup][3], snode[4][24];
  __shared__ int sbox[4][6 * kIpGroup];
  double (*ip)[3] = sip[threadIdx.x >> 6];
  if (lane < kIpGroup) {
    double Xi[24];
    for (int q = 0; q < 24; ++q) Xi[q] = nodex[24 * (size_t)i + q];
    double t[3] = {0.0, 0.0, 0.0};
    if (lane < ng) interp_site(spa_int + 8 * (g0 + lane), Xi, t);
    ip[lane][0] = t[0];
    ip[lane][1] = t[1];
    ip[lane][2] = t[2];
  }
  wave_sync();
  const unsigned all = (1u << ng) - 1u;
  uint64_t *out = tmp + ((size_t)i * ig.nint + g0) * maxnip;
  int cnt[kIpGroup];
#pragma unroll
  for (int g = 0; g < kIpGroup; ++g) cnt[g] = 0;
  const int nel = enumb[i] + 1;  // candidates, then the element itself
  for (int l0 = 0; l0 < nel; l0 += 64) {
    const int l = l0 + lane;
    int j = i, img = kImgSelf;
    unsigned keep = 0;
    if (l < nel) {
      if (l < nel - 1) {
        const uint32_t e = eneigh[(size_t)l * ig.ne + i];
        j = (int)(e & 0x7FFFFFFu);
        img = (int)(e >> 27);
      }
      int s[3];
      img_shift(img, s);
      const double cx = shifted(erad[4 * j], s[0], b.prd[0]);
      const double cy = shifted(erad[4 * j + 1], s[1], b.prd[1]);
      const double cz = shifted(erad[4 * j + 2], s[2], b.prd[2]);
      const double reach = cutneigh + erad[4 * j + 3];
#pragma unroll
      for (int g = 0; g < kIpGroup; ++g) {
        const double dx = ip[g][0] - cx, dy = ip[g][1] - cy, dz = ip[g][2] - cz;
        if (dx * dx + dy * dy + dz * dz <= reach * reach && hull_hit(hull, j, dx, dy, dz)) keep |= 1u << g;
      }
      keep &= all;
    }
    unsigned long long cm = __ballot(keep != 0);
    while (cm) {
      const int src = __ffsll((long long)cm) - 1;
      cm &= cm - 1;
      const int cj = __shfl(j, src);
      const int cimg = __shfl(img, src);
      const unsigned gm = (unsigned)__shfl((int)keep, src);
      scan_sites_group(ig, nodex, erad, hull, b, cj, cimg, cj == i && cimg == kImgSelf, ip, gm, cutneigh, cnt, maxnip,
                       out, bitmap, snode[threadIdx.x >> 6], sbox[threadIdx.x >> 6], lane);
    }
  }
  bool touched = false;
  for (int g = 0; g < ng; ++g) {
    int c = 0;  // cnt[g] without dynamic register indexing
#pragma unroll
    for (int h = 0; h < kIpGroup; ++h)
      if (h == g) c = cnt[h];
    const double p[3] = {ip[g][0], ip[g][1], ip[g][2]};
    if (coupled) {
      const int c0 = c;
      c = scan_atoms(ab, b, p, cutneigh, c, maxnip, out + (size_t)g * maxnip, lane);
      touched |= c > c0;
    }
    if (lane == 0) {
      ipnum[(size_t)i * ig.nint + g0 + g] = c;
      if (c > maxnip) {
        atomicOr(&flags[4], 1);
        atomicMax(&flags[5], c);
      }
    }
  }
  if (lane == 0 && touched) etouch[i] = 1;
}

// Element cells for the coupled atom searches (grid false: fewer than 3
// element cells per dimension, every element and image is a candidate).
struct ElemBins {
  int ne;
  const double *erad, *hull;
  bool grid;
  Grid3 g;
  const uint32_t *bin;
  const int *bs, *be;
};

// Coupled atom lists, step 1: flag the atoms that some element image can
// reach (|x - c_j'| <= cutneigh + R_j and inside the expanded hull).  One
// thread per atom.
__global__ __launch_bounds__(256) void k_atom_iface(int n, const double *__restrict__ x4, ElemBins eb, Box3 b,
                                                    double cutneigh, int *__restrict__ iflag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q[3] = {x4[4 * (size_t)i], x4[4 * (size_t)i + 1], x4[4 * (size_t)i + 2]};
  auto hit = [&](int j, int img) {
    int s[3];
    img_shift(img, s);
    const double dx = q[0] - shifted(eb.erad[4 * j], s[0], b.prd[0]);
    const double dy = q[1] - shifted(eb.erad[4 * j + 1], s[1], b.prd[1]);
    const double dz = q[2] - shifted(eb.erad[4 * j + 2], s[2], b.prd[2]);
    const double reach = cutneigh + eb.erad[4 * j + 3];
    return dx * dx + dy * dy + dz * dz <= reach * reach && hull_hit(eb.hull, j, dx, dy, dz);
  };
  int found = 0;
  if (eb.grid) {
    const int ci[3] = {cell_coord(q[0], b.lo[0], eb.g.inv[0], eb.g.n[0]),
                       cell_coord(q[1], b.lo[1], eb.g.inv[1], eb.g.n[1]),
                       cell_coord(q[2], b.lo[2], eb.g.inv[2], eb.g.n[2])};
    // the 27 ranges are read up front (independent loads), then scanned
    int rs[27], re[27], im[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int c = cell_around(ci, t, eb.g, &im[t]);
      rs[t] = eb.bs[c];
      re[t] = eb.be[c];
    }
#pragma unroll
    for (int t = 0; t < 27; ++t)
      for (int qq = rs[t]; qq < re[t] && !found; ++qq) found = hit((int)eb.bin[qq], im[t]);
  } else {
    for (int img = 0; img < 27 && !found; ++img)
      for (int j = 0; j < eb.ne && !found; ++j) found = hit(j, img);
  }
  iflag[i] = found;
}

// The same flags, one wavefront per non-empty atom cell (ag, abin, abs/abe:
// the atom cells of this build).  The cell's atoms have a bounding sphere
// (center m, radius h); an element image that reaches one of them has
// |c_j' - m| <= cutneigh + R_j + h and |(J^-1 (m - c_j'))_d| <= bound_d +
// h |row_d|, and its center lies in the 27 element cells around m's (the
// host uses this kernel only when the element cells are at least
// cutneigh + R_max + h wide).  Lanes test 64 candidate images at a time
// against the sphere; each survivor is then tested exactly (as in
// k_atom_iface) by every atom of the cell, a lane per atom.  Cells deep in
// the atomistic region see no candidate at all.
__global__ __launch_bounds__(256) void k_atom_iface_cells(int ncells, const int *__restrict__ cells,
                                                          const uint32_t *__restrict__ abin,
                                                          const int *__restrict__ abs_, const int *__restrict__ abe,
                                                          const double *__restrict__ x4, ElemBins eb, Box3 b,
                                                          double cutneigh, int *__restrict__ iflag) {
  const int lane = threadIdx.x & 63;
  const int w = (xcd_block() * (int)blockDim.x + (int)threadIdx.x) >> 6;
  if (w >= ncells) return;  // whole wavefront
  const int c = cells[w], hs = abs_[c], na = abe[c] - hs;
  // bounding box of the cell's atoms -> sphere (m, h)
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int a = lane; a < na; a += 64) {
    const int i = (int)abin[hs + a];
    for (int d = 0; d < 3; ++d) {
      const double x = x4[4 * (size_t)i + d];
      lo[d] = fmin(lo[d], x);
      hi[d] = fmax(hi[d], x);
    }
  }
  double m[3], h2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], o));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], o));
    }
    m[d] = 0.5 * (lo[d] + hi[d]);
    const double e = 0.5 * (hi[d] - lo[d]);
    h2 += e * e;
  }
  const double h = sqrt(h2) * (1.0 + 1e-9) + 1e-6;
  int ci[3], qs = 0, len = 0, img = 0;
  for (int d = 0; d < 3; ++d) ci[d] = cell_coord(m[d], b.lo[d], eb.g.inv[d], eb.g.n[d]);
  if (lane < 27) {
    const int cc = cell_around(ci, lane, eb.g, &img);
    qs = eb.bs[cc];
    len = eb.be[cc] - qs;
  }
  const int incl = lane_prefix32(len, lane), excl = incl - len;
  const int total = __shfl(incl, 26);
  for (int a0 = 0; a0 < na; a0 += 64) {
    const bool valid = a0 + lane < na;
    const int i = valid ? (int)abin[hs + a0 + lane] : 0;
    const double q[3] = {x4[4 * (size_t)i], x4[4 * (size_t)i + 1], x4[4 * (size_t)i + 2]};
    int found = 0;
    for (int f0 = 0; f0 < total; f0 += 64) {
      const int f = f0 + lane;
      const int r = range_of(incl, f);
      const int rs = __shfl(qs, r), rx = __shfl(excl, r), rimg = __shfl(img, r);
      const int j = f < total ? (int)eb.bin[rs + f - rx] : 0;
      bool pass = false;
      if (f < total) {
        int s[3];
        img_shift(rimg, s);
        const double dx = m[0] - shifted(eb.erad[4 * j], s[0], b.prd[0]);
        const double dy = m[1] - shifted(eb.erad[4 * j + 1], s[1], b.prd[1]);
        const double dz = m[2] - shifted(eb.erad[4 * j + 2], s[2], b.prd[2]);
        const double reach = cutneigh + eb.erad[4 * j + 3] + h;
        pass = dx * dx + dy * dy + dz * dz <= reach * reach;
        const double *hh = eb.hull + 16 * (size_t)j;
        for (int d = 0; d < 3 && pass; ++d) {
          const double u = hh[3 * d] * dx + hh[3 * d + 1] * dy + hh[3 * d + 2] * dz;
          pass = fabs(u) <= hh[9 + d] + h * hh[12 + d];
        }
      }
      unsigned long long sm = __ballot(pass);
      while (sm) {
        const int src = __ffsll((long long)sm) - 1;
        sm &= sm - 1;
        const int cj = __builtin_amdgcn_readfirstlane(__shfl(j, src));
        const int cimg = __builtin_amdgcn_readfirstlane(__shfl(rimg, src));
        int s[3];
        img_shift(cimg, s);
        const double dx = q[0] - shifted(eb.erad[4 * cj], s[0], b.prd[0]);
        const double dy = q[1] - shifted(eb.erad[4 * cj + 1], s[1], b.prd[1]);
        const double dz = q[2] - shifted(eb.erad[4 * cj + 2], s[2], b.prd[2]);
        const double reach = cutneigh + eb.erad[4 * cj + 3];
        if (dx * dx + dy * dy + dz * dz <= reach * reach && hull_hit(eb.hull, cj, dx, dy, dz)) found = 1;
      }
    }
    if (valid) iflag[i] = found;
  }
}

__global__ void k_iface_compact(int n, const int *__restrict__ iflag, const int *__restrict__ ipos,
                                int *__restrict__ iface) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && iflag[i]) iface[ipos[i]] = i;
}

// Coupled atom lists, step 2: the element sites within cutneigh of each
// flagged atom (raw entries, after the atom's own atom entries, capacity
// maxneigh in total).  One wavefront per kIpGroup consecutive flagged atoms
// (consecutive in index, so close in space).  The atoms of the group that
// share an element cell share its candidate walk: each candidate element
// image (the 27 cells around that cell, in cell then bin order; every image
// of every element for a box under 3 element cells wide) is tested against
// each of those atoms (radius, then hull), and each survivor is scanned once
// for the atoms that reach it (scan_sites_group).  Per atom the entries come
// in candidate order, then site index order.
__global__ __launch_bounds__(256) void k_atom_sites(int niface, const int *__restrict__ iface,
                                                    const double *__restrict__ x4, const int *__restrict__ anum,
                                                    IpGeom ig, const double *__restrict__ nodex,
                                                    ElemBins eb, Box3 b,
                                                    double cutneigh, int maxneigh, uint64_t *__restrict__ atmp,
                                                    int *__restrict__ asnum, uint32_t *__restrict__ bitmap,
                                                    int *__restrict__ etouch, int *__restrict__ flags) {
  const int lane = threadIdx.x & 63;
  const int r0 = ((xcd_block() * (int)blockDim.x + (int)threadIdx.x) >> 6) * kIpGroup;
  if (r0 >= niface) return;  // whole wavefront
  const int ng = niface - r0 < kIpGroup ? niface - r0 : kIpGroup;
  __shared__ double sq[4][kIpGroup][3], snode[4][24];
  __shared__ int sbox[4][6 * kIpGroup];
  double (*q)[3] = sq[threadIdx.x >> 6];
  int cell = 0;
  if (lane < ng) {
    const int i = iface[r0 + lane];
    const double p[3] = {x4[4 * (size_t)i], x4[4 * (size_t)i + 1], x4[4 * (size_t)i + 2]};
    q[lane][0] = p[0];
    q[lane][1] = p[1];
    q[lane][2] = p[2];
    if (eb.grid)
      cell = (cell_coord(p[2], b.lo[2], eb.g.inv[2], eb.g.n[2]) * eb.g.n[1] +
              cell_coord(p[1], b.lo[1], eb.g.inv[1], eb.g.n[1])) * eb.g.n[0] +
             cell_coord(p[0], b.lo[0], eb.g.inv[0], eb.g.n[0]);
  }
  wave_sync();
  uint64_t *out = atmp + (size_t)r0 * maxneigh;
  int cnt[kIpGroup];
#pragma unroll
  for (int g = 0; g < kIpGroup; ++g) cnt[g] = 0;
  unsigned todo = (1u << ng) - 1u;
  while (todo) {
    const int lc = __shfl(cell, __ffs((int)todo) - 1);
    const unsigned gmask = (unsigned)__ballot(lane < ng && cell == lc) & todo;
    todo &= ~gmask;
    auto visit = [&](bool valid, int j, int img) {
      unsigned keep = 0;
      if (valid) {
        int s[3];
        img_shift(img, s);
        const double cx = shifted(eb.erad[4 * j], s[0], b.prd[0]);
        const double cy = shifted(eb.erad[4 * j + 1], s[1], b.prd[1]);
        const double cz = shifted(eb.erad[4 * j + 2], s[2], b.prd[2]);
        const double reach = cutneigh + eb.erad[4 * j + 3];
#pragma unroll
        for (int g = 0; g < kIpGroup; ++g)
          if ((gmask >> g) & 1u) {
            const double dx = q[g][0] - cx, dy = q[g][1] - cy, dz = q[g][2] - cz;
            if (dx * dx + dy * dy + dz * dz <= reach * reach && hull_hit(eb.hull, j, dx, dy, dz)) keep |= 1u << g;
          }
      }
      unsigned long long cm = __ballot(keep != 0);
      while (cm) {
        const int src = __ffsll((long long)cm) - 1;
        cm &= cm - 1;
        const int cj = __shfl(j, src), cimg = __shfl(img, src);
        const unsigned gm = (unsigned)__shfl((int)keep, src);
        int before = 0, after = 0;
#pragma unroll
        for (int g = 0; g < kIpGroup; ++g) before += cnt[g];
        scan_sites_group(ig, nodex, eb.erad, eb.hull, b, cj, cimg, false, q, gm, cutneigh, cnt, maxneigh, out,
                         bitmap, snode[threadIdx.x >> 6], sbox[threadIdx.x >> 6], lane);
#pragma unroll
        for (int g = 0; g < kIpGroup; ++g) after += cnt[g];
        if (lane == 0 && after > before) etouch[cj] = 1;
      }
    };
    if (eb.grid) {
      int ci[3], qs = 0, len = 0, img = 0;
      cell_split(lc, eb.g, ci);
      if (lane < 27) {
        const int c = cell_around(ci, lane, eb.g, &img);
        qs = eb.bs[c];
        len = eb.be[c] - qs;
      }
      const int incl = lane_prefix32(len, lane), excl = incl - len;
      const int total = __shfl(incl, 26);
      for (int f0 = 0; f0 < total; f0 += 64) {
        const int f = f0 + lane;
        const int r = range_of(incl, f);
        const int rs = __shfl(qs, r), rx = __shfl(excl, r), rimg = __shfl(img, r);
        visit(f < total, f < total ? (int)eb.bin[rs + f - rx] : 0, rimg);
      }
    } else {
      const int tot = 27 * eb.ne;  // image-major, then element index
      for (int f0 = 0; f0 < tot; f0 += 64) {
        const int f = f0 + lane;
        const bool v = f < tot;
        visit(v, v ? f % eb.ne : 0, v ? f / eb.ne : 0);
      }
    }
  }
  for (int g = 0; g < ng; ++g) {
    int c = 0;  // cnt[g] without dynamic register indexing
#pragma unroll
    for (int h = 0; h < kIpGroup; ++h)
      if (h == g) c = cnt[h];
    if (lane == 0) {
      asnum[r0 + g] = c;
      const int tot = anum[iface[r0 + g]] + c;
      if (tot > maxneigh) {
        atomicOr(&flags[0], 1);
        atomicMax(&flags[1], tot);
      }
    }
  }
}

// Active-site numbering: per element, the referenced sites in site order;
// elements in index order.  wprefix[j][w] = first id of word w.  One
// wavefront per element, a lane per bitmap word, 64 words at a time with a
// wavefront prefix sum of the word counts.
__device__ __forceinline__ int wave_incl_sum(int v, int lane) {
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

__global__ void k_active_count(int ne, int nwords, const uint32_t *__restrict__ bitmap, int *__restrict__ wprefix,
                               int *__restrict__ ecount) {
  const int j = (int)((blockIdx.x * (long)blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (j >= ne) return;  // whole wavefront
  int run = 0;
  for (int w0 = 0; w0 < nwords; w0 += 64) {
    const int w = w0 + lane;
    const int c = w < nwords ? __popc(bitmap[(size_t)j * nwords + w]) : 0;
    const int incl = wave_incl_sum(c, lane);
    if (w < nwords) wprefix[(size_t)j * nwords + w] = run + incl - c;
    run += __shfl(incl, 63);
  }
  if (lane == 0) ecount[j] = run;
}

__global__ void k_active_finalize(int ne, int nwords, const uint32_t *__restrict__ bitmap,
                                  const int *__restrict__ estart, int *__restrict__ wprefix,
                                  int2 *__restrict__ asite) {
  const int j = (int)((blockIdx.x * (long)blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (j >= ne) return;  // whole wavefront
  const int base = estart[j];
  for (int w = lane; w < nwords; w += 64) {
    const size_t q = (size_t)j * nwords + w;
    int id = wprefix[q] + base;
    wprefix[q] = id;
    uint32_t bits = bitmap[q];
    while (bits) {
      const int bit = __ffs(bits) - 1;
      bits &= bits - 1;
      asite[id++] = make_int2(j, w * 32 + bit);
    }
  }
}

// Raw entry -> particle entry (row of x4 | image << 27): an atom j keeps
// its index, site `site` of element j becomes n + its active id.
__device__ __forceinline__ uint32_t particle_entry(uint64_t e, int n, int nwords, const uint32_t *__restrict__ bitmap,
                                                   const int *__restrict__ wprefix) {
  const int j = (int)(e & 0x7FFFFFFu);
  const uint32_t img = (uint32_t)((e >> 27) & 31);
  const uint32_t site = (uint32_t)(e >> 32);
  if (site == kAtomSite) return (uint32_t)j | (img << 27);
  const size_t w = (size_t)j * nwords + (site >> 5);
  const int aid = wprefix[w] + __popc(bitmap[w] & ((1u << (site & 31)) - 1u));
  return (uint32_t)(n + aid) | (img << 27);
}

// Integration-point lists: raw entries -> particle entries in the blocked
// layout.  One wavefront per group of 8 rows: lane q of a pass takes entry
// q/8 of row q%8, so the writes are contiguous and the reads 8 runs.
__global__ void k_ip_translate(int nip, int maxnip, const uint64_t *__restrict__ tmp, const int *__restrict__ ipnum,
                               int n, int nwords, const uint32_t *__restrict__ bitmap,
                               const int *__restrict__ wprefix, uint32_t *__restrict__ out) {
  const int g = (int)((blockIdx.x * (long)blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (g * kRowGroup >= nip) return;  // whole wavefront
  const int pr = g * kRowGroup + (lane & 7);
  int c = lane < kRowGroup && pr < nip ? min(ipnum[pr], maxnip) : 0;
  int kmax = c;
  for (int o = 1; o < kRowGroup; o <<= 1) kmax = max(kmax, __shfl_xor(kmax, o));
  kmax = __shfl(kmax, 0);
  c = __shfl(c, lane & 7);  // the count of this lane's row
  for (int q = lane; q < kRowGroup * kmax; q += 64) {
    const int k = q >> 3;
    if (k < c)
      out[((size_t)g * maxnip + k) * kRowGroup + (q & 7)] =
          particle_entry(tmp[(size_t)pr * maxnip + k], n, nwords, bitmap, wprefix);
  }
}

// Coupled atom lists: the raw site entries of flagged atom r go after the
// atom's atom entries, and each pair also to the compact array:
// sent[sbase[r] + k] = (active site) << 32 | (i | img << 27).
__global__ void k_atom_site_translate(int niface, int maxneigh, const int *__restrict__ iface,
                                      const uint64_t *__restrict__ atmp, const int *__restrict__ asnum,
                                      const int *__restrict__ anum, int n, int nwords,
                                      const uint32_t *__restrict__ bitmap, const int *__restrict__ wprefix,
                                      uint32_t *__restrict__ aneigh, const int *__restrict__ sbase,
                                      uint64_t *__restrict__ sent) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= niface) return;
  const int i = iface[r], a = anum[i];
  const uint64_t *row = atmp + (size_t)r * maxneigh;
  uint64_t *out = sent + sbase[r];
  for (int k = 0; k < asnum[r]; ++k) {
    const uint32_t e = particle_entry(row[k], n, nwords, bitmap, wprefix);
    if (a + k < maxneigh) aneigh[atom_slot(i, a + k, maxneigh)] = e;
    out[k] = ((uint64_t)((e & 0x7FFFFFFu) - (uint32_t)n) << 32) | ((uint32_t)i | (e & ~0x7FFFFFFu));
  }
}

__global__ void k_atom_site_count(int niface, const int *__restrict__ iface, const int *__restrict__ asnum,
                                  int *__restrict__ anum) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < niface) anum[iface[r]] += asnum[r];
}

// Every entry of a blocked list must name a row of x4 (< nrow) with an image
// code < 27: the pair passes gather with it.  The first bad entry found is
// reported as (row, slot, entry) in flags[8..10] with flags[7] set.  One
// thread per row; rows come in groups of G.
template <int G>
__global__ void k_validate_list(int rows, int M, const uint32_t *__restrict__ list, const int *__restrict__ num,
                                uint32_t nrow, int *__restrict__ flags) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= rows) return;
  const int c = num[p];
  if (c < 0 || c > M) {
    if (atomicCAS(&flags[7], 0, 1) == 0) {
      flags[8] = p;
      flags[9] = -1;
      flags[10] = c;
    }
    return;
  }
  const uint32_t *row = list + (G == kAtomGroup ? atom_slot(p, 0, M) : (size_t)(p / G) * M * G + (p % G));
  for (int k = 0; k < c; ++k) {
    const uint32_t e = row[G == kAtomGroup ? atom_off(k) : (size_t)k * G];
    if ((e & 0x7FFFFFFu) >= nrow || (e >> 27) >= 27u) {
      if (atomicCAS(&flags[7], 0, 1) == 0) {
        flags[8] = p;
        flags[9] = k;
        flags[10] = (int)e;
      }
      return;
    }
  }
}

// Neighbor::check_distance, neighbor.cpp:1176-1217: flag if any atom or node
// moved more than sqrt(triggersq) since the last build.
__global__ void k_check_disp(int n, const double *__restrict__ x4, const double *__restrict__ xhold, int ne,
                             const double *__restrict__ nodex, const double *__restrict__ nodexhold,
                             double triggersq, int *__restrict__ flags) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false;
  if (t < n) {
    double dx = x4[4 * t] - xhold[t];
    double dy = x4[4 * t + 1] - xhold[n + t];
    double dz = x4[4 * t + 2] - xhold[2 * (size_t)n + t];
    hit = dx * dx + dy * dy + dz * dz > triggersq;
  } else if (t < n + 8 * ne) {
    size_t q = (size_t)(t - n);
    double dx = nodex[3 * q] - nodexhold[3 * q];
    double dy = nodex[3 * q + 1] - nodexhold[3 * q + 1];
    double dz = nodex[3 * q + 2] - nodexhold[3 * q + 2];
    hit = dx * dx + dy * dy + dz * dz > triggersq;
  }
  if (hit) atomicOr(&flags[6], 1);
}

inline unsigned grid_for(long n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

// Cells of edge >= cut*1.0001 over the box; 0 if fewer than 3 per dim.
bool make_grid(const cac_ctx *ctx, double cut, Grid3 *g) {
  bool ok = true;
  for (int d = 0; d < 3; ++d) {
    g->n[d] = (int)(ctx->prd[d] / (cut * 1.0001));
    if (g->n[d] < 3) ok = false;
    g->inv[d] = g->n[d] > 0 ? g->n[d] / ctx->prd[d] : 0.0;
  }
  return ok;
}

// Bin points (stride doubles apart) into g: sorted point ids and per-cell
// [start, end) ranges, ascending id within a cell (stable radix sort).
int bin_points(cac_ctx *ctx, int np, const double *pts, int stride, const Grid3 &g, cac::DevBuf &bk,
               cac::DevBuf &bv, cac::DevBuf &bk2, cac::DevBuf &bv2, cac::DevBuf &bcs, cac::DevBuf &bce,
               uint32_t **sorted_out, int **cs_out, int **ce_out) {
  uint32_t *key, *val, *key2, *val2;
  int *cs, *ce;
  const long ncell = (long)g.n[0] * g.n[1] * g.n[2];
  int rc;
  if ((rc = ensure(ctx, bk, np, &key)) || (rc = ensure(ctx, bv, np, &val)) || (rc = ensure(ctx, bk2, np, &key2)) ||
      (rc = ensure(ctx, bv2, np, &val2)) || (rc = ensure(ctx, bcs, ncell, &cs)) || (rc = ensure(ctx, bce, ncell, &ce)))
    return rc;
  Box3 b = box_of(ctx);
  if (np > 0) {
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_for(np)), dim3(kBlock), 0, ctx->stream, np, pts, stride, b, g, key, val);
    int bits = 1;
    while ((1L << bits) < ncell) ++bits;
    size_t tmp_bytes = 0;
    CAC_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, key2, val, val2, np, 0, bits,
                                                    ctx->stream));
    char *tmp;
    if ((rc = ensure(ctx, ctx->b_sorttmp, tmp_bytes, &tmp))) return rc;
    CAC_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key2, val, val2, np, 0, bits,
                                                    ctx->stream));
  }
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(ncell)), dim3(kBlock), 0, ctx->stream, (int)ncell, cs, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(ncell)), dim3(kBlock), 0, ctx->stream, (int)ncell, ce, 0);
  if (np > 0)
    hipLaunchKernelGGL(k_cell_bounds, dim3(grid_for(np)), dim3(kBlock), 0, ctx->stream, np, key2, cs, ce);
  CAC_HIP(ctx, hipGetLastError());
  *sorted_out = val2;
  *cs_out = cs;
  *ce_out = ce;
  return CAC_OK;
}

AtomBins atom_bins(const cac_ctx *ctx) {
  AtomBins ab;
  ab.n = ctx->n;
  ab.x4 = ctx->x4;
  for (int d = 0; d < 3; ++d) {
    ab.g.n[d] = ctx->acell[d];
    ab.g.inv[d] = ctx->acell[d] > 0 ? ctx->acell[d] / ctx->prd[d] : 0.0;
  }
  ab.bin = ctx->abin;
  ab.bs = ctx->abin_s;
  ab.be = ctx->abin_e;
  return ab;
}

ElemBins elem_bins(const cac_ctx *ctx) {
  ElemBins eb;
  eb.ne = ctx->ne;
  eb.erad = ctx->erad;
  eb.hull = ctx->ehull;
  eb.grid = ctx->ecell[0] > 0;
  for (int d = 0; d < 3; ++d) {
    eb.g.n[d] = ctx->ecell[d];
    eb.g.inv[d] = ctx->ecell[d] > 0 ? ctx->ecell[d] / ctx->prd[d] : 0.0;
  }
  eb.bin = ctx->ebin;
  eb.bs = ctx->ebin_s;
  eb.be = ctx->ebin_e;
  return eb;
}

IpGeom ip_geom(const cac_ctx *ctx) {
  IpGeom ig;
  ig.ne = ctx->ne;
  ig.nint = ctx->nint;
  for (int d = 0; d < 3; ++d) {
    ig.nae[d] = ctx->nae[d];
    ig.pstep[d] = 2 / ((double)ctx->nae[d] - 1);
  }
  ig.nwords = ctx->nwords;
  return ig;
}

}  // namespace

int launch_pbc_and_centers(cac_ctx *ctx) {
  Box3 b = box_of(ctx);
  if (ctx->n > 0)
    hipLaunchKernelGGL(k_pbc_atoms, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->x4, b);
  if (ctx->ne > 0)
    hipLaunchKernelGGL(k_pbc_elements, dim3(grid_for(ctx->ne)), dim3(kBlock), 0, ctx->stream, ctx->ne,
                       ctx->nodex, ctx->xc, b);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

int launch_atom_neighbors(cac_ctx *ctx) {
  const int n = ctx->n;
  if (n == 0) return CAC_OK;
  Grid3 g;
  if (!make_grid(ctx, ctx->cutneigh, &g))
    return fail(ctx, CAC_ERR_BOX, "box must be at least 3 neighbor cutoffs wide in every dimension");
  for (int d = 0; d < 3; ++d) ctx->acell[d] = g.n[d];
  uint32_t *sorted;
  int *cs, *ce, rc;
  if ((rc = bin_points(ctx, n, ctx->x4, 4, g, ctx->b_akey, ctx->b_aval, ctx->b_akey2, ctx->b_aval2, ctx->b_acs,
                       ctx->b_ace, &sorted, &cs, &ce)))
    return rc;
  ctx->maxneigh = round_up(ctx->maxneigh, kLanes);
  const size_t rows = (size_t)round_up(n, kAtomGroup);
  if ((rc = ensure(ctx, ctx->b_aneigh, (size_t)ctx->maxneigh * rows, &ctx->aneigh)) ||
      (rc = ensure(ctx, ctx->b_anum, rows, &ctx->anum)) || (rc = ensure(ctx, ctx->b_xhold, 3 * (size_t)n, &ctx->xhold)))
    return rc;
  ctx->abin = sorted;
  ctx->abin_s = cs;
  ctx->abin_e = ce;
  hipEvent_t t0;
  timer_begin(ctx, "atom_neigh", &t0);
  int *hflag, *cells;
  if ((rc = ensure(ctx, ctx->b_ahead, 2 * ((size_t)n + 1), &hflag)) || (rc = ensure(ctx, ctx->b_acells, n, &cells)))
    return rc;
  int *hpos = hflag + n + 1;
  const uint32_t *key2 = static_cast<const uint32_t *>(ctx->b_akey2.p);
  hipLaunchKernelGGL(k_cell_heads, dim3(grid_for(n + 1)), dim3(kBlock), 0, ctx->stream, n, key2, hflag);
  size_t tmp_bytes = 0;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, hflag, hpos, n + 1, ctx->stream));
  char *tmp;
  if ((rc = ensure(ctx, ctx->b_scantmp, tmp_bytes, &tmp))) return rc;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, hflag, hpos, n + 1, ctx->stream));
  CAC_HIP(ctx, hipMemsetAsync(ctx->dflags + 13, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL(k_cell_list, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, key2, hflag, hpos, cells);
  hipLaunchKernelGGL(k_cell_maxcand, dim3(kMaxcandBlocks), dim3(kBlock), 0, ctx->stream, hpos + n, cells, g, cs, ce,
                     ctx->dflags + 13);
  int hv[2] = {0, 0};
  CAC_HIP(ctx, hipMemcpyAsync(&hv[0], hpos + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipMemcpyAsync(&hv[1], ctx->dflags + 13, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int ncells = hv[0], maxcand = hv[1];
  ctx->acells = cells;
  ctx->nacells = ncells;
  int cap2 = 64;
  while (cap2 < ctx->maxneigh) cap2 <<= 1;
  // LDS: the four wavefronts' key rows, then up to ccap staged candidates
  // (28 B each); cells with more candidates read them from global memory.
  const size_t kLdsMax = 64 * 1024, key_bytes = (size_t)kNeighWaves * cap2 * sizeof(uint64_t);
  if (key_bytes + 64 * 28 > kLdsMax)
    return fail(ctx, CAC_ERR_OVERFLOW, "atom list capacity above 1024 entries per atom");
  int ccap = round_up(maxcand, 64);
  const int ccap_max = (int)((kLdsMax - key_bytes) / 28) / 64 * 64;
  ccap = ccap < ccap_max ? ccap : ccap_max;
  const size_t lds = key_bytes + (size_t)ccap * 28;
  const double qinv = ctx->list_spacing > 0.0 ? 1.0 / ctx->list_spacing : 0.0;
  // code box of the rounded displacements within cutneigh (one spare unit
  // each way); 0 = rank every row by the sweep
  int qk = 0;
  if (qinv > 0.0) {
    const double kmax = ceil(ctx->cutneigh * qinv) + 1.0;
    if (kmax <= 16.0 && (2.0 * kmax + 1.0) * (2.0 * kmax + 1.0) * (2.0 * kmax + 1.0) <= 32.0 * kCodeWords)
      qk = (int)kmax;
  }
  if (qk > 0 && ctx->crank_qk != qk) {
    // the codes' ranks in key order (a table per box size, made once)
    const int D = 2 * qk + 1, nc = D * D * D;
    std::vector<uint32_t> hiw(nc);
    std::vector<int> idx(nc);
    for (int c = 0; c < nc; ++c) {
      hiw[c] = list_key_hi(c % D - qk, (c / D) % D - qk, c / (D * D) - qk);
      idx[c] = c;
    }
    std::sort(idx.begin(), idx.end(), [&](int u, int v) { return hiw[u] < hiw[v]; });
    ctx->crank_host.assign(nc, 0);
    for (int r = 0; r < nc; ++r) ctx->crank_host[idx[r]] = (int16_t)r;
    if ((rc = ensure(ctx, ctx->b_crank, (size_t)nc, &ctx->crank))) return rc;
    CAC_HIP(ctx, hipMemcpyAsync(ctx->crank, ctx->crank_host.data(), (size_t)nc * sizeof(int16_t),
                                hipMemcpyHostToDevice, ctx->stream));
    CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->crank_qk = qk;
  }
  hipLaunchKernelGGL(k_atom_neigh, dim3(xcd_grid(ncells, 1)), dim3(64 * kNeighWaves), lds, ctx->stream, ncells, cells,
                     ctx->x4, box_of(ctx), g, sorted, cs, ce, ctx->cutneigh * ctx->cutneigh, ctx->maxneigh, cap2, ccap,
                     qinv, qk, qk > 0 ? ctx->crank : nullptr, n, ctx->aneigh, ctx->anum, ctx->xhold, ctx->dflags);
  timer_end(ctx, "atom_neigh", t0);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

int launch_element_neighbors(cac_ctx *ctx) {
  const int ne = ctx->ne;
  if (ne == 0) return CAC_OK;
  int rc;
  unsigned long long *rmax_bits;
  if ((rc = ensure(ctx, ctx->b_erad, 4 * (size_t)ne, &ctx->erad)) ||
      (rc = ensure(ctx, ctx->b_ehull, 16 * (size_t)ne, &ctx->ehull)))
    return rc;
  // scratch for the radius max lives in the reduction buffer
  if ((rc = ensure(ctx, ctx->b_red, 64 * 16, &ctx->red))) return rc;
  rmax_bits = reinterpret_cast<unsigned long long *>(ctx->red);
  CAC_HIP(ctx, hipMemsetAsync(rmax_bits, 0, sizeof(unsigned long long), ctx->stream));
  hipLaunchKernelGGL(k_elem_radius, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->nodex, ctx->xc,
                     ctx->cutneigh, ctx->erad, ctx->ehull, rmax_bits);
  unsigned long long hbits = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&hbits, rmax_bits, sizeof(hbits), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double rmax;
  memcpy(&rmax, &hbits, sizeof(rmax));
  ctx->rmax = rmax;
  if ((rc = ensure(ctx, ctx->b_eneigh, (size_t)ctx->maxelist * ne, &ctx->eneigh)) ||
      (rc = ensure(ctx, ctx->b_enumb, ne, &ctx->enumb)))
    return rc;
  Grid3 g;
  const double ecut = ctx->cutneigh + 2.0 * rmax;
  hipEvent_t t0;
  if (make_grid(ctx, ecut, &g)) {
    for (int d = 0; d < 3; ++d) ctx->ecell[d] = g.n[d];
    uint32_t *sorted;
    int *cs, *ce;
    if ((rc = bin_points(ctx, ne, ctx->erad, 4, g, ctx->b_ekey, ctx->b_eval, ctx->b_ekey2, ctx->b_eval2, ctx->b_ecs,
                         ctx->b_ece, &sorted, &cs, &ce)))
      return rc;
    ctx->ebin = sorted;
    ctx->ebin_s = cs;
    ctx->ebin_e = ce;
    timer_begin(ctx, "elem_neigh", &t0);
    hipLaunchKernelGGL(k_elem_neigh_cells, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->erad,
                       box_of(ctx), g, sorted, cs, ce, ctx->cutneigh, ctx->maxelist, ctx->eneigh, ctx->enumb,
                       ctx->dflags);
    timer_end(ctx, "elem_neigh", t0);
  } else {
    for (int d = 0; d < 3; ++d) ctx->ecell[d] = 0;
    ctx->ebin = nullptr;
    ctx->ebin_s = ctx->ebin_e = nullptr;
    for (int d = 0; d < 3; ++d)
      if (ctx->prd[d] < ctx->cutneigh)
        return fail(ctx, CAC_ERR_BOX, "box narrower than the neighbor cutoff");
    timer_begin(ctx, "elem_neigh", &t0);
    hipLaunchKernelGGL(k_elem_neigh_brute, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->erad,
                       box_of(ctx), ctx->cutneigh, ctx->maxelist, ctx->eneigh, ctx->enumb, ctx->dflags);
    timer_end(ctx, "elem_neigh", t0);
  }
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

int launch_ip_neighbors(cac_ctx *ctx) {
  const int ne = ctx->ne;
  if (ne == 0) return CAC_OK;
  const int nip = ne * ctx->nint;
  int rc;
  ctx->maxnip = round_up(ctx->maxnip, kLanes);
  ctx->nwords = (ctx->nsite + 31) / 32;
  const size_t nbits = (size_t)ne * ctx->nwords;
  if ((rc = ensure(ctx, ctx->b_iptmp, (size_t)ctx->maxnip * nip, &ctx->iptmp)) ||
      (rc = ensure(ctx, ctx->b_ipnum, (size_t)round_up(nip, kRowGroup), &ctx->ipnum)) ||
      (rc = ensure(ctx, ctx->b_bitmap, nbits, &ctx->bitmap)) ||
      (rc = ensure(ctx, ctx->b_nodexhold, 24 * (size_t)ne, &ctx->nodexhold)) ||
      (rc = ensure(ctx, ctx->b_etouch, (size_t)ne, &ctx->etouch)))
    return rc;
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for((long)nbits)), dim3(kBlock), 0, ctx->stream, (int)nbits,
                     reinterpret_cast<int *>(ctx->bitmap), 0);
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->etouch, 0);
  const bool coupled = ctx->coupled && ctx->n > 0;
  const IpGeom ig = ip_geom(ctx);
  hipEvent_t t0;
  timer_begin(ctx, "ip_neigh", &t0);
  const long waves = (long)ne * ((ctx->nint + kIpGroup - 1) / kIpGroup);
  hipLaunchKernelGGL(k_ip_neigh, dim3(xcd_grid(64L * waves, kBlock)), dim3(kBlock), 0, ctx->stream, ig, ctx->nodex,
                     ctx->spa_int, ctx->erad, ctx->ehull, ctx->eneigh, ctx->enumb, box_of(ctx),
                     ctx->cutneigh, ctx->maxnip, ctx->iptmp, ctx->ipnum, ctx->bitmap, ctx->dflags, coupled,
                     atom_bins(ctx), ctx->etouch);
  timer_end(ctx, "ip_neigh", t0);
  CAC_HIP(ctx, hipMemcpyAsync(ctx->nodexhold, ctx->nodex, sizeof(double) * 24 * (size_t)ne,
                              hipMemcpyDeviceToDevice, ctx->stream));
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

// Coupled atom lists: flag the atoms some element can reach, compact them,
// and collect their element sites (raw entries in atmp).  Needs the element
// geometry and cells of this build and the site bitmap cleared by
// launch_ip_neighbors.
int launch_atom_sites(cac_ctx *ctx) {
  const int n = ctx->n, ne = ctx->ne;
  ctx->niface = 0;
  if (!ctx->coupled || n == 0 || ne == 0) return CAC_OK;
  int rc;
  if ((rc = ensure(ctx, ctx->b_iflag, (size_t)n + 1, &ctx->iflag)) ||
      (rc = ensure(ctx, ctx->b_ipos, (size_t)n + 1, &ctx->ipos)))
    return rc;
  hipEvent_t t0;
  timer_begin(ctx, "atom_sites", &t0);
  // per atom cell when the element cells are wide enough (see k_atom_iface_cells)
  double hcell2 = 0.0;
  for (int d = 0; d < 3; ++d) hcell2 += 0.25 * (ctx->prd[d] / ctx->acell[d]) * (ctx->prd[d] / ctx->acell[d]);
  bool by_cell = ctx->ecell[0] > 0 && ctx->nacells > 0;
  for (int d = 0; d < 3 && by_cell; ++d)
    by_cell = ctx->prd[d] / ctx->ecell[d] >= (ctx->cutneigh + ctx->rmax + sqrt(hcell2)) * 1.001 + 1e-6;
  if (by_cell)
    hipLaunchKernelGGL(k_atom_iface_cells, dim3(xcd_grid(64L * ctx->nacells, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->nacells, ctx->acells, ctx->abin, ctx->abin_s, ctx->abin_e, ctx->x4, elem_bins(ctx),
                       box_of(ctx), ctx->cutneigh, ctx->iflag);
  else
    hipLaunchKernelGGL(k_atom_iface, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, ctx->x4, elem_bins(ctx),
                       box_of(ctx), ctx->cutneigh, ctx->iflag);
  CAC_HIP(ctx, hipMemsetAsync(ctx->iflag + n, 0, sizeof(int), ctx->stream));
  size_t tmp_bytes = 0;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, ctx->iflag, ctx->ipos, n + 1, ctx->stream));
  char *tmp;
  if ((rc = ensure(ctx, ctx->b_scantmp, tmp_bytes, &tmp))) return rc;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, ctx->iflag, ctx->ipos, n + 1, ctx->stream));
  int niface = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&niface, ctx->ipos + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->niface = niface;
  if (niface == 0) {
    timer_end(ctx, "atom_sites", t0);
    return CAC_OK;
  }
  if ((rc = ensure(ctx, ctx->b_iface, (size_t)niface, &ctx->iface)) ||
      (rc = ensure(ctx, ctx->b_asnum, (size_t)niface + 1, &ctx->asnum)) ||
      (rc = ensure(ctx, ctx->b_atmp, (size_t)niface * ctx->maxneigh, &ctx->atmp)))
    return rc;
  hipLaunchKernelGGL(k_iface_compact, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, ctx->iflag, ctx->ipos,
                     ctx->iface);
  const long groups = (niface + kIpGroup - 1) / kIpGroup;
  hipLaunchKernelGGL(k_atom_sites, dim3(xcd_grid(64L * groups, kBlock)), dim3(kBlock), 0, ctx->stream, niface,
                     ctx->iface, ctx->x4, ctx->anum, ip_geom(ctx), ctx->nodex,
                     elem_bins(ctx), box_of(ctx), ctx->cutneigh, ctx->maxneigh, ctx->atmp, ctx->asnum, ctx->bitmap,
                     ctx->etouch, ctx->dflags);
  timer_end(ctx, "atom_sites", t0);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

// After a successful search: number the referenced sites, build the
// active-site table, place the site rows in x4 after the atoms and rewrite
// the lists as particle entries.
int launch_active_sites(cac_ctx *ctx) {
  const int ne = ctx->ne;
  ctx->nsent = 0;
  if (ne == 0) return CAC_OK;
  const int nip = ne * ctx->nint;
  const int nw = ctx->nwords;
  int rc;
  if ((rc = ensure(ctx, ctx->b_wprefix, (size_t)ne * nw, &ctx->wprefix)) ||
      (rc = ensure(ctx, ctx->b_ecount, (size_t)ne + 1, &ctx->ecount)) ||
      (rc = ensure(ctx, ctx->b_estart, (size_t)ne + 1, &ctx->estart)))
    return rc;
  int *estart = ctx->estart;
  hipLaunchKernelGGL(k_active_count, dim3(grid_for(64L * ne)), dim3(kBlock), 0, ctx->stream, ne, nw, ctx->bitmap,
                     ctx->wprefix, ctx->ecount);
  size_t tmp_bytes = 0;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, ctx->ecount, estart, ne + 1, ctx->stream));
  char *tmp;
  if ((rc = ensure(ctx, ctx->b_scantmp, tmp_bytes, &tmp))) return rc;
  // ecount[ne] = 0 so estart[ne] is the total
  CAC_HIP(ctx, hipMemsetAsync(ctx->ecount + ne, 0, sizeof(int), ctx->stream));
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, ctx->ecount, estart, ne + 1, ctx->stream));
  int total = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&total, estart + ne, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->nactive = total;
  const int n = ctx->n;
  if ((long)n + total >= (1L << 27)) return fail(ctx, CAC_ERR_OVERFLOW, "more than 2^27-1 atoms plus active sites");
  if ((rc = ensure(ctx, ctx->b_asite, (size_t)total, &ctx->asite)) ||
      (rc = ensure_keep(ctx, ctx->b_x4, 4 * ((size_t)n + total), 4 * (size_t)n, &ctx->x4)) ||
      (rc = ensure(ctx, ctx->b_ipneigh, (size_t)ctx->maxnip * round_up(nip, kRowGroup), &ctx->ipneigh)))
    return rc;
  ctx->xs4 = ctx->x4 + 4 * (size_t)n;
  hipLaunchKernelGGL(k_active_finalize, dim3(grid_for(64L * ne)), dim3(kBlock), 0, ctx->stream, ne, nw, ctx->bitmap,
                     estart, ctx->wprefix, ctx->asite);
  hipLaunchKernelGGL(k_ip_translate, dim3(grid_for(64L * ((nip + kRowGroup - 1) / kRowGroup))), dim3(kBlock), 0,
                     ctx->stream, nip, ctx->maxnip, ctx->iptmp,
                     ctx->ipnum, n, nw, ctx->bitmap, ctx->wprefix, ctx->ipneigh);
  if (ctx->niface > 0) {
    // bases of the interface atoms' pairs in the compact array (asnum[niface] = 0: the total)
    const int ni = ctx->niface;
    if ((rc = ensure(ctx, ctx->b_sbase, (size_t)ni + 1, &ctx->sbase))) return rc;
    CAC_HIP(ctx, hipMemsetAsync(ctx->asnum + ni, 0, sizeof(int), ctx->stream));
    size_t sb = 0;
    CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, sb, ctx->asnum, ctx->sbase, ni + 1, ctx->stream));
    if ((rc = ensure(ctx, ctx->b_scantmp, sb, &tmp))) return rc;
    CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, sb, ctx->asnum, ctx->sbase, ni + 1, ctx->stream));
    int ns = 0;
    CAC_HIP(ctx, hipMemcpyAsync(&ns, ctx->sbase + ni, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->nsent = ns;
    if ((rc = ensure(ctx, ctx->b_sent, (size_t)(ns > 0 ? ns : 1), &ctx->sent))) return rc;
    hipLaunchKernelGGL(k_atom_site_translate, dim3(grid_for(ctx->niface)), dim3(kBlock), 0, ctx->stream, ctx->niface,
                       ctx->maxneigh, ctx->iface, ctx->atmp, ctx->asnum, ctx->anum, n, nw, ctx->bitmap, ctx->wprefix,
                       ctx->aneigh, ctx->sbase, ctx->sent);
    hipLaunchKernelGGL(k_atom_site_count, dim3(grid_for(ctx->niface)), dim3(kBlock), 0, ctx->stream, ctx->niface,
                       ctx->iface, ctx->asnum, ctx->anum);
  }
  CAC_HIP(ctx, hipGetLastError());
  return validate_lists(ctx);
}

// ------------------------------------------------------------ clusters
// Per vertex cluster (one workgroup of 256 threads): the entries of its
// integration points' lists (<= 8 rows of the blocked ipneigh layout) are
// sorted in LDS (bitonic, 32-bit), the distinct values become the cluster's
// union (ascending: atoms first, then active sites element by element), and
// every list entry is rewritten as its index in the union (uint16, same
// layout).  The largest union size goes to flags[12]: it sets the LDS rows of
// the cluster passes.
struct NodeList8 {
  int k[8];
};

constexpr int kClBuildThreads = 256;

__device__ void bitonic_sort_lds32(uint32_t *a, int n2, int tid) {
  for (int size = 2; size <= n2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = tid; t < (n2 >> 1); t += kClBuildThreads) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint32_t x = a[lo], y = a[hi];
        if ((x > y) == up) {
          a[lo] = y;
          a[hi] = x;
        }
      }
      __syncthreads();
    }
}

__global__ __launch_bounds__(kClBuildThreads) void k_cl_build(
    int ncl, const int *__restrict__ cl_start, const int *__restrict__ cl_slot, int nint, NodeList8 nl,
    const uint32_t *__restrict__ ipneigh, const int *__restrict__ ipnum, int M, int cap2,
    uint32_t *__restrict__ cl_union, int *__restrict__ cl_nu, uint16_t *__restrict__ iplocal,
    int *__restrict__ flags) {
  extern __shared__ uint32_t cl_lds[];  // [cap2] sorted entries, then [cap2] union
  uint32_t *keys = cl_lds, *uni = cl_lds + cap2;
  __shared__ int s_u;
  const int c = blockIdx.x;
  if (c >= ncl) return;  // whole workgroup
  const int tid = (int)threadIdx.x, lane = tid & 63;
  const int s0 = cl_start[c], ns = cl_start[c + 1] - s0;
  int pid[kClMax], off[kClMax + 1];
  off[0] = 0;
#pragma unroll
  for (int i = 0; i < kClMax; ++i) {
    pid[i] = 0;
    int cnt = 0;
    if (i < ns) {
      const int slot = cl_slot[s0 + i];
      pid[i] = (slot >> 3) * nint + nl.k[slot & 7];
      cnt = min(ipnum[pid[i]], M);
    }
    off[i + 1] = off[i] + cnt;
  }
  const int m = off[kClMax];
  int n2 = 64;  // <= cap2: m <= kClMax * M
  while (n2 < m) n2 <<= 1;
  auto entry_ptr = [&](int p, int k) -> size_t { return ((size_t)(p >> 3) * M + k) * kRowGroup + (p & 7); };
  for (int t = tid; t < n2; t += kClBuildThreads) {
    uint32_t v = 0xFFFFFFFFu;  // above every entry (image codes are < 27)
    if (t < m) {
      int i = 0;
      while (t >= off[i + 1]) ++i;
      v = ipneigh[entry_ptr(pid[i], t - off[i])];
    }
    keys[t] = v;
  }
  __syncthreads();
  bitonic_sort_lds32(keys, n2, tid);
  if (tid < 64) {  // the first wavefront compacts the distinct values
    int u = 0;
    for (int t0 = 0; t0 < m; t0 += 64) {
      const int t = t0 + lane;
      const bool first = t < m && (t == 0 || keys[t] != keys[t - 1]);
      const unsigned long long bm = __ballot(first);
      if (first) uni[u + __popcll(bm & ((1ull << lane) - 1ull))] = keys[t];
      u += __popcll(bm);
    }
    if (lane == 0) {
      s_u = u;
      cl_nu[c] = u;
      atomicMax(&flags[12], u);
    }
  }
  __syncthreads();
  const int u = s_u;
  for (int t = tid; t < u; t += kClBuildThreads) cl_union[(size_t)c * cap2 + t] = uni[t];
  for (int t = tid; t < m; t += kClBuildThreads) {
    int i = 0;
    while (t >= off[i + 1]) ++i;
    const size_t q = entry_ptr(pid[i], t - off[i]);
    const uint32_t v = ipneigh[q];
    int lo = 0, hi = u;  // first union index with value >= v
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (uni[mid] < v) lo = mid + 1; else hi = mid;
    }
    iplocal[local_slot(pid[i], t - off[i], M)] = (uint16_t)lo;
  }
}

// Active sites that atom lists refer to (atom-element coupling): marks, then
// the ascending list of marked ids.
__global__ void k_mark_sent(long ns, const uint64_t *__restrict__ sent, int *__restrict__ mark) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ns) mark[sent[t] >> 32] = 1;
}

__global__ void k_compact_marks(int na, const int *__restrict__ mark, const int *__restrict__ pos,
                                int *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < na && mark[q]) out[pos[q]] = q;
}

// Check both lists after a build (rows of x4: n atoms + nactive sites).
int validate_lists(cac_ctx *ctx) {
  const uint32_t nrow = (uint32_t)(ctx->n + ctx->nactive);
  CAC_HIP(ctx, hipMemsetAsync(ctx->dflags + 7, 0, 4 * sizeof(int), ctx->stream));
  const int nip = ctx->ne * ctx->nint;
  if (ctx->n)
    hipLaunchKernelGGL(k_validate_list<kAtomGroup>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n,
                       ctx->maxneigh, ctx->aneigh, ctx->anum, nrow, ctx->dflags);
  int bad[4];
  CAC_HIP(ctx, hipMemcpyAsync(bad, ctx->dflags + 7, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (bad[0])
    return fail(ctx, CAC_ERR_DATA, "atom list entry out of range: row " + std::to_string(bad[1]) + " slot " +
                                       std::to_string(bad[2]) + " value " + std::to_string((unsigned)bad[3]) +
                                       " (n " + std::to_string(ctx->n) + ", active sites " +
                                       std::to_string(ctx->nactive) + ")");
  if (nip)
    hipLaunchKernelGGL(k_validate_list<kRowGroup>, dim3(grid_for(nip)), dim3(kBlock), 0, ctx->stream, nip,
                       ctx->maxnip, ctx->ipneigh, ctx->ipnum, nrow, ctx->dflags);
  CAC_HIP(ctx, hipMemcpyAsync(bad, ctx->dflags + 7, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (bad[0])
    return fail(ctx, CAC_ERR_DATA, "integration-point list entry out of range: row " + std::to_string(bad[1]) +
                                       " slot " + std::to_string(bad[2]) + " value " +
                                       std::to_string((unsigned)bad[3]) + " (n " + std::to_string(ctx->n) +
                                       ", active sites " + std::to_string(ctx->nactive) + ")");
  return CAC_OK;
}


// Cluster unions and union-index lists after a build (see k_cl_build), and
// the active sites atom lists refer to.
int launch_clusters(cac_ctx *ctx) {
  ctx->cl_ok = false;
  ctx->naref = 0;
  const int nip = ctx->ne * ctx->nint;
  if (ctx->ncl == 0 || !ctx->ip_on_nodes || nip == 0) return CAC_OK;
  int rc;
  const int M = ctx->maxnip;
  int cap2 = 64;
  while (cap2 < kClMax * M) cap2 <<= 1;
  const size_t lds = (size_t)cap2 * 2 * sizeof(uint32_t);
  if (lds > ctx->lds_bytes) return CAC_OK;  // lists too long for the cluster passes
  if ((rc = ensure(ctx, ctx->b_cl_union, (size_t)ctx->ncl * cap2, &ctx->cl_union)) ||
      (rc = ensure(ctx, ctx->b_cl_nu, (size_t)ctx->ncl, &ctx->cl_nu)) ||
      (rc = ensure(ctx, ctx->b_iplocal, local_stride(M) * (size_t)nip, &ctx->iplocal)))
    return rc;
  ctx->cl_stride = cap2;
  NodeList8 nl;
  for (int k = 0; k < 8; ++k) nl.k[k] = ctx->node_list[k];
  CAC_HIP(ctx, hipMemsetAsync(ctx->dflags + 11, 0, 2 * sizeof(int), ctx->stream));
  CAC_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_cl_build),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_cl_build, dim3(ctx->ncl), dim3(kClBuildThreads), lds, ctx->stream, ctx->ncl, ctx->cl_start,
                     ctx->cl_slot,
                     ctx->nint, nl, ctx->ipneigh, ctx->ipnum, M, cap2, ctx->cl_union, ctx->cl_nu, ctx->iplocal,
                     ctx->dflags);
  CAC_HIP(ctx, hipGetLastError());
  if (ctx->coupled && ctx->niface > 0 && ctx->nactive > 0) {
    const int na = ctx->nactive;
    int *pos;
    // marks [0, na] (the last one stays 0) and their exclusive sums [na+1, 2na+2)
    if ((rc = ensure(ctx, ctx->b_amark, 2 * (size_t)na + 2, &ctx->amark)) ||
        (rc = ensure(ctx, ctx->b_aref, (size_t)na, &ctx->aref)))
      return rc;
    pos = ctx->amark + na + 1;
    CAC_HIP(ctx, hipMemsetAsync(ctx->amark, 0, ((size_t)na + 1) * sizeof(int), ctx->stream));
    if (ctx->nsent > 0)
      hipLaunchKernelGGL(k_mark_sent, dim3(grid_for(ctx->nsent)), dim3(kBlock), 0, ctx->stream, ctx->nsent, ctx->sent,
                         ctx->amark);
    size_t tmp_bytes = 0;
    CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, ctx->amark, pos, na + 1, ctx->stream));
    char *tmp;
    if ((rc = ensure(ctx, ctx->b_scantmp, tmp_bytes, &tmp))) return rc;
    CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, ctx->amark, pos, na + 1, ctx->stream));
    hipLaunchKernelGGL(k_compact_marks, dim3(grid_for(na)), dim3(kBlock), 0, ctx->stream, na, ctx->amark, pos,
                       ctx->aref);
    CAC_HIP(ctx, hipMemcpyAsync(&ctx->naref, pos + na, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  }
  int fl[2] = {0, 0};
  CAC_HIP(ctx, hipMemcpyAsync(fl, ctx->dflags + 11, sizeof(fl), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->cl_umax = fl[1];
  // LDS rows of 32 bytes per union entry, one wavefront per cluster: past
  // 48 KB a CU would hold too few wavefronts to hide the gathers
  if ((size_t)fl[1] * 32 > 48 * 1024) return CAC_OK;
  // launch order (one launch per pass, LDS rows for unions of at most
  // kUSmall entries): the larger unions, whose rows live in global scratch,
  // go to the start of the 8 XCDs' shares of the grid (xcd_block() gives
  // XCD x the logical blocks [x*q, (x+1)*q) in order), so each starts early
  // and runs beside the small ones; the small unions fill the rest in order.
  std::vector<int> nu(ctx->ncl), big(ctx->ncl, -1), order(ctx->ncl, -1);
  CAC_HIP(ctx, hipMemcpy(nu.data(), ctx->cl_nu, (size_t)ctx->ncl * sizeof(int), hipMemcpyDeviceToHost));
  int small_max = 0, nbig = 0;
  const int q = (ctx->ncl + 7) / 8;
  for (int c = 0; c < ctx->ncl; ++c) {
    if (nu[c] <= kUSmall) {
      small_max = std::max(small_max, nu[c]);
      continue;
    }
    int pos = (nbig % 8) * q + nbig / 8;
    if (pos >= ctx->ncl || order[pos] >= 0) pos = 0;
    while (order[pos] >= 0) ++pos;  // a free slot exists: fewer placed than ncl
    order[pos] = c;
    big[c] = nbig++;
  }
  for (int c = 0, pos = 0; c < ctx->ncl; ++c) {
    if (nu[c] > kUSmall) continue;
    while (order[pos] >= 0) ++pos;
    order[pos] = c;
  }
  ctx->cl_nbig = nbig;
  ctx->cl_umax_small = small_max;
  const size_t sstride = (size_t)4 * ctx->cl_umax + kClMax * (24 + 8);
  if ((rc = ensure(ctx, ctx->b_cl_order, (size_t)ctx->ncl, &ctx->cl_order)) ||
      (rc = ensure(ctx, ctx->b_cl_scratch, sstride * (nbig > 0 ? nbig : 1), &ctx->cl_scratch)))
    return rc;
  std::vector<int> start((size_t)ctx->ncl + 1);
  CAC_HIP(ctx, hipMemcpy(start.data(), ctx->cl_start, start.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int4> meta(ctx->ncl);
  for (int o = 0; o < ctx->ncl; ++o) {
    const int c = order[o];
    meta[o] = make_int4(c, start[c], nu[c], big[c] * 16 + (start[c + 1] - start[c]));
  }
  CAC_HIP(ctx, hipMemcpy(ctx->cl_order, meta.data(), (size_t)ctx->ncl * sizeof(int4), hipMemcpyHostToDevice));
  ctx->cl_ok = true;
  return CAC_OK;
}

// ------------------------------------------------------------ reactions
// Reverse lists of the atom-site pairs (coupling): per active site the
// atoms whose lists name it (atom | img << 27, the image the atom's list
// gives the site), ascending, so the per-step reaction sum has a fixed
// order.  The compact pair array sent (site << 32 | atom entry) is counted
// per site, the counts scanned into ranges, and the array radix-sorted:
// sorted, its low words are the ranges' entries in order.
__global__ void k_rev_count(long ns, const uint64_t *__restrict__ sent, int *__restrict__ rcount) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ns) atomicAdd(&rcount[sent[t] >> 32], 1);
}

// rkey from the sorted pairs; rcount[a] becomes the "has entries" flag.
__global__ void k_rev_keys(long ns, int na, const uint64_t *__restrict__ sorted, uint32_t *__restrict__ rkey,
                           int *__restrict__ rcount) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ns) rkey[t] = (uint32_t)sorted[t];
  if (t < na) rcount[t] = rcount[t] > 0 ? 1 : 0;
}

// Elements of the reaction sites (ascending site ids are grouped by
// element): head flags, then the element list and its ranges in rsite; and
// each reaction site's entry range in rkey.
__global__ void k_rel_heads(int nrs, const int *__restrict__ rsite, const int2 *__restrict__ asite,
                            const int *__restrict__ rstart, int *__restrict__ head, int *__restrict__ rsl,
                            int4 *__restrict__ rmeta) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > nrs) return;
  if (q == nrs) { head[q] = 0; return; }
  const int a = rsite[q];
  const int2 js = asite[a];
  rsl[q] = js.y;
  rmeta[q] = make_int4(a, rstart[a], rstart[a + 1], 0);
  head[q] = (q == 0 || js.x != asite[rsite[q - 1]].x) ? 1 : 0;
}

__global__ void k_rel_compact(int nrs, const int *__restrict__ rsite, const int2 *__restrict__ asite,
                              const int *__restrict__ head, const int *__restrict__ pos, int *__restrict__ relem,
                              int *__restrict__ rel_start) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nrs) return;
  if (head[q]) {
    relem[pos[q]] = asite[rsite[q]].x;
    rel_start[pos[q]] = q;
  }
  if (q == nrs - 1) rel_start[pos[nrs]] = nrs;
}

int launch_reaction_lists(cac_ctx *ctx) {
  ctx->nrs = ctx->nre = 0;
  ctx->nrev = 0;
  const int na = ctx->nactive;
  const long ns = ctx->nsent;
  if (!ctx->coupled || ctx->niface == 0 || na == 0 || ns == 0) return CAC_OK;
  int rc;
  if ((rc = ensure(ctx, ctx->b_rcount, (size_t)na + 1, &ctx->rcount)) ||
      (rc = ensure(ctx, ctx->b_rstart, (size_t)na + 1, &ctx->rstart)) ||
      (rc = ensure(ctx, ctx->b_rsite, (size_t)na, &ctx->rsite)) ||
      (rc = ensure(ctx, ctx->b_rkey, (size_t)ns, &ctx->rkey)) ||
      (rc = ensure(ctx, ctx->b_sent2, (size_t)ns, &ctx->sent2)))
    return rc;
  int *tmpi;
  if ((rc = ensure(ctx, ctx->b_rtmp, 2 * ((size_t)na + 1), &tmpi))) return rc;
  uint64_t *sorted = ctx->sent2;
  CAC_HIP(ctx, hipMemsetAsync(ctx->rcount, 0, ((size_t)na + 1) * sizeof(int), ctx->stream));
  hipLaunchKernelGGL(k_rev_count, dim3(grid_for(ns)), dim3(kBlock), 0, ctx->stream, ns, ctx->sent, ctx->rcount);
  int bits = 32;
  while ((1L << (bits - 32)) < na) ++bits;
  size_t tmp_bytes = 0, sort_bytes = 0;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, ctx->rcount, ctx->rstart, na + 1, ctx->stream));
  CAC_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, ctx->sent, sorted, (int)ns, 0, bits,
                                                 ctx->stream));
  char *tmp;
  if ((rc = ensure(ctx, ctx->b_scantmp, std::max(tmp_bytes, sort_bytes), &tmp))) return rc;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, ctx->rcount, ctx->rstart, na + 1, ctx->stream));
  CAC_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(tmp, sort_bytes, ctx->sent, sorted, (int)ns, 0, bits, ctx->stream));
  ctx->nrev = ns;
  hipLaunchKernelGGL(k_rev_keys, dim3(grid_for(std::max(ns, (long)na))), dim3(kBlock), 0, ctx->stream, ns, na, sorted,
                     ctx->rkey, ctx->rcount);
  int *pos = tmpi;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, ctx->rcount, pos, na + 1, ctx->stream));
  hipLaunchKernelGGL(k_compact_marks, dim3(grid_for(na)), dim3(kBlock), 0, ctx->stream, na, ctx->rcount, pos,
                     ctx->rsite);
  int nrs = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&nrs, pos + na, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (nrs == 0) return CAC_OK;
  if ((rc = ensure(ctx, ctx->b_relem, (size_t)nrs, &ctx->relem)) ||
      (rc = ensure(ctx, ctx->b_rel_start, (size_t)nrs + 1, &ctx->rel_start)) ||
      (rc = ensure(ctx, ctx->b_rforce, 3 * (size_t)nrs, &ctx->rforce)) ||
      (rc = ensure(ctx, ctx->b_rsl, (size_t)nrs, &ctx->rsl)) ||
      (rc = ensure(ctx, ctx->b_rmeta, (size_t)nrs, &ctx->rmeta)))
    return rc;
  // head flags [nrs + 1] in tmpi[0 ..], their exclusive sums after them
  int *head = tmpi, *hpos = tmpi + (nrs + 1);
  hipLaunchKernelGGL(k_rel_heads, dim3(grid_for((long)nrs + 1)), dim3(kBlock), 0, ctx->stream, nrs, ctx->rsite,
                     ctx->asite, ctx->rstart, head, ctx->rsl, ctx->rmeta);
  tmp_bytes = 0;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, head, hpos, nrs + 1, ctx->stream));
  if ((rc = ensure(ctx, ctx->b_scantmp, tmp_bytes, &tmp))) return rc;
  CAC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, head, hpos, nrs + 1, ctx->stream));
  hipLaunchKernelGGL(k_rel_compact, dim3(grid_for(nrs)), dim3(kBlock), 0, ctx->stream, nrs, ctx->rsite, ctx->asite,
                     head, hpos, ctx->relem, ctx->rel_start);
  int nre = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&nre, hpos + nrs, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  CAC_HIP(ctx, hipGetLastError());
  ctx->nrs = nrs;
  ctx->nre = nre;
  return CAC_OK;
}

int launch_check_displacement(cac_ctx *ctx, double triggersq) {
  long tot = (long)ctx->n + 8L * ctx->ne;
  if (tot == 0) return CAC_OK;
  hipLaunchKernelGGL(k_check_disp, dim3(grid_for(tot)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->x4, ctx->xhold,
                     ctx->ne, ctx->nodex, ctx->nodexhold, triggersq, ctx->dflags);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

}  // namespace cac
