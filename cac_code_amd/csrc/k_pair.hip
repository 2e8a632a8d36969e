// k_pair.hip -- EAM and LJ force passes over atoms and element integration
// points, and the integration-point -> node scatter.
//
// FP64 path: the reference's arithmetic term by term (pair_eam.cpp:584-746,
// 927-1137; pair_lj_cut.cpp:227-288, 295-442).  Unlike the neighbor build,
// this file lets the compiler fuse multiplies and adds (fp contract fast;
// the force pass's VALU work falls by about a tenth), except where bits must
// match a stored table (the knot-form spline coefficients).  Results differ
// from the CPU oracle through the order of the neighbor sums and those
// fused roundings, both far below the 1e-10 parity bar.
// MIXED path (P = float): r, the spline polynomials, F'(rho) and the pair
// force are evaluated in FP32; positions, site interpolation, rho, force,
// energy and virial sums stay FP64.
//
// Work mapping: one thread per atom (atom lists in lattice order, see the
// atom section); 8 lanes per node / integration point, lane l taking list
// entries l, l+8, ... so the 64 lanes of a wavefront read 64 consecutive
// entries per iteration, partial sums combined with a fixed xor butterfly
// (every lane ends with the same bits).  Results are reproducible run to
// run.
//
// Element path: the sites that some list refers to (the "active sites",
// numbered at rebuild) are interpolated once per step (k_site_pos:
// x = sum_k N_k X_k; k_site_fp: rho = sum_k N_k erho_k and F'(rho)) into the
// rows of x4 after the atoms, so every pass gathers 32-byte (x, y, z, fp)
// rows of one particle array whether the neighbor is an atom or a site (the
// coupled lists mix them).  The image of a neighbor is x_j + s*prd.
//
// Tables on the device: per r-knot, the density pass reads the 4 value
// coefficients of rhor (32 bytes, 16 in FP32) and the force pass the 4 value
// coefficients of rhor and of z2r (64 bytes); the derivative polynomial is
// formed from them, (3*c3*p + 2*c4)*p + c5 times 1/dr, which is the
// reference's (c0*p + c1)*p + c2 with c0 = 3*c3/dr, c1 = 2*c4/dr,
// c2 = c5/dr (pair_eam.cpp:574-578).  frho rows keep all 7 coefficients.
#include <type_traits>

#include "cac_internal.h"

#pragma clang fp contract(fast)

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

struct Prd {
  double v[3];
};

template <class P>
struct Tabs {
  const P *rv;  // [(nr+1)][4]  rhor c3..c6
  const P *rz;  // [(nr+1)][8]  rhor c3..c6, z2r c3..c6
  const P *fr;  // [(nrho+1)][8]
  const double *rk;  // [(nr+1)][4] knot form (FP64 force passes), null = rows
  int nr, nrho;
  P rdr, rdrho;
  double cutforcesq, rhomax;
};

// r-table knot and fraction, pair_eam.cpp:649-653
template <class P>
__device__ __forceinline__ int rknot(const Tabs<P> &t, P r, P *pout) {
  P p = r * t.rdr + P(1.0);
  int m = (int)p;
  m = m < t.nr - 1 ? m : t.nr - 1;
  p -= (P)m;
  p = p < P(1.0) ? p : P(1.0);
  *pout = p;
  return m;
}

// rho-table row and fraction, pair_eam.cpp:672-676
template <class P>
__device__ __forceinline__ const P *rhorow(const Tabs<P> &t, P rho, P *pout) {
  P p = rho * t.rdrho + P(1.0);
  int m = (int)p;
  m = m < t.nrho - 1 ? m : t.nrho - 1;
  m = m > 1 ? m : 1;
  p -= (P)m;
  p = p < P(1.0) ? p : P(1.0);
  *pout = p;
  return t.fr + 8 * (size_t)m;
}

// Table window in LDS: rows m0 .. nr-1 of a per-knot table (stride P
// values per row) copied into shared memory at kernel start; rows below m0
// (short pair distances, rarely reached) are read from the global table.
// Gathering coefficient rows through LDS instead of the vector L1 keeps the
// force passes off the texture-address path, which 64 distinct rows per
// wavefront instruction otherwise saturate.
template <class P>
using lds_cptr = const __attribute__((address_space(3))) P *;

template <class P>
struct Win {
  lds_cptr<P> lds;
  int m0;
};

template <class P>
__device__ __forceinline__ Win<P> make_win(const unsigned char *smem, int m0) {
  return Win<P>{(lds_cptr<P>)(const __attribute__((address_space(3))) unsigned char *)smem, m0};
}

template <class P>
__device__ void load_window(const P *__restrict__ table, int stride, int m0, int nr, P *lds) {
  const int units = (nr - m0) * stride * (int)sizeof(P) / 16;
  const uint4 *src = reinterpret_cast<const uint4 *>(table + (size_t)m0 * stride);
  uint4 *dst = reinterpret_cast<uint4 *>(lds);
  for (int q = (int)threadIdx.x; q < units; q += (int)blockDim.x) dst[q] = src[q];
}

// Row copies, one per address space: the two sources are read in separate
// branches so the LDS side compiles to ds_read and the global side to
// global_load (a pointer select between them would force flat loads, which
// go through the texture-address path like any global gather).
template <class P, int K>
struct Row {
  P c[K];
};
template <class P, int K>
__device__ __forceinline__ Row<P, K> fetch_row(const Tabs<P> &t, const P *__restrict__ gtab, const Win<P> &w, int m) {
  Row<P, K> out;
  if (m >= w.m0) {
    lds_cptr<P> q = w.lds + K * (m - w.m0);
#pragma unroll
    for (int k = 0; k < K; ++k) out.c[k] = q[k];
  } else {
    const P *q = gtab + K * (size_t)m;
#pragma unroll
    for (int k = 0; k < K; ++k) out.c[k] = q[k];
  }
  return out;
}

// rho(r) = ((c3*p + c4)*p + c5)*p + c6, pair_eam.cpp:655
template <class P>
__device__ __forceinline__ P rho_of_r_w(const Tabs<P> &t, const Win<P> &w, P r) {
  P p;
  const int m = rknot<P>(t, r, &p);
  const Row<P, 4> R = fetch_row<P, 4>(t, t.rv, w, m);
  const P *c = R.c;
  return ((c[0] * p + c[1]) * p + c[2]) * p + c[3];
}

// rho(r) and rho'(r), z2'(r), z2(r) from the global rows (integration-point
// passes), pair_eam.cpp:655, 720-726
template <class P>
__device__ __forceinline__ P rho_of_r(const Tabs<P> &t, P r) {
  P p;
  const P *c = t.rv + 4 * (size_t)rknot<P>(t, r, &p);
  return ((c[0] * p + c[1]) * p + c[2]) * p + c[3];
}

template <class P>
__device__ __forceinline__ void rz_of_r(const Tabs<P> &t, P r, P *rhoip, P *z2p, P *z2) {
  P p;
  const P *c = t.rz + 8 * (size_t)rknot<P>(t, r, &p);
  *rhoip = ((P(3.0) * c[0] * p + P(2.0) * c[1]) * p + c[2]) * t.rdr;
  *z2p = ((P(3.0) * c[4] * p + P(2.0) * c[5]) * p + c[6]) * t.rdr;
  *z2 = ((c[4] * p + c[5]) * p + c[6]) * p + c[7];
}

// rho'(r), z2'(r), z2(r) from the knot form: row m's c3, c4 of rhor and z2r
// recomputed from knots m and m+1 with array2spline's operations
// (pair_eam.cpp:538-541; same bits as the stored rows), c5 = slope,
// c6 = value.  32 bytes per knot instead of 64 per row, so an LDS window
// covers twice the r range.
__device__ __forceinline__ void rz_of_r_knots(const Tabs<double> &t, const Win<double> &w, double r, double *rhoip,
                                              double *z2p, double *z2) {
  double p;
  const int m = rknot<double>(t, r, &p);
  double k[8];
  if (m >= w.m0) {
    lds_cptr<double> q = w.lds + 4 * (m - w.m0);
#pragma unroll
    for (int c = 0; c < 8; ++c) k[c] = q[c];
  } else {
    const double *q = t.rk + 4 * (size_t)m;
#pragma unroll
    for (int c = 0; c < 8; ++c) k[c] = q[c];
  }
  double r4, r3, z4, z3;
  {  // array2spline's operations, unfused: the stored c3, c4 bit for bit
#pragma clang fp contract(off)
    const double dr = k[4] - k[0], dz = k[6] - k[2];
    r4 = 3.0 * dr - 2.0 * k[1] - k[5];
    r3 = k[1] + k[5] - 2.0 * dr;
    z4 = 3.0 * dz - 2.0 * k[3] - k[7];
    z3 = k[3] + k[7] - 2.0 * dz;
  }
  *rhoip = ((3.0 * r3 * p + 2.0 * r4) * p + k[1]) * t.rdr;
  *z2p = ((3.0 * z3 * p + 2.0 * z4) * p + k[3]) * t.rdr;
  *z2 = ((z3 * p + z4) * p + k[3]) * p + k[2];
}

// rho'(r), z2'(r), z2(r) of one pair, pair_eam.cpp:720-726 (knot form in
// FP64 when the table has it, stored rows otherwise)
template <class P>
__device__ __forceinline__ void rz_of_r_w(const Tabs<P> &t, const Win<P> &w, P r, P *rhoip, P *z2p, P *z2) {
  if constexpr (std::is_same<P, double>::value) {
    if (t.rk) {
      rz_of_r_knots(t, w, r, rhoip, z2p, z2);
      return;
    }
  }
  P p;
  const int m = rknot<P>(t, r, &p);
  const Row<P, 8> R = fetch_row<P, 8>(t, t.rz, w, m);
  const P *c = R.c;
  *rhoip = ((P(3.0) * c[0] * p + P(2.0) * c[1]) * p + c[2]) * t.rdr;
  *z2p = ((P(3.0) * c[4] * p + P(2.0) * c[5]) * p + c[6]) * t.rdr;
  *z2 = ((c[4] * p + c[5]) * p + c[6]) * p + c[7];
}

template <class P>
__device__ __forceinline__ P embed_fp(const Tabs<P> &t, double rho) {
  P p;
  const P *c = rhorow<P>(t, (P)rho, &p);
  return (c[0] * p + c[1]) * p + c[2];
}

// Embedding energy with the linear extension past rhomax (build-defined; the
// reference computes no energy).
template <class P>
__device__ __forceinline__ double embed_energy(const Tabs<P> &t, double rho, double fp) {
  P p;
  const P *c = rhorow<P>(t, (P)rho, &p);
  double e = (double)(((c[3] * p + c[4]) * p + c[5]) * p + c[6]);
  if (rho > t.rhomax) e += fp * (rho - t.rhomax);
  return e;
}

// r and 1/r of a pair from r^2: the hardware reciprocal square root refined
// by Newton-Raphson steps for 1/sqrt(x), y <- y + y*(1 - x*y^2)/2 (two in
// FP64, one in FP32; each step squares the relative error), then r = x*y.
// Within a few ulps of sqrt and 1/sqrt, for a fraction of the instructions
// of a correctly rounded square root and division.  Pair distances here are
// far from zero, overflow and denormals.
template <class P>
struct RInvP {
  P r, rinv;
};
template <class P>
__device__ __forceinline__ RInvP<P> r_and_inv(double rsq);
template <>
__device__ __forceinline__ RInvP<double> r_and_inv<double>(double rsq) {
  double y = __builtin_amdgcn_rsq(rsq);
  double e = __builtin_fma(-(rsq * y), y, 1.0);
  y = __builtin_fma(0.5 * y, e, y);
  e = __builtin_fma(-(rsq * y), y, 1.0);
  y = __builtin_fma(0.5 * y, e, y);
  return RInvP<double>{rsq * y, y};
}
template <>
__device__ __forceinline__ RInvP<float> r_and_inv<float>(double rsq) {
  const float x = (float)rsq;
  float y = __builtin_amdgcn_rsqf(x);
  const float e = __builtin_fmaf(-(x * y), y, 1.0f);
  y = __builtin_fmaf(0.5f * y, e, y);
  return RInvP<float>{x * y, y};
}
template <class P>
__device__ __forceinline__ P sqrt_p(double rsq) {
  return r_and_inv<P>(rsq).r;
}

// Sum over the 8 lanes of a group; the xor butterfly pairs the same partial
// sums on every lane, so all lanes hold identical bits.
__device__ __forceinline__ double group_sum(double v) {
  v += __shfl_xor(v, 1, kLanes);
  v += __shfl_xor(v, 2, kLanes);
  v += __shfl_xor(v, 4, kLanes);
  return v;
}

__device__ __forceinline__ const uint32_t *list_row(const uint32_t *base, int r, int M) {
  return base + (size_t)(r >> 3) * M * kRowGroup + (r & 7);
}

struct Pt {
  double x, y, z;
};

// Entries per lane fetched together: the list words and then the neighbor
// gathers of kU entries are issued back to back before any arithmetic, so a
// lane keeps several memory requests in flight instead of one dependent chain.
constexpr int kU = 4;
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;  // image field 31: never a real entry

// List words are streamed once per pass: non-temporal loads keep them from
// displacing the reused neighbor rows in the caches.
__device__ __forceinline__ uint32_t get_entry(const uint32_t *p) { return __builtin_nontemporal_load(p); }

template <int U>
__device__ __forceinline__ void fetch_entries(const uint32_t *__restrict__ row, int k0, int jn, uint32_t e[U]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int k = k0 + u * kLanes;
    e[u] = k < jn ? get_entry(row + (size_t)k * kRowGroup) : kNoEntry;
  }
}

// Position of list entry e = (j | img << 27): x_j, or x_j + s*prd for a
// periodic image.
__device__ __forceinline__ Pt load_xyz(const double *__restrict__ x4, uint32_t e, const Prd &prd) {
  const size_t j = e & 0x7FFFFFFu;
  const int img = (int)(e >> 27);
  const double2 xy = reinterpret_cast<const double2 *>(x4)[2 * j];
  const double z = x4[4 * j + 2];
  if (img == kImgSelf) return Pt{xy.x, xy.y, z};
  int s[3];
  img_shift(img, s);
  return Pt{shifted(xy.x, s[0], prd.v[0]), shifted(xy.y, s[1], prd.v[1]), shifted(z, s[2], prd.v[2])};
}

// position and fourth slot (fp) of list entry e
__device__ __forceinline__ Pt load_xyzw(const double *__restrict__ x4, uint32_t e, const Prd &prd, double *w) {
  const size_t j = e & 0x7FFFFFFu;
  const int img = (int)(e >> 27);
  const double2 xy = reinterpret_cast<const double2 *>(x4)[2 * j];
  const double2 zw = reinterpret_cast<const double2 *>(x4)[2 * j + 1];
  *w = zw.y;
  if (img == kImgSelf) return Pt{xy.x, xy.y, zw.x};
  int s[3];
  img_shift(img, s);
  return Pt{shifted(xy.x, s[0], prd.v[0]), shifted(xy.y, s[1], prd.v[1]), shifted(zw.x, s[2], prd.v[2])};
}

struct Virial {
  double v[6];
  __device__ void zero() {
    for (int c = 0; c < 6; ++c) v[c] = 0.0;
  }
  // share: the part of the pair's virial this side tallies (1/2, or all of
  // it for an atom-site pair, whose site side tallies nothing)
  __device__ void add(double dx, double dy, double dz, double fpair, double share = 0.5) {
    v[0] += share * dx * dx * fpair;
    v[1] += share * dy * dy * fpair;
    v[2] += share * dz * dz * fpair;
    v[3] += share * dx * dy * fpair;
    v[4] += share * dx * dz * fpair;
    v[5] += share * dy * dz * fpair;
  }
  __device__ void reduce() {
    for (int c = 0; c < 6; ++c) v[c] = group_sum(v[c]);
  }
  __device__ void store(double *out, size_t stride, size_t idx, double w) const {
    for (int c = 0; c < 6; ++c) out[c * stride + idx] = w * v[c];
  }
};

// EAM pair terms of the atom force pass, pair_eam.cpp:706-738 (the
// integration-point passes evaluate the same terms, :1083-1120).
template <class P>
__device__ __forceinline__ double eam_fpair(const Tabs<P> &t, const Win<P> &w, double rsq, P fpi, P fpj,
                                            P *phi_out) {
  const RInvP<P> rr = r_and_inv<P>(rsq);
  P r = rr.r, rhoip, z2p, z2;
  rz_of_r_w<P>(t, w, r, &rhoip, &z2p, &z2);
  P recip = rr.rinv;
  P phi = z2 * recip;
  P phip = z2p * recip - phi * recip;
  P psip = fpi * rhoip + fpj * rhoip + phip;
  *phi_out = phi;
  return (double)(-psip * recip);
}

// The same from the global rows (the same bits as the knot form), for the
// reactions, which have no LDS window: the rows need no coefficient
// recompute.
template <class P>
__device__ __forceinline__ double eam_fpair_global(const Tabs<P> &t, double rsq, P fpi, P fpj, P *phi_out) {
  const RInvP<P> rr = r_and_inv<P>(rsq);
  P rhoip, z2p, z2;
  rz_of_r<P>(t, rr.r, &rhoip, &z2p, &z2);
  P recip = rr.rinv;
  P phi = z2 * recip;
  P phip = z2p * recip - phi * recip;
  P psip = fpi * rhoip + fpj * rhoip + phip;
  *phi_out = phi;
  return (double)(-psip * recip);
}

// ------------------------------------------------------------- atoms
// One thread per atom.  Atom lists are in the atom_slot layout and every row
// is in lattice order, so the 64 atoms of a wavefront (consecutive atoms,
// spatially close) read their k-th entries from one 256-byte run and mostly
// gather neighboring x4 rows.  Each atom sums its entries in list order, so
// results are reproducible run to run.  kUA entries are fetched (list words,
// then neighbor rows) before any arithmetic.
constexpr int kUA = 8;

__host__ __device__ __forceinline__ size_t mask_slot(int i, int w, int W) {
  return ((size_t)(i >> 6) * W + w) * kAtomGroup + (i & 63);
}

// One past the last set bit of atom i's mask words over its jn entries (0:
// none).  The rows run shell by shell (k_atom_neigh's key), so the entries
// a pass needs fill the front of the row and the pass stops here instead
// of walking the skin shells to the end.
__device__ __forceinline__ int mask_end(const uint32_t *__restrict__ m, int i, int W, int jn) {
  int end = 0;
  const int nw = (jn + 31) >> 5;
#pragma unroll 4
  for (int w = 0; w < nw; ++w) {
    const uint32_t v = m[mask_slot(i, w, W)];
    if (v) end = 32 * w + 32 - __clz(v);
  }
  return end;
}

// U entries from k0 (a multiple of 4): one 16-byte load per run of 4 (the
// run lies inside the row, whose length is a multiple of 8; entries past jn
// are dropped)
template <int U>
__device__ __forceinline__ void fetch_atom_entries(const uint32_t *__restrict__ row, int k0, int jn, uint32_t e[U]) {
  static_assert(U % kAtomRun == 0, "whole runs");
#pragma unroll
  for (int c = 0; c < U / kAtomRun; ++c) {
    const int k = k0 + kAtomRun * c;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 v = {kNoEntry, kNoEntry, kNoEntry, kNoEntry};
    if (k < jn) v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(row + atom_off(k)));
    e[kAtomRun * c] = v.x;
    e[kAtomRun * c + 1] = k + 1 < jn ? v.y : kNoEntry;
    e[kAtomRun * c + 2] = k + 2 < jn ? v.z : kNoEntry;
    e[kAtomRun * c + 3] = k + 3 < jn ? v.w : kNoEntry;
  }
}

// rho pass + fp pass fused (pair_eam.cpp:629-680); fp stored in x4.w, which
// the force pass reads for ghosts too (forward_comm_pair, :683).
//
// In-range mask: the density pass also records which entries have
// r^2 < cutforcesq, one bit per entry (word w of atom i at
// mask_slot(i, w, W), W = maxneigh/32 words), and the force pass gathers
// neighbor rows only for those.  Positions do not change between the
// density and force passes of one force call, so the force pass evaluates
// exactly the pairs the full list would give it, in list order (same
// bits), while skipping the row gathers of the skin shell (about 46% of the
// entries at 5.95 A vs 4.95 A).
//
// Inner mask (a second, shorter Verlet list inside the neighbor list): a
// full density pass (*ibad != 0) gathers every entry and also records the
// entries with r^2 < cutinsq = (cutforce + margin)^2 and every element site
// (which moves with its nodes), and the positions (xhold_in).  While no atom
// has moved more than margin/2 since, no atom pair outside that set can be
// within cutforce, so the next density passes (*ibad == 0) gather only the
// entries of that set: the same pairs in the same order, so the same bits.
// Both passes stop after the last entry their mask keeps (mask_end).
//
// The table window fills a CU's LDS, so one workgroup runs per CU at a time.
// A large system runs one thread per atom (many rounds of workgroups; the
// block order keeps each XCD's workgroups on neighboring atoms).  When the
// atoms fill fewer than two rounds, the last round would leave CUs idle:
// then one workgroup per CU takes an equal contiguous range of `per` atoms
// (a multiple of 64) in strides of kAtomWG.
constexpr int kAtomWG = 1024;

struct AtomGrid {
  unsigned blocks;
  int per;
};

AtomGrid atom_grid(const cac_ctx *ctx, int n) {
  AtomGrid g = {0, 0};
  if (n <= 0) return g;  // no atom launches
  long nb = ((long)n + kAtomWG - 1) / kAtomWG;
  if (nb <= 2L * ctx->ncu && nb > ctx->ncu) nb = ctx->ncu;
  nb = ((nb + 7) / 8) * 8;  // xcd_block() needs a multiple of 8
  g.blocks = (unsigned)nb;
  g.per = (int)((((long)n + nb - 1) / nb + 63) / 64 * 64);
  return g;
}

template <class P, bool EV>
__global__ __launch_bounds__(kAtomWG) void k_eam_atom_rho(int n, double *__restrict__ x4, Prd prd,
                                                          const uint32_t *__restrict__ neigh, int M,
                                                          const int *__restrict__ num, Tabs<P> t, int m0,
                                                          double *__restrict__ rho_out, double *__restrict__ eatom,
                                                          uint32_t *__restrict__ mask, const int *__restrict__ ibad,
                                                          uint32_t *__restrict__ mask_in,
                                                          double *__restrict__ xhold_in, double cutinsq, int per,
                                                          const int *__restrict__ skip) {
  if (*skip) return;  // this step rebuilds first: the caller discards these results
  extern __shared__ __align__(16) unsigned char smem[];
  const Win<P> w = make_win<P>(smem, m0);
  load_window<P>(t.rv, 4, m0, t.nr, reinterpret_cast<P *>(smem));
  __syncthreads();
  const bool inner = *ibad == 0;
  const int i0 = xcd_block() * per, i1 = min(n, i0 + per);
  for (int i = i0 + (int)threadIdx.x; i < i1; i += kAtomWG) {
    const double xi = x4[4 * (size_t)i], yi = x4[4 * (size_t)i + 1], zi = x4[4 * (size_t)i + 2];
    const int jn = num[i];
    const uint32_t *row = neigh + atom_slot(i, 0, M);
    const int W = (M + 31) >> 5;
    uint32_t bits = 0, bits_in = 0;
    double rho = 0.0;
    // software pipeline: the list words (and inner-mask word) of chunk
    // k0 + kUA are loaded while the neighbor rows of chunk k0 are in flight
    uint32_t e[kUA];
    fetch_atom_entries<kUA>(row, 0, jn, e);
    uint32_t in_w = (inner && jn > 0) ? mask_in[mask_slot(i, 0, W)] : 0u;
    const int kend = inner ? min(jn, mask_end(mask_in, i, W, jn)) : jn;
    for (int k0 = 0; k0 < kend; k0 += kUA) {
      if (inner && (k0 & 31) == 0) bits_in = in_w;
      if (inner) {  // element sites (rows >= n, moved by their nodes) are always gathered
        const uint32_t sel = bits_in >> (k0 & 31);
  #pragma unroll
        for (int u = 0; u < kUA; ++u)
          if (!((sel >> u) & 1u) && (e[u] & 0x7FFFFFFu) < (uint32_t)n) e[u] = kNoEntry;
      }
      Pt xj[kUA];
  #pragma unroll
      for (int u = 0; u < kUA; ++u)
        if (e[u] != kNoEntry) xj[u] = load_xyz(x4, e[u], prd);
      __asm__ volatile("" ::: "memory");  // keep the next chunk's loads behind this chunk's gathers
      const int k1 = k0 + kUA;
      uint32_t en[kUA];
      fetch_atom_entries<kUA>(row, k1, jn, en);
      if (inner && (k1 & 31) == 0 && k1 < jn) in_w = mask_in[mask_slot(i, k1 >> 5, W)];
  #pragma unroll
      for (int u = 0; u < kUA; ++u) {
        if (e[u] == kNoEntry) continue;
        const double delx = xi - xj[u].x, dely = yi - xj[u].y, delz = zi - xj[u].z;
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t.cutforcesq) {
          rho += (double)rho_of_r_w<P>(t, w, sqrt_p<P>(rsq));
          bits |= 1u << ((k0 + u) & 31);
        }
        if (!inner && (rsq < cutinsq || (e[u] & 0x7FFFFFFu) >= (uint32_t)n)) bits_in |= 1u << ((k0 + u) & 31);
      }
      if (((k0 + kUA) & 31) == 0 || k0 + kUA >= kend) {
        mask[mask_slot(i, k0 >> 5, W)] = bits;
        if (!inner) mask_in[mask_slot(i, k0 >> 5, W)] = bits_in;
        bits = 0;
        bits_in = 0;
      }
  #pragma unroll
      for (int u = 0; u < kUA; ++u) e[u] = en[u];
    }
    // the words past the walk keep no entry
    for (int wd = (kend + 31) >> 5; wd < ((jn + 31) >> 5); ++wd) mask[mask_slot(i, wd, W)] = 0u;
    if (!inner) {
      xhold_in[i] = xi;
      xhold_in[n + i] = yi;
      xhold_in[2 * (size_t)n + i] = zi;
    }
    rho_out[i] = rho;
    const double fp = (double)embed_fp<P>(t, rho);
    x4[4 * (size_t)i + 3] = fp;
    if (EV) eatom[i] = embed_energy<P>(t, rho, fp);
  }
}

// force pass without Newton's third law, pair_eam.cpp:688-745, gathering
// only the entries the density pass marked in range
template <class P, bool EV>
__global__ __launch_bounds__(kAtomWG) void k_eam_atom_force(int n, const double *__restrict__ x4, Prd prd,
                                                            const uint32_t *__restrict__ neigh, int M,
                                                            const int *__restrict__ num, Tabs<P> t, int m0,
                                                            double *__restrict__ f, double *__restrict__ eatom,
                                                            double *__restrict__ vatom,
                                                            const uint32_t *__restrict__ mask, int per,
                                                            int *__restrict__ ibad, const int *__restrict__ skip) {
  if (*skip) return;  // this step rebuilds first: the caller discards these results
  extern __shared__ __align__(16) unsigned char smem[];
  // the density pass has read the inner-mask flag: its inner mask now holds
  // for these positions (see k_eam_atom_rho)
  if (blockIdx.x == 0 && threadIdx.x == 0) *ibad = 0;
  const Win<P> w = make_win<P>(smem, m0);
  if (t.rk)
    load_window<double>(t.rk, 4, m0, t.nr + 1, reinterpret_cast<double *>(smem));
  else
    load_window<P>(t.rz, 8, m0, t.nr, reinterpret_cast<P *>(smem));
  __syncthreads();
  const int i0 = xcd_block() * per, i1 = min(n, i0 + per);
  for (int i = i0 + (int)threadIdx.x; i < i1; i += kAtomWG) {
    double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
    Virial vir;
    vir.zero();
    const double xi = x4[4 * (size_t)i], yi = x4[4 * (size_t)i + 1], zi = x4[4 * (size_t)i + 2];
    const P fpi = (P)x4[4 * (size_t)i + 3];
    const int jn = num[i];
    const uint32_t *row = neigh + atom_slot(i, 0, M);
    const int W = (M + 31) >> 5;
    // software pipeline: the list words and mask word of chunk k0 + kUA are
    // loaded while the neighbor rows of chunk k0 are in flight
    uint32_t e[kUA];
    fetch_atom_entries<kUA>(row, 0, jn, e);
    uint32_t bits = jn > 0 ? mask[mask_slot(i, 0, W)] : 0u;
    const int kend = min(jn, mask_end(mask, i, W, jn));
    for (int k0 = 0; k0 < kend; k0 += kUA) {
      const uint32_t sel = bits >> (k0 & 31);
  #pragma unroll
      for (int u = 0; u < kUA; ++u)
        if (!((sel >> u) & 1u)) e[u] = kNoEntry;
      Pt xjv[kUA];
      double fpjv[kUA];
  #pragma unroll
      for (int u = 0; u < kUA; ++u)
        if (e[u] != kNoEntry) xjv[u] = load_xyzw(x4, e[u], prd, &fpjv[u]);
      __asm__ volatile("" ::: "memory");  // keep the next chunk's loads behind this chunk's gathers
      const int k1 = k0 + kUA;
      uint32_t en[kUA];
      fetch_atom_entries<kUA>(row, k1, jn, en);
      const uint32_t bits_n = ((k1 & 31) == 0 && k1 < jn) ? mask[mask_slot(i, k1 >> 5, W)] : bits;
  #pragma unroll
      for (int u = 0; u < kUA; ++u) {
        if (e[u] == kNoEntry) continue;
        const double delx = xi - xjv[u].x, dely = yi - xjv[u].y, delz = zi - xjv[u].z;
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t.cutforcesq) {
          P phi;
          const double fpair = eam_fpair<P>(t, w, rsq, fpi, (P)fpjv[u], &phi);
          fx += delx * fpair;
          fy += dely * fpair;
          fz += delz * fpair;
          if (EV) {  // an atom-site pair's whole energy and virial go to the atom
            const double share = (e[u] & 0x7FFFFFFu) < (uint32_t)n ? 0.5 : 1.0;
            ep += share * (double)phi;
            vir.add(delx, dely, delz, fpair, share);
          }
        }
      }
  #pragma unroll
      for (int u = 0; u < kUA; ++u) e[u] = en[u];
      bits = bits_n;
    }
    f[i] = fx;
    f[(size_t)n + i] = fy;
    f[2 * (size_t)n + i] = fz;
    if (EV) {
      eatom[i] += ep;
      vir.store(vatom, (size_t)n, (size_t)i, 1.0);
    }
  }
}

// ------------------------------------------------------------- atoms, LJ

// PairLJCut::compute, pair_lj_cut.cpp:255-287
template <class P, bool EV>
__global__ __launch_bounds__(kBlock) void k_lj_atom_force(int n, const double *__restrict__ x4, Prd prd,
                                                          const uint32_t *__restrict__ neigh, int M,
                                                          const int *__restrict__ num, LjParams lj,
                                                          double *__restrict__ f, double *__restrict__ eatom,
                                                          double *__restrict__ vatom) {
  const int i = xcd_block() * kBlock + (int)threadIdx.x;
  if (i >= n) return;
  const P lj1 = (P)lj.lj1, lj2 = (P)lj.lj2, lj3 = (P)lj.lj3, lj4 = (P)lj.lj4;
  double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
  Virial vir;
  vir.zero();
  const double xi = x4[4 * (size_t)i], yi = x4[4 * (size_t)i + 1], zi = x4[4 * (size_t)i + 2];
  const int jn = num[i];
  const uint32_t *row = neigh + atom_slot(i, 0, M);
  for (int k0 = 0; k0 < jn; k0 += kUA) {
    uint32_t e[kUA];
    fetch_atom_entries<kUA>(row, k0, jn, e);
    Pt xjv[kUA];
#pragma unroll
    for (int u = 0; u < kUA; ++u)
      if (e[u] != kNoEntry) xjv[u] = load_xyz(x4, e[u], prd);
#pragma unroll
    for (int u = 0; u < kUA; ++u) {
      if (e[u] == kNoEntry) continue;
      const double delx = xi - xjv[u].x, dely = yi - xjv[u].y, delz = zi - xjv[u].z;
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < lj.cutsq) {
        P r2inv = P(1.0) / (P)rsq;
        P r6inv = r2inv * r2inv * r2inv;
        P forcelj = r6inv * (lj1 * r6inv - lj2);
        double fpair = (double)(P(1.0) * forcelj * r2inv);
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        if (EV) {  // an atom-site pair's whole energy and virial go to the atom
          const double share = (e[u] & 0x7FFFFFFu) < (uint32_t)n ? 0.5 : 1.0;
          ep += share * (double)(r6inv * (lj3 * r6inv - lj4));
          vir.add(delx, dely, delz, fpair, share);
        }
      }
    }
  }
  f[i] = fx;
  f[(size_t)n + i] = fy;
  f[2 * (size_t)n + i] = fz;
  if (EV) {
    eatom[i] = ep;
    vir.store(vatom, (size_t)n, (size_t)i, 1.0);
  }
}

// ------------------------------------------------------------- elements
struct ElemArgs {
  int ne, nint, nip, nactive, M;
  int natoms;                // particle rows below natoms are atoms
  const double *nodex;       // [ne][8][3]
  const double *spa_site;    // [nsite][8]
  const double *spa_int;     // [nint][8]
  const double *weight;      // [nint]
  const uint32_t *ipneigh;   // blocked, entries (active id | img << 27)
  const int *ipnum;          // [nip]
  const int2 *asite;         // [nactive] (element, site)
  int node_list[8];
  int nae[3];
  double pstep[3];           // parent-coordinate step 2/(nae-1) per direction
  // optional subsets (surrogate mode): elements for the erho / force /
  // scatter passes, active sites for the site passes; null = all
  const int *elist;
  int nelist;
  const int *slist;
  int nslist;
  __device__ __forceinline__ int n_elem() const { return elist ? nelist : ne; }
  __device__ __forceinline__ int elem(int o) const { return elist ? elist[o] : o; }
  __device__ __forceinline__ int n_site() const { return slist ? nslist : nactive; }
  __device__ __forceinline__ int site(int o) const { return slist ? slist[o] : o; }
};

// Active-site positions: sum_k N_k(site) X_k in node order
// (pair_eam.cpp:1071-1076, element.cpp:1251-1271).
__global__ __launch_bounds__(kBlock) void k_site_pos(ElemArgs a, double *__restrict__ xs4) {
  const int o = xcd_block() * kBlock + (int)threadIdx.x;
  if (o >= a.n_site()) return;
  const int q = a.site(o);
  const int2 js = a.asite[q];
  const double *X = a.nodex + 24 * (size_t)js.x;
  const double *N = a.spa_site + 8 * (size_t)js.y;
  double x = 0.0, y = 0.0, z = 0.0;
  for (int k = 0; k < 8; ++k) {
    x += N[k] * X[3 * k];
    y += N[k] * X[3 * k + 1];
    z += N[k] * X[3 * k + 2];
  }
  reinterpret_cast<double2 *>(xs4)[2 * (size_t)q] = make_double2(x, y);
  xs4[4 * (size_t)q + 2] = z;
}

// Active-site density and embedding derivative: rho_s = sum_k N_k erho_k,
// F'(rho_s) (pair_eam.cpp:1077, 1089-1096).
template <class P>
__global__ __launch_bounds__(kBlock) void k_site_fp(ElemArgs a, Tabs<P> t, const double *__restrict__ erho,
                                                    double *__restrict__ xs4) {
  const int o = xcd_block() * kBlock + (int)threadIdx.x;
  if (o >= a.n_site()) return;
  const int q = a.site(o);
  const int2 js = a.asite[q];
  const double *E = erho + 8 * (size_t)js.x;
  const double *N = a.spa_site + 8 * (size_t)js.y;
  double rho = 0.0;
  for (int k = 0; k < 8; ++k) rho += N[k] * E[k];
  xs4[4 * (size_t)q + 3] = (double)embed_fp<P>(t, rho);
}

// Node density, pair_eam.cpp:927-986: erho[i][k] from the list of the
// integration point at node k, measured from the node itself.
//
// When every integration point sits on a node (mask != null), the pass also
// records for the force pass which entries are in range: lane l of the 8
// serving row p walks entries l, l+8, ... and sets bit j of its words
// mask[(8p + l)*Wl + j/32] for its j-th entry (Wl = ceil(M/256)).  The node
// position equals the interpolated integration-point position bit for bit
// there (the shape functions are exactly 0 and 1).
//
// The integration-point passes read the spline tables through the cache
// hierarchy: with an LDS table window they would run one workgroup per CU,
// and these gather-latency-bound passes measured slower at that occupancy.
template <class P>
__global__ __launch_bounds__(kBlock) void k_eam_erho(ElemArgs a, Prd prd, Tabs<P> t, const double *__restrict__ px4,
                                                     double *__restrict__ erho, uint32_t *__restrict__ mask) {
  const int gt = xcd_block() * kBlock + (int)threadIdx.x;
  const int o = gt >> 3, l = gt & 7;
  const bool live = o < 8 * a.n_elem();
  const int i = live ? a.elem(o >> 3) : 0, k = o & 7;
  const int q = 8 * i + k;
  double acc = 0.0;
  if (live) {
    const double ix = a.nodex[24 * (size_t)i + 3 * k], iy = a.nodex[24 * (size_t)i + 3 * k + 1],
                 iz = a.nodex[24 * (size_t)i + 3 * k + 2];
    const int p = i * a.nint + a.node_list[k];
    const int jn = a.ipnum[p];
    const uint32_t *row = list_row(a.ipneigh, p, a.M);
    const int Wl = (a.M + 8 * 32 - 1) / (8 * 32);
    uint32_t *mrow = mask ? mask + ((size_t)p * kLanes + l) * Wl : nullptr;
    uint32_t bits = 0;
    for (int k0 = l, j0 = 0; k0 < jn; k0 += kU * kLanes, j0 += kU) {
      uint32_t e[kU];
      fetch_entries<kU>(row, k0, jn, e);
      Pt xj[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (e[u] != kNoEntry) xj[u] = load_xyz(px4, e[u], prd);
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (e[u] == kNoEntry) continue;
        const double delx = ix - xj[u].x, dely = iy - xj[u].y, delz = iz - xj[u].z;
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t.cutforcesq) {
          acc += (double)rho_of_r<P>(t, sqrt_p<P>(rsq));
          bits |= 1u << ((j0 + u) & 31);
        }
      }
      if (mrow && (((j0 + kU) & 31) == 0 || k0 + kU * kLanes >= jn)) {
        mrow[j0 >> 5] = bits;
        bits = 0;
      }
    }
  }
  acc = group_sum(acc);
  if (live && l == 0) erho[q] = acc;
}

// Integration-point force, pair_eam.cpp:1013-1129.  Writes f_g to ipf
// ([3][nip]) and the weighted energy / virial tallies.
// With mask != null only the entries the node-density pass marked in range
// are gathered.
template <class P, bool EV>
__global__ __launch_bounds__(kBlock) void k_eam_ip_force(ElemArgs a, Prd prd, Tabs<P> t,
                                                         const double *__restrict__ erho,
                                                         const double *__restrict__ px4, double *__restrict__ ipf,
                                                         double *__restrict__ eip, double *__restrict__ vip,
                                                         const uint32_t *__restrict__ mask) {
  const int gt = xcd_block() * kBlock + (int)threadIdx.x;
  const int o = gt >> 3, l = gt & 7;
  const bool live = o < a.n_elem() * a.nint;
  const int i = live ? a.elem(o / a.nint) : 0, g = o % a.nint;
  const int p = i * a.nint + g;
  double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0, rhoi = 0.0;
  P fpi = P(0);
  Virial vir;
  vir.zero();
  if (live) {
    const double *Ng = a.spa_int + 8 * g;
    double ix = 0.0, iy = 0.0, iz = 0.0;
    for (int kk = 0; kk < 8; ++kk) {
      ix += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk];
      iy += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk + 1];
      iz += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk + 2];
      rhoi += Ng[kk] * erho[8 * (size_t)i + kk];
    }
    fpi = embed_fp<P>(t, rhoi);
    const int jn = a.ipnum[p];
    const uint32_t *row = list_row(a.ipneigh, p, a.M);
    const int Wl = (a.M + 8 * 32 - 1) / (8 * 32);
    const uint32_t *mrow = mask ? mask + ((size_t)p * kLanes + l) * Wl : nullptr;
    uint32_t bits = 0xFFFFFFFFu;
    for (int k0 = l, j0 = 0; k0 < jn; k0 += kU * kLanes, j0 += kU) {
      if (mrow && (j0 & 31) == 0) bits = mrow[j0 >> 5];
      const uint32_t sel = bits >> (j0 & 31);
      uint32_t e[kU];
      fetch_entries<kU>(row, k0, jn, e);
      // atoms count in the node densities only: their pair forces reach the
      // element as the reactions of the atom pass (k_site_react_*)
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (!((sel >> u) & 1u) || (e[u] & 0x7FFFFFFu) < (uint32_t)a.natoms) e[u] = kNoEntry;
      Pt xjv[kU];
      double fpjv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (e[u] != kNoEntry) xjv[u] = load_xyzw(px4, e[u], prd, &fpjv[u]);
#pragma unroll
      for (int u = 0; u < kU; ++u) {
      if (e[u] == kNoEntry) continue;
      const Pt xj = xjv[u];
      const double fpj = fpjv[u];
      const double delx = ix - xj.x, dely = iy - xj.y, delz = iz - xj.z;
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < t.cutforcesq) {
        P phi;
        // psip = (fpi + fpj)*rhoip + phip, pair_eam.cpp:1118
        const RInvP<P> rr = r_and_inv<P>(rsq);
        P r = rr.r, rhoip, z2p, z2;
        rz_of_r<P>(t, r, &rhoip, &z2p, &z2);
        P recip = rr.rinv;
        phi = z2 * recip;
        P phip = z2p * recip - phi * recip;
        P psip = (fpi + (P)fpj) * rhoip + phip;
        const double fpair = (double)(-psip * recip);
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        if (EV) {
          ep += 0.5 * (double)phi;
          vir.add(delx, dely, delz, fpair);
        }
      }
      }
    }
  }
  fx = group_sum(fx);
  fy = group_sum(fy);
  fz = group_sum(fz);
  if (EV) {
    ep = group_sum(ep);
    vir.reduce();
  }
  if (live && l == 0) {
    ipf[p] = fx;
    ipf[(size_t)a.nip + p] = fy;
    ipf[2 * (size_t)a.nip + p] = fz;
    if (EV) {
      const double w = a.weight[p % a.nint];
      eip[p] = w * (embed_energy<P>(t, rhoi, (double)fpi) + ep);
      vir.store(vip, (size_t)a.nip, (size_t)p, w);
    }
  }
}

// ------------------------------------------------------------- clusters
// Element passes by vertex cluster (one wavefront per cluster of <= 8
// integration points at one mesh vertex, whose lists overlap almost
// entirely).  The cluster's union of list entries (built at rebuild) is evaluated once into LDS rows (x, y, z, fp): an atom row is read
// from x4, a site is interpolated from its element's nodes (and, for the
// force pass, its density from the element's node densities, then F'),
// with exactly k_site_pos's and k_site_fp's operations, so the rows equal
// the x4 rows those kernels would write.  The 8 lanes of each integration
// point then walk their list entries l, l+8, ... as in k_eam_erho /
// k_eam_ip_force, reading union indices (uint16) and LDS rows instead of
// gathering x4 rows from HBM, and sum in the same order: the same bits.
struct ClArgs {
  int ncl, n, stride;
  const int *slot;
  const uint32_t *uni;     // [ncl][stride]
  const uint16_t *local;   // ipneigh layout
  // per launch slot: cluster, slot start, union size, (scratch slot or -1)
  // * 16 + slot count
  const int4 *order;
  int count;
  int rows;                // LDS rows of this launch (its largest small union)
  // clusters whose union exceeds the LDS rows keep their rows in global
  // memory instead (a scratch slot in order[o].w), slot stride sstride
  // doubles, srows rows
  double *scratch;
  size_t sstride;
  int srows;
  const int *skip;  // nonzero: this step rebuilds first, return at once
};

// s / d for 0 <= s < 2^16 (an element has fewer than 65536 sites) and
// d >= 1: a float reciprocal estimate, at these sizes within one of the
// quotient, and one correction with 24-bit multiplies (full rate; a 32-bit
// integer division takes about 25 instructions, six of them quarter-rate
// multiplies)
__device__ __forceinline__ int site_div(int s, int d) {
  const int q = (int)((float)s * __builtin_amdgcn_rcpf((float)d));
  const int r = s - (int)__umul24((unsigned)q, (unsigned)d);
  return q + (r >= d ? 1 : 0) - (r < 0 ? 1 : 0);
}

// Shape functions of site s, computed as the host computes the site table
// (Element::inter_ele + shape_array_inter_ele, element.cpp:1361-1423: the
// same operations in the same order, unfused, so the same bits as
// spa_site).
__device__ __forceinline__ void site_shape(const ElemArgs &a, int s, double N[8]) {
#pragma clang fp contract(off)
  const int n12 = a.nae[1] * a.nae[2];
  const int ii = site_div(s, n12), rem = s - (int)__umul24((unsigned)ii, (unsigned)n12);
  const int jj = site_div(rem, a.nae[2]), kk = rem - (int)__umul24((unsigned)jj, (unsigned)a.nae[2]);
  const double px = a.pstep[0] * (double)ii - 1, py = a.pstep[1] * (double)jj - 1,
               pz = a.pstep[2] * (double)kk - 1;
  N[0] = 0.125 * (1 - px) * (1 - py) * (1 - pz);
  N[1] = 0.125 * (1 + px) * (1 - py) * (1 - pz);
  N[2] = 0.125 * (1 + px) * (1 + py) * (1 - pz);
  N[3] = 0.125 * (1 - px) * (1 + py) * (1 - pz);
  N[4] = 0.125 * (1 - px) * (1 - py) * (1 + pz);
  N[5] = 0.125 * (1 + px) * (1 - py) * (1 + pz);
  N[6] = 0.125 * (1 + px) * (1 + py) * (1 + pz);
  N[7] = 0.125 * (1 - px) * (1 + py) * (1 + pz);
}

// Union indices of a lane's next kU list entries k, k+8, ... (steps
// u0 .. u0+kU-1 of the lane, u0 a multiple of 4: one 8-byte load from the
// lane's run), -1 past the end or where the selection bits are clear.
template <int U>
__device__ __forceinline__ void cl_indices(const uint16_t *__restrict__ lrow, int u0, int k, int jn, uint32_t sel,
                                           int q[U]) {
  static_assert(U == 4, "one run of 4 steps");
  typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
  u16x4 v = {0, 0, 0, 0};
  if (k < jn) v = *reinterpret_cast<const u16x4 *>(lrow + (size_t)(u0 >> 2) * 32);
#pragma unroll
  for (int u = 0; u < U; ++u) q[u] = (k + u * kLanes < jn && ((sel >> u) & 1u)) ? (int)v[u] : -1;
}
__device__ __forceinline__ uint32_t k0_mask(const uint32_t *__restrict__ mrow, int jn, int l) {
  return l < jn ? mrow[0] : 0u;
}

// Site position x = sum_k N_k X_k and (FP) density rho = sum_k N_k E_k, in
// node order, k_site_pos's and k_site_fp's operations.
template <bool FP, class PX>
__device__ __forceinline__ void site_interp(PX X, PX E, const double N[8], double *x, double *y, double *z,
                                            double *rho) {
  double xx = 0.0, yy = 0.0, zz = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    xx += N[k] * X[3 * k];
    yy += N[k] * X[3 * k + 1];
    zz += N[k] * X[3 * k + 2];
  }
  *x = xx;
  *y = yy;
  *z = zz;
  if (FP) {
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) r += N[k] * E[k];
    *rho = r;
  }
}

// LDS of a cluster wavefront: rows [c.rows][4] (x, y, z, fp of every union
// entry), then the nodes [8][24] and node densities [8][8] of the cluster's
// own elements (the elements whose nodes meet at the vertex: nearly every
// union site lies in one of them).
//
// A cluster pass is a chain of dependent loads (cluster bounds, slots and
// union entries, then the node copy, the atom rows and the sites' elements,
// then the lists), each a full memory latency for a wavefront that has
// nothing else to do.  The fill is therefore written in stages that issue
// the loads of one level together (a lane's first three union rows, the node
// copy and the integration point's list header) before waiting on any.

// One union row's sources, requested before the node copy lands in LDS: the
// union entry and, for a site, its (element, site) pair (js.x = -1 for an
// atom, whose x4 row is read when the row is written).
struct ClRow {
  uint32_t e;
  int2 js;
};

__device__ __forceinline__ void cl_row_src(const ClArgs &c, const ElemArgs &a, int cl, int u, int nu, ClRow &r) {
  r.e = 0;
  r.js = make_int2(-1, 0);
  if (u >= nu) return;
  r.e = c.uni[(size_t)cl * c.stride + u];
  const int j = (int)(r.e & 0x7FFFFFFu);
  if (j >= c.n) r.js = a.asite[j - c.n];
}

// The row itself: a site is interpolated from its element's nodes (from the
// LDS copy when the element is one of the cluster's), then shifted by its
// periodic image.
template <class P, bool FP, bool LDS>
__device__ __forceinline__ void cl_row_put(const ClArgs &c, const ElemArgs &a, const Prd &prd, const Tabs<P> &t,
                                           const double *__restrict__ x4, const double *__restrict__ erho,
                                           const int (&el)[kClMax], const double *nodes, const double *enode,
                                           const ClRow &r, double *rows, int u) {
  double x, y, z, w = 0.0;
  if (r.js.x < 0) {
    const int j = (int)(r.e & 0x7FFFFFFu);
    const double2 xy = reinterpret_cast<const double2 *>(x4)[2 * (size_t)j];
    const double2 zw = reinterpret_cast<const double2 *>(x4)[2 * (size_t)j + 1];
    x = xy.x;
    y = xy.y;
    z = zw.x;
    // the force pass reads the atom's F'; the density pass marks atom rows
    w = FP ? zw.y : 1.0;
  } else {
    double N[8];
    site_shape(a, r.js.y, N);
    int li = -1;
#pragma unroll
    for (int i = 0; i < kClMax; ++i)
      if (el[i] == r.js.x) li = i;
    // one branch per memory space (the cluster's own elements from LDS,
    // any other from global memory): a pointer selected between the two
    // would be read with flat loads, which take the texture-address path
    // even for LDS data
    double rho = 0.0;
    using gptr = const __attribute__((address_space(1))) double *;
    if (li < 0)
      site_interp<FP>((gptr)(a.nodex + 24 * (size_t)r.js.x), (gptr)(FP ? erho + 8 * (size_t)r.js.x : nullptr), N, &x,
                      &y, &z, &rho);
    else if (LDS)
      site_interp<FP>((lds_cptr<double>)(nodes + 24 * li), (lds_cptr<double>)(enode + 8 * li), N, &x, &y, &z, &rho);
    else
      site_interp<FP>((gptr)(nodes + 24 * li), (gptr)(enode + 8 * li), N, &x, &y, &z, &rho);
    if (FP) w = (double)embed_fp<P>(t, rho);
  }
  const int img = (int)(r.e >> 27);
  if (img != kImgSelf) {
    int s[3];
    img_shift(img, s);
    x = shifted(x, s[0], prd.v[0]);
    y = shifted(y, s[1], prd.v[1]);
    z = shifted(z, s[2], prd.v[2]);
  }
  double *o = rows + 4 * u;
  o[0] = x;
  o[1] = y;
  o[2] = z;
  o[3] = w;
}

// Stage 1: the cluster's elements, the lane's first three union rows (most
// unions have fewer than 192 entries) and the node copy are requested
// together; the node copy is then stored to LDS.
struct ClFill {
  int el[kClMax];
  ClRow r0, r1, r2;
};


template <bool FP>
__device__ __forceinline__ void cl_fill_issue(const ClArgs &c, int cl, int s0, int ns, int nu, const ElemArgs &a,
                                              const double *__restrict__ x4, const double *__restrict__ erho,
                                              double *nodes, double *enode, ClFill &f, int lane) {
#pragma unroll
  for (int i = 0; i < kClMax; ++i) f.el[i] = i < ns ? (c.slot[s0 + i] >> 3) : -1;
  cl_row_src(c, a, cl, lane, nu, f.r0);
  cl_row_src(c, a, cl, lane + 64, nu, f.r1);
  cl_row_src(c, a, cl, lane + 128, nu, f.r2);
  // ns * 24 <= 3 * 64 node coordinates and ns * 8 <= 64 node densities
  double nv[3], ev = 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int q = lane + 64 * m;
    nv[m] = q < ns * 24 ? a.nodex[24 * (size_t)(c.slot[s0 + q / 24] >> 3) + q % 24] : 0.0;
  }
  if (FP && lane < ns * 8) ev = erho[8 * (size_t)(c.slot[s0 + lane / 8] >> 3) + lane % 8];
#pragma unroll
  for (int m = 0; m < 3; ++m)
    if (lane + 64 * m < ns * 24) nodes[lane + 64 * m] = nv[m];
  if (FP && lane < ns * 8) enode[lane] = ev;
}

// Stage 2 (after the barrier that publishes the node copy): every row.
template <class P, bool FP, bool LDS>
__device__ __forceinline__ void cl_fill_rows(const ClArgs &c, int cl, int nu, const ElemArgs &a, const Prd &prd,
                                             const Tabs<P> &t, const double *__restrict__ x4,
                                             const double *__restrict__ erho, double *rows, const double *nodes,
                                             const double *enode, const ClFill &f, int lane) {
  for (int u = lane, h = 0; u < nu; u += 64, ++h) {
    // the prefetched rows field by field (a selected struct would be
    // copied through scratch memory)
    ClRow r;
    r.e = h == 0 ? f.r0.e : h == 1 ? f.r1.e : f.r2.e;
    r.js.x = h == 0 ? f.r0.js.x : h == 1 ? f.r1.js.x : f.r2.js.x;
    r.js.y = h == 0 ? f.r0.js.y : h == 1 ? f.r1.js.y : f.r2.js.y;
    if (h > 2) cl_row_src(c, a, cl, u, nu, r);
    cl_row_put<P, FP, LDS>(c, a, prd, t, x4, erho, f.el, nodes, enode, r, rows, u);
  }
}

// rows: the cluster's union rows, in LDS or (a large union) in global
// memory; cap: their capacity.  Inlined once per memory space.
template <class P, bool LDS>
__device__ __forceinline__ void cl_erho_body(const ClArgs &c, const ElemArgs &a, const Prd &prd, const Tabs<P> &t,
                                             const double *__restrict__ px4, double *__restrict__ erho,
                                             uint32_t *__restrict__ mask, double *rows, int cap, int cl, int s0,
                                             int ns, int nu, int lane) {
  double *nodes = rows + 4 * cap, *enode = nodes + kClMax * 24;
  const int gi = lane >> 3, l = lane & 7;
  // the integration point of the lane's group: node k of element i (the
  // cluster's gi-th slot)
  const bool live = gi < ns;
  const int slot = live ? c.slot[s0 + gi] : 0;
  ClFill f;
  cl_fill_issue<false>(c, cl, s0, ns, nu, a, px4, nullptr, nodes, enode, f, lane);
  const int i = slot >> 3, k = slot & 7;
  const int p = i * a.nint + a.node_list[k];
  const int jn = live ? a.ipnum[p] : 0;
  const uint16_t *lrow = c.local + local_slot(p, l, a.M);
  // the lane's first three runs of union indices (its steps 0..11): a walk
  // of at most 12 steps needs no index loads in its loop
  const int steps = jn > l ? (jn - l + kLanes - 1) / kLanes : 0;
  const uint64_t *runs = reinterpret_cast<const uint64_t *>(lrow);
  const uint64_t w0 = steps > 0 ? runs[0] : 0, w1 = steps > 4 ? runs[8] : 0, w2 = steps > 8 ? runs[16] : 0;
  int qn[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) qn[u] = u < steps ? (int)((w0 >> (16 * u)) & 0xFFFFu) : -1;
  __syncthreads();
  cl_fill_rows<P, false, LDS>(c, cl, nu, a, prd, t, px4, nullptr, rows, nodes, enode, f, lane);
  __syncthreads();
  double acc = 0.0;
  if (live) {
    // the node's position from the LDS node copy: the shape-function sums
    // over its unit shape row give exactly the node's coordinates
    const double ix = nodes[24 * gi + 3 * k], iy = nodes[24 * gi + 3 * k + 1], iz = nodes[24 * gi + 3 * k + 2];
    const int Wl = (a.M + 8 * 32 - 1) / (8 * 32);
    uint32_t *mrow = mask + ((size_t)p * kLanes + l) * Wl;
    uint32_t bits = 0;
    // software pipeline: the next chunk's union indices are requested
    // before this chunk's spline rows
    for (int k0 = l, j0 = 0; k0 < jn; k0 += kU * kLanes, j0 += kU) {
      int q[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) q[u] = qn[u];
      if (steps <= 12) {
        const int j1 = j0 + kU;
        const uint64_t w = j1 == 4 ? w1 : w2;
#pragma unroll
        for (int u = 0; u < kU; ++u) qn[u] = j1 + u < steps ? (int)((w >> (16 * u)) & 0xFFFFu) : -1;
      } else {
        cl_indices<kU>(lrow, j0 + kU, k0 + kU * kLanes, jn, 0xFFFFFFFFu, qn);
      }
      // distances and knots, then the rows (all in flight), then the sums
      P pp[kU];
      int mm[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        mm[u] = -1;
        if (q[u] < 0) continue;
        const double *r = rows + 4 * q[u];
        const double delx = ix - r[0], dely = iy - r[1], delz = iz - r[2];
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t.cutforcesq) {
          mm[u] = rknot<P>(t, sqrt_p<P>(rsq), &pp[u]);
          // the force pass takes element sites only (atom rows carry 1.0)
          if (r[3] == 0.0) bits |= 1u << ((j0 + u) & 31);
        }
      }
      P c[kU][4];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (mm[u] >= 0) {
          const P *src = t.rv + 4 * (size_t)mm[u];
#pragma unroll
          for (int k = 0; k < 4; ++k) c[u][k] = src[k];
        }
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (mm[u] >= 0) {
          const P p = pp[u];  // rho_of_r's polynomial, pair_eam.cpp:655
          acc += (double)(((c[u][0] * p + c[u][1]) * p + c[u][2]) * p + c[u][3]);
        }
      if (((j0 + kU) & 31) == 0 || k0 + kU * kLanes >= jn) {
        mrow[j0 >> 5] = bits;
        bits = 0;
      }
    }
  }
  acc = group_sum(acc);
  if (live && l == 0) erho[slot] = acc;
}

template <class P>
__global__ __launch_bounds__(64) void k_cl_erho(ClArgs c, ElemArgs a, Prd prd, Tabs<P> t,
                                                const double *__restrict__ px4, double *__restrict__ erho,
                                                uint32_t *__restrict__ mask) {
  extern __shared__ __align__(16) double lds_rows[];
  const int o = xcd_block();
  if (o >= c.count) return;  // whole workgroup
  // the skip flag and the slot's cluster are requested together
  const int skip = *c.skip;
  const int4 m = c.order[o];
  if (skip) return;
  const int big = m.w >> 4, ns = m.w & 15;
  if (big >= 0)
    cl_erho_body<P, false>(c, a, prd, t, px4, erho, mask, c.scratch + (size_t)big * c.sstride, c.srows, m.x, m.y, ns,
                           m.z, (int)threadIdx.x);
  else
    cl_erho_body<P, true>(c, a, prd, t, px4, erho, mask, lds_rows, c.rows, m.x, m.y, ns, m.z, (int)threadIdx.x);
}

template <class P, bool EV, bool LDS>
__device__ __forceinline__ void cl_force_body(const ClArgs &c, const ElemArgs &a, const Prd &prd, const Tabs<P> &t,
                                              const double *__restrict__ erho, const double *__restrict__ px4,
                                              double *__restrict__ ipf, double *__restrict__ eip,
                                              double *__restrict__ vip, const uint32_t *__restrict__ mask,
                                              double *__restrict__ nodef, double *rows, int cap, int cl, int s0,
                                              int ns, int nu, int lane) {
  double *nodes = rows + 4 * cap, *enode = nodes + kClMax * 24;
  const int gi = lane >> 3, l = lane & 7;
  // the integration point of the lane's group is node k of element i (the
  // cluster's gi-th slot, whose nodes and node densities the fill puts in
  // LDS): the shape-function sums over its unit shape row give exactly the
  // node's position and density
  const bool live = gi < ns;
  const int slot = live ? c.slot[s0 + gi] : 0;
  ClFill f;
  cl_fill_issue<true>(c, cl, s0, ns, nu, a, px4, erho, nodes, enode, f, lane);
  const int i = slot >> 3, k = slot & 7;
  const int p = i * a.nint + a.node_list[k];
  const int jn = live ? a.ipnum[p] : 0;
  const uint16_t *lrow = c.local + local_slot(p, l, a.M);
  const int Wl = (a.M + 8 * 32 - 1) / (8 * 32);
  const uint32_t *mrow = mask + ((size_t)p * kLanes + l) * Wl;
  // the lane's first three runs of union indices (steps 0..11) and its
  // first mask word
  const int steps = jn > l ? (jn - l + kLanes - 1) / kLanes : 0;
  uint32_t bits = k0_mask(mrow, jn, l);
  const uint64_t *runs = reinterpret_cast<const uint64_t *>(lrow);
  const uint64_t w0 = steps > 0 ? runs[0] : 0, w1 = steps > 4 ? runs[8] : 0, w2 = steps > 8 ? runs[16] : 0;
  __syncthreads();
  const double rhoi = live ? enode[8 * gi + k] : 0.0;
  const P fpi = live ? embed_fp<P>(t, rhoi) : P(0);
  cl_fill_rows<P, true, LDS>(c, cl, nu, a, prd, t, px4, erho, rows, nodes, enode, f, lane);
  __syncthreads();
  double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
  Virial vir;
  vir.zero();
  if (live) {
    const double ix = nodes[24 * gi + 3 * k], iy = nodes[24 * gi + 3 * k + 1], iz = nodes[24 * gi + 3 * k + 2];
    // the pair terms of up to kU union rows (-1: none), in order
    auto pairs = [&](const int (&qc)[kU]) {
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (qc[u] < 0) continue;
        const double *r = rows + 4 * qc[u];
        const double delx = ix - r[0], dely = iy - r[1], delz = iz - r[2];
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t.cutforcesq) {
          P phi;
          const RInvP<P> rr = r_and_inv<P>(rsq);
          P rhoip, z2p, z2;
          rz_of_r<P>(t, rr.r, &rhoip, &z2p, &z2);
          P recip = rr.rinv;
          phi = z2 * recip;
          P phip = z2p * recip - phi * recip;
          P psip = (fpi + (P)r[3]) * rhoip + phip;
          const double fpair = (double)(-psip * recip);
          fx += delx * fpair;
          fy += dely * fpair;
          fz += delz * fpair;
          if (EV) {
            ep += 0.5 * (double)phi;
            vir.add(delx, dely, delz, fpair);
          }
        }
      }
    };
    if (steps <= 12) {
      // the whole walk is in registers: only the kept entries (the mask's
      // set bits: in range, element sites), four at a time in step order,
      // so the sums run in the order of the full walk without its skipped
      // entries and in fewer rounds
      while (bits) {
        int qc[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          qc[u] = -1;
          if (bits) {
            const int s = __builtin_ctz(bits);
            bits &= bits - 1;
            const uint64_t w = s < 4 ? w0 : s < 8 ? w1 : w2;
            qc[u] = (int)((w >> (16 * (s & 3))) & 0xFFFFu);
          }
        }
        pairs(qc);
      }
    } else {
      // a longer walk: every step, the skipped entries masked out; the union
      // indices and mask word of the next chunk are requested before this
      // chunk's spline rows, so they arrive while this chunk computes
      int q[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) q[u] = ((bits >> u) & 1u) ? (int)((w0 >> (16 * u)) & 0xFFFFu) : -1;
      for (int k0 = l, j0 = 0; k0 < jn; k0 += kU * kLanes, j0 += kU) {
        int qc[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) qc[u] = q[u];
        const int j1 = j0 + kU, k1 = k0 + kU * kLanes;
        if ((j1 & 31) == 0 && k1 < jn) bits = mrow[j1 >> 5];
        cl_indices<kU>(lrow, j1, k1, jn, bits >> (j1 & 31), q);
        pairs(qc);
      }
    }
  }
  fx = group_sum(fx);
  fy = group_sum(fy);
  fz = group_sum(fz);
  if (EV) {
    ep = group_sum(ep);
    vir.reduce();
  }
  if (live && l == 0) {
    ipf[p] = fx;
    ipf[(size_t)a.nip + p] = fy;
    ipf[2 * (size_t)a.nip + p] = fz;
    // the scatter of this node (every integration point on a node, shape
    // rows exact unit vectors): k_scatter's sum has this one nonzero term
    const int g = p % a.nint, k = slot & 7;
    const double wn = a.weight[g] * a.spa_int[8 * g + k];
    nodef[3 * (size_t)slot] = wn * fx;
    nodef[3 * (size_t)slot + 1] = wn * fy;
    nodef[3 * (size_t)slot + 2] = wn * fz;
    if (EV) {
      const double w = a.weight[p % a.nint];
      eip[p] = w * (embed_energy<P>(t, rhoi, (double)fpi) + ep);
      vir.store(vip, (size_t)a.nip, (size_t)p, w);
    }
  }
}

template <class P, bool EV>
__global__ __launch_bounds__(64) void k_cl_force(ClArgs c, ElemArgs a, Prd prd, Tabs<P> t,
                                                 const double *__restrict__ erho, const double *__restrict__ px4,
                                                 double *__restrict__ ipf, double *__restrict__ eip,
                                                 double *__restrict__ vip, const uint32_t *__restrict__ mask,
                                                 double *__restrict__ nodef) {
  extern __shared__ __align__(16) double lds_rows[];
  const int o = xcd_block();
  if (o >= c.count) return;  // whole workgroup
  const int skip = *c.skip;
  const int4 m = c.order[o];
  if (skip) return;
  const int big = m.w >> 4, ns = m.w & 15;
  if (big >= 0)
    cl_force_body<P, EV, false>(c, a, prd, t, erho, px4, ipf, eip, vip, mask, nodef,
                                c.scratch + (size_t)big * c.sstride, c.srows, m.x, m.y, ns, m.z, (int)threadIdx.x);
  else
    cl_force_body<P, EV, true>(c, a, prd, t, erho, px4, ipf, eip, vip, mask, nodef, lds_rows, c.rows, m.x, m.y, ns,
                               m.z, (int)threadIdx.x);
}

// LJ integration-point force, pair_lj_cut.cpp:355-422.
template <class P, bool EV>
__global__ __launch_bounds__(kBlock) void k_lj_ip_force(ElemArgs a, Prd prd, LjParams lj,
                                                        const double *__restrict__ px4, double *__restrict__ ipf,
                                                        double *__restrict__ eip, double *__restrict__ vip) {
  const int gt = xcd_block() * kBlock + (int)threadIdx.x;
  const int o = gt >> 3, l = gt & 7;
  const bool live = o < a.n_elem() * a.nint;
  const int i = live ? a.elem(o / a.nint) : 0, g = o % a.nint;
  const int p = i * a.nint + g;
  const P lj1 = (P)lj.lj1, lj2 = (P)lj.lj2, lj3 = (P)lj.lj3, lj4 = (P)lj.lj4;
  double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
  Virial vir;
  vir.zero();
  if (live) {
    const double *Ng = a.spa_int + 8 * g;
    double ix = 0.0, iy = 0.0, iz = 0.0;
    for (int kk = 0; kk < 8; ++kk) {
      ix += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk];
      iy += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk + 1];
      iz += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk + 2];
    }
    const int jn = a.ipnum[p];
    const uint32_t *row = list_row(a.ipneigh, p, a.M);
    for (int kk = l; kk < jn; kk += kLanes) {
      const uint32_t e = row[(size_t)kk * kRowGroup];
      if ((e & 0x7FFFFFFu) < (uint32_t)a.natoms) continue;  // atoms: reactions of the atom pass
      const Pt xj = load_xyz(px4, e, prd);
      const double delx = ix - xj.x, dely = iy - xj.y, delz = iz - xj.z;
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < lj.cutsq) {
        P r2inv = P(1.0) / (P)rsq;
        P r6inv = r2inv * r2inv * r2inv;
        P forcelj = r6inv * (lj1 * r6inv - lj2);
        const double fpair = (double)(forcelj * r2inv);
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        if (EV) {
          ep += 0.5 * (double)(r6inv * (lj3 * r6inv - lj4));
          vir.add(delx, dely, delz, fpair);
        }
      }
    }
  }
  fx = group_sum(fx);
  fy = group_sum(fy);
  fz = group_sum(fz);
  if (EV) {
    ep = group_sum(ep);
    vir.reduce();
  }
  if (live && l == 0) {
    ipf[p] = fx;
    ipf[(size_t)a.nip + p] = fy;
    ipf[2 * (size_t)a.nip + p] = fz;
    if (EV) {
      const double w = a.weight[p % a.nint];
      eip[p] = w * ep;
      vir.store(vip, (size_t)a.nip, (size_t)p, w);
    }
  }
}

// Scatter to nodes, pair_eam.cpp:1130-1134 / pair_lj_cut.cpp:424-428:
// nodef[i][k] = sum_g (w_g*N_k(g))*f_g in integration-point order, starting
// from the cleared force (Verlet::node_force_clear, verlet.cpp:131-144).
// One thread per node, no atomics, fixed order.
__global__ void k_scatter(ElemArgs a, const double *__restrict__ ipf, double *__restrict__ nodef) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= 8 * a.n_elem()) return;
  const int i = a.elem(o >> 3), k = o & 7;
  const int q = 8 * i + k;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int g = 0; g < a.nint; ++g) {
    const int p = i * a.nint + g;
    const double wn = a.weight[g] * a.spa_int[8 * g + k];
    fx += wn * ipf[p];
    fy += wn * ipf[(size_t)a.nip + p];
    fz += wn * ipf[2 * (size_t)a.nip + p];
  }
  nodef[3 * (size_t)q] = fx;
  nodef[3 * (size_t)q + 1] = fy;
  nodef[3 * (size_t)q + 2] = fz;
}

// Reactions of the atom-site pair forces (atom-element coupling).  The atom
// pass gives atom i the force f_is of every element site s in its list; the
// site takes -f_is, and site s passes it to its element's nodes through its
// shape functions, nodef[e][k] += N_k(s) * R_s, so atoms and nodes receive
// equal and opposite totals (sum_k N_k = 1).  k_site_react_* recompute f_is
// for the atoms of each site's reverse list (rkey, ascending) with the atom
// pass's own operations: del = x_i - (x_s + img*prd), the same pair
// arithmetic and arguments in the same order (the global rows, whose
// coefficients the atom pass's knot form reproduces bit for bit).
struct ReactArgs {
  int nrs, n;
  const int4 *meta;  // [nrs] active site a, rkey range [rstart[a], rstart[a+1])
  const uint32_t *rkey;
  double *R;  // [3][nrs]
  const int *skip;
};

// 8 lanes per reaction site (the layout of the integration-point passes):
// lane l takes the site's reverse entries l, l+8, ..., fetching kU of them
// and their atom rows before the arithmetic; the lanes' partial sums are
// combined by group_sum.
template <class P>
__global__ __launch_bounds__(kBlock) void k_site_react_eam(ReactArgs r, const double *__restrict__ x4, Prd prd,
                                                           Tabs<P> t) {
  const int gt = xcd_block() * kBlock + (int)threadIdx.x;
  const int q = gt >> 3, l = gt & 7;
  const bool live = q < r.nrs;
  // the skip flag and the site's entry range are requested together
  const int skip = *r.skip;
  const int4 m = live ? r.meta[q] : make_int4(0, 0, 0, 0);
  if (skip) return;
  double rx = 0.0, ry = 0.0, rz = 0.0;
  if (live) {
    const size_t row = (size_t)r.n + m.x;
    const double2 sxy = reinterpret_cast<const double2 *>(x4)[2 * row];
    const double2 szw = reinterpret_cast<const double2 *>(x4)[2 * row + 1];
    const P fps = (P)szw.y;
    const int k1 = m.z;
    for (int k0 = m.y + l; k0 < k1; k0 += kU * kLanes) {
      uint32_t key[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) key[u] = k0 + u * kLanes < k1 ? r.rkey[k0 + u * kLanes] : kNoEntry;
      double2 ixy[kU], izw[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (key[u] != kNoEntry) {
          const size_t i = key[u] & 0x7FFFFFFu;
          ixy[u] = reinterpret_cast<const double2 *>(x4)[2 * i];
          izw[u] = reinterpret_cast<const double2 *>(x4)[2 * i + 1];
        }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (key[u] == kNoEntry) continue;
        const int img = (int)(key[u] >> 27);
        double sx = sxy.x, sy = sxy.y, sz = szw.x;
        if (img != kImgSelf) {
          int sh[3];
          img_shift(img, sh);
          sx = shifted(sx, sh[0], prd.v[0]);
          sy = shifted(sy, sh[1], prd.v[1]);
          sz = shifted(sz, sh[2], prd.v[2]);
        }
        const double delx = ixy[u].x - sx, dely = ixy[u].y - sy, delz = izw[u].x - sz;
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t.cutforcesq) {
          P phi;
          const double fpair = eam_fpair_global<P>(t, rsq, (P)izw[u].y, fps, &phi);
          rx -= delx * fpair;
          ry -= dely * fpair;
          rz -= delz * fpair;
        }
      }
    }
  }
  rx = group_sum(rx);
  ry = group_sum(ry);
  rz = group_sum(rz);
  if (live && l == 0) {
    r.R[q] = rx;
    r.R[(size_t)r.nrs + q] = ry;
    r.R[2 * (size_t)r.nrs + q] = rz;
  }
}

template <class P>
__global__ __launch_bounds__(kBlock) void k_site_react_lj(ReactArgs r, const double *__restrict__ x4, Prd prd,
                                                          LjParams lj) {
  const int gt = xcd_block() * kBlock + (int)threadIdx.x;
  const int q = gt >> 3, l = gt & 7;
  const bool live = q < r.nrs;
  const P lj1 = (P)lj.lj1, lj2 = (P)lj.lj2;
  double rx = 0.0, ry = 0.0, rz = 0.0;
  if (live) {
    const int4 m = r.meta[q];
    const size_t row = (size_t)r.n + m.x;
    const double sx0 = x4[4 * row], sy0 = x4[4 * row + 1], sz0 = x4[4 * row + 2];
    for (int k = m.y + l; k < m.z; k += kLanes) {
      const uint32_t key = r.rkey[k];
      const size_t i = key & 0x7FFFFFFu;
      const int img = (int)(key >> 27);
      double sx = sx0, sy = sy0, sz = sz0;
      if (img != kImgSelf) {
        int sh[3];
        img_shift(img, sh);
        sx = shifted(sx, sh[0], prd.v[0]);
        sy = shifted(sy, sh[1], prd.v[1]);
        sz = shifted(sz, sh[2], prd.v[2]);
      }
      const double delx = x4[4 * i] - sx, dely = x4[4 * i + 1] - sy, delz = x4[4 * i + 2] - sz;
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < lj.cutsq) {
        P r2inv = P(1.0) / (P)rsq;
        P r6inv = r2inv * r2inv * r2inv;
        P forcelj = r6inv * (lj1 * r6inv - lj2);
        const double fpair = (double)(P(1.0) * forcelj * r2inv);
        rx -= delx * fpair;
        ry -= dely * fpair;
        rz -= delz * fpair;
      }
    }
  }
  rx = group_sum(rx);
  ry = group_sum(ry);
  rz = group_sum(rz);
  if (live && l == 0) {
    r.R[q] = rx;
    r.R[(size_t)r.nrs + q] = ry;
    r.R[2 * (size_t)r.nrs + q] = rz;
  }
}

// One wavefront per element with reaction sites: lane l takes the sites
// l, l+64, ... of the element's range and sums N_k(s) * R_s for its 8 nodes
// (24 partial sums); a fixed xor butterfly combines the lanes (every lane ends
// with the same bits), and lanes 0..23 add node k = l/3, component l%3 to the
// node force written by the integration-point pass.  rsl[q]: the site index
// of reaction site q in its element.  The site indices and reactions of four
// steps are requested before the first is summed (an element has a few
// hundred reaction sites, so this is the whole range), in the same order.
__global__ __launch_bounds__(64) void k_site_react_nodes(ReactArgs r, int nre, const int *__restrict__ relem,
                                                         const int *__restrict__ rel_start,
                                                         const int *__restrict__ rsl, ElemArgs a,
                                                         double *__restrict__ nodef) {
  const int ri = xcd_block();
  if (ri >= nre) return;  // whole wavefront
  const int skip = *r.skip;
  const int q0 = rel_start[ri], q1 = rel_start[ri + 1], e = relem[ri];
  if (skip) return;
  const int lane = (int)threadIdx.x;
  double acc[24];
#pragma unroll
  for (int c = 0; c < 24; ++c) acc[c] = 0.0;
  constexpr int kSteps = 4;
  for (int qb = q0 + lane; qb < q1; qb += kSteps * 64) {
    int si[kSteps];
    double R[kSteps][3];
#pragma unroll
    for (int h = 0; h < kSteps; ++h) {
      const int q = qb + 64 * h;
      si[h] = 0;
      R[h][0] = R[h][1] = R[h][2] = 0.0;
      if (q < q1) {
        si[h] = rsl[q];
        R[h][0] = r.R[q];
        R[h][1] = r.R[(size_t)r.nrs + q];
        R[h][2] = r.R[2 * (size_t)r.nrs + q];
      }
    }
#pragma unroll
    for (int h = 0; h < kSteps; ++h) {
      if (qb + 64 * h >= q1) break;
      double N[8];
      site_shape(a, si[h], N);
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[3 * k + d] += N[k] * R[h][d];
    }
  }
#pragma unroll
  for (int c = 0; c < 24; ++c)
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) acc[c] += __shfl_xor(acc[c], m, 64);
  double mine = 0.0;
#pragma unroll
  for (int c = 0; c < 24; ++c)
    if (lane == c) mine = acc[c];
  if (lane < 24) nodef[24 * (size_t)e + lane] += mine;
}

inline unsigned grid_for(long n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

Prd prd_of(const cac_ctx *ctx) {
  Prd p;
  for (int d = 0; d < 3; ++d) p.v[d] = ctx->prd[d];
  return p;
}

template <class P>
Tabs<P> tabs_of(const cac_ctx *ctx);
template <>
Tabs<double> tabs_of<double>(const cac_ctx *ctx) {
  Tabs<double> t;
  t.rv = ctx->eam.rv;
  t.rz = ctx->eam.rz;
  t.fr = ctx->eam.fr;
  t.rk = ctx->eam.knots ? ctx->eam.rk : nullptr;
  t.nr = ctx->eam.nr;
  t.nrho = ctx->eam.nrho;
  t.rdr = ctx->eam.rdr;
  t.rdrho = ctx->eam.rdrho;
  t.cutforcesq = ctx->eam.cutforcesq;
  t.rhomax = ctx->eam.rhomax;
  return t;
}
template <>
Tabs<float> tabs_of<float>(const cac_ctx *ctx) {
  Tabs<float> t;
  t.rv = ctx->eam.rv_f;
  t.rz = ctx->eam.rz_f;
  t.fr = ctx->eam.fr_f;
  t.rk = nullptr;
  t.nr = ctx->eam.nr;
  t.nrho = ctx->eam.nrho;
  t.rdr = (float)ctx->eam.rdr;
  t.rdrho = (float)ctx->eam.rdrho;
  t.cutforcesq = ctx->eam.cutforcesq;
  t.rhomax = ctx->eam.rhomax;
  return t;
}

ElemArgs elem_args(const cac_ctx *ctx) {
  ElemArgs a;
  a.ne = ctx->ne;
  a.nint = ctx->nint;
  a.nip = ctx->ne * ctx->nint;
  a.nactive = ctx->nactive;
  a.natoms = ctx->n;
  a.M = ctx->maxnip;
  a.nodex = ctx->nodex;
  a.spa_site = ctx->spa_site;
  a.spa_int = ctx->spa_int;
  a.weight = ctx->weight;
  a.ipneigh = ctx->ipneigh;
  a.ipnum = ctx->ipnum;
  a.asite = ctx->asite;
  for (int k = 0; k < 8; ++k) a.node_list[k] = ctx->node_list[k];
  for (int d = 0; d < 3; ++d) {
    a.nae[d] = ctx->nae[d];
    a.pstep[d] = 2 / ((double)ctx->nae[d] - 1);
  }
  a.elist = nullptr;
  a.nelist = 0;
  a.slist = nullptr;
  a.nslist = 0;
  return a;
}

long host_nsite(const ElemArgs &a) { return a.slist ? a.nslist : a.nactive; }
long host_nelem(const ElemArgs &a) { return a.elist ? a.nelist : a.ne; }

ElemArgs with_elems(ElemArgs a, const int *el, int n) {
  a.elist = el;
  a.nelist = n;
  return a;
}
ElemArgs with_sites(ElemArgs a, const int *sl, int n) {
  a.slist = sl;
  a.nslist = n;
  return a;
}

int site_positions(cac_ctx *ctx, const ElemArgs &a) {
  if (host_nsite(a) == 0) return CAC_OK;
  hipEvent_t t0;
  timer_begin(ctx, "site_pos", &t0);
  hipLaunchKernelGGL(k_site_pos, dim3(xcd_grid(host_nsite(a), kBlock)), dim3(kBlock), 0, ctx->stream, a, ctx->xs4);
  timer_end(ctx, "site_pos", t0);
  return CAC_OK;
}

int scatter(cac_ctx *ctx, const ElemArgs &a) {
  const long nf = host_nelem(a);
  if (nf == 0) return CAC_OK;
  hipEvent_t t0;
  timer_begin(ctx, "scatter", &t0);
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(8L * nf)), dim3(kBlock), 0, ctx->stream, a, ctx->ipf, ctx->nodef);
  timer_end(ctx, "scatter", t0);
  return CAC_OK;
}

// The reactions of one force call, after every node force is written.
template <class P>
int site_reactions(cac_ctx *ctx, bool eam, const ElemArgs &a) {
  if (ctx->nrs == 0 || ctx->nre == 0) return CAC_OK;
  ReactArgs r;
  r.nrs = ctx->nrs;
  r.n = ctx->n;
  r.meta = ctx->rmeta;
  r.rkey = ctx->rkey;
  r.R = ctx->rforce;
  r.skip = ctx->dflags + 15;
  hipEvent_t t0;
  timer_begin(ctx, "site_react", &t0);
  const unsigned g = xcd_grid((long)ctx->nrs * kLanes, kBlock);
  if (eam)
    hipLaunchKernelGGL((k_site_react_eam<P>), dim3(g), dim3(kBlock), 0, ctx->stream, r, ctx->x4, prd_of(ctx),
                       tabs_of<P>(ctx));
  else
    hipLaunchKernelGGL((k_site_react_lj<P>), dim3(g), dim3(kBlock), 0, ctx->stream, r, ctx->x4, prd_of(ctx),
                       ctx->lj);
  hipLaunchKernelGGL(k_site_react_nodes, dim3(xcd_grid(64L * ctx->nre, 64)), dim3(64), 0, ctx->stream, r, ctx->nre,
                     ctx->relem, ctx->rel_start, ctx->rsl, a, ctx->nodef);
  timer_end(ctx, "site_react", t0);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

// LDS window of a per-knot table with rows of `rowbytes`: the last rows up to
// the device's shared-memory limit per workgroup (rows m0 .. nr-1).
struct TabWindow {
  int m0;
  size_t bytes;
};
TabWindow table_window(const cac_ctx *ctx, size_t rowbytes) {
  const int nr = ctx->eam.nr;
  long rows = (long)(ctx->lds_bytes / rowbytes);
  if (rows > nr - 1) rows = nr - 1;
  if (rows < 1) rows = 1;
  TabWindow w;
  w.m0 = nr - (int)rows;
  w.bytes = (size_t)rows * rowbytes;
  return w;
}

// LDS window of the knot form: knots m0 .. nr (row m reads knots m, m+1).
TabWindow knot_window(const cac_ctx *ctx) {
  const int nr = ctx->eam.nr;
  long kn = (long)(ctx->lds_bytes / (4 * sizeof(double)));
  if (kn > nr) kn = nr;
  TabWindow w;
  w.m0 = nr + 1 - (int)kn;
  w.bytes = (size_t)kn * 4 * sizeof(double);
  return w;
}

// The atom density pass (see k_eam_atom_rho): full or inner, as the device
// flag says; the inner mask is then valid for the current positions, and
// the atom force pass that follows clears the flag.
template <class P, bool EV>
int atom_rho(cac_ctx *ctx, const Tabs<P> &t, AtomGrid agrid, const TabWindow &wr) {
  const double skin = ctx->cutneigh - sqrt(ctx->eam.cutforcesq);
  const double margin = 0.4 * (skin > 0.0 ? skin : 0.0);
  const double cin = sqrt(ctx->eam.cutforcesq) + margin;
  const double cutinsq = cin * cin;
  int *ibad = ctx->dflags + 14;
  if (cutinsq != ctx->inner_cutsq) {  // new cutoff or skin: the inner mask is stale
    ctx->inner_cutsq = cutinsq;
    CAC_HIP(ctx, hipMemsetAsync(ibad, 0xFF, sizeof(int), ctx->stream));
  }
  ctx->inner_margin = margin;
  ctx->inner_cut = cin;
  hipLaunchKernelGGL((k_eam_atom_rho<P, EV>), dim3(agrid.blocks), dim3(kAtomWG), wr.bytes, ctx->stream, ctx->n,
                     ctx->x4, prd_of(ctx), ctx->aneigh, ctx->maxneigh, ctx->anum, t, wr.m0, ctx->rho, ctx->eatom,
                     ctx->amask, ibad, ctx->amask_in, ctx->xhold_in, cutinsq, agrid.per, ctx->dflags + 15);
  return CAC_OK;
}

// In-range masks written by the density passes: atoms one bit per list
// entry (mask_slot layout), integration points one word group per lane
// (only when every integration point is a node point).
int ensure_masks(cac_ctx *ctx) {
  int rc;
  if (ctx->n) {
    const size_t rows = (size_t)round_up(ctx->n, kAtomGroup);
    const size_t words = rows * ((ctx->maxneigh + 31) / 32);
    if ((rc = ensure(ctx, ctx->b_amask, words, &ctx->amask)) ||
        (rc = ensure(ctx, ctx->b_amask_in, words, &ctx->amask_in)) ||
        (rc = ensure(ctx, ctx->b_xhold_in, 3 * (size_t)ctx->n, &ctx->xhold_in)))
      return rc;
  }
  const long nip = (long)ctx->ne * ctx->nint;
  if (nip && ctx->ip_on_nodes) {
    const int wl = (ctx->maxnip + 8 * 32 - 1) / (8 * 32);
    if ((rc = ensure(ctx, ctx->b_ipmask, (size_t)nip * kLanes * wl, &ctx->ipmask))) return rc;
  }
  return CAC_OK;
}

ClArgs cl_args(const cac_ctx *ctx) {
  ClArgs c;
  c.ncl = ctx->ncl;
  c.n = ctx->n;
  c.stride = ctx->cl_stride;
  c.slot = ctx->cl_slot;
  c.uni = ctx->cl_union;
  c.local = ctx->iplocal;
  // one launch: LDS rows for the small unions (most of them), so the few
  // large unions at interfaces do not lower the occupancy of all; those keep
  // their rows in global memory and start first on every XCD (see
  // launch_clusters), running beside the rest
  c.order = ctx->cl_order;
  c.count = ctx->ncl;
  c.rows = ctx->cl_umax_small > 0 ? ctx->cl_umax_small : 1;
  c.scratch = ctx->cl_scratch;
  c.srows = ctx->cl_umax;
  c.sstride = (size_t)4 * ctx->cl_umax + kClMax * (24 + 8);
  c.skip = ctx->dflags + 15;
  return c;
}
size_t cl_lds(const ClArgs &c) { return ((size_t)4 * c.rows + kClMax * (24 + 8)) * sizeof(double); }

// One EAM force call with the element passes by vertex cluster (every
// element in the passes, the cluster structures built): only the active
// sites that atom lists refer to are still interpolated into x4.
template <class P, bool EV>
int eam_clusters(cac_ctx *ctx, const ElemArgs &base) {
  Tabs<P> t = tabs_of<P>(ctx);
  const int n = ctx->n;
  const AtomGrid agrid = atom_grid(ctx, n);
  const TabWindow wr = table_window(ctx, 4 * sizeof(P));
  const TabWindow wz = t.rk ? knot_window(ctx) : table_window(ctx, 8 * sizeof(P));
  hipEvent_t t0;
  int rc;
  if ((rc = ensure_masks(ctx))) return rc;
  const ElemArgs aref = with_sites(base, ctx->aref, ctx->naref);
  const ClArgs c = cl_args(ctx);

  if (ctx->naref) site_positions(ctx, aref);
  if (n) {
    CAC_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eam_atom_rho<P, EV>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)wr.bytes));
    CAC_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eam_atom_force<P, EV>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)wz.bytes));
    timer_begin(ctx, "atom_rho", &t0);
    if ((rc = atom_rho<P, EV>(ctx, t, agrid, wr))) return rc;
    timer_end(ctx, "atom_rho", t0);
  }
  timer_begin(ctx, "ip_erho", &t0);
  if (c.count)
    hipLaunchKernelGGL((k_cl_erho<P>), dim3(xcd_grid((long)c.count * 64, 64)), dim3(64), cl_lds(c), ctx->stream, c,
                       base, prd_of(ctx), t, ctx->x4, ctx->erho, ctx->ipmask);
  timer_end(ctx, "ip_erho", t0);
  if (ctx->naref) {
    timer_begin(ctx, "site_fp", &t0);
    hipLaunchKernelGGL((k_site_fp<P>), dim3(xcd_grid(ctx->naref, kBlock)), dim3(kBlock), 0, ctx->stream, aref, t,
                       ctx->erho, ctx->xs4);
    timer_end(ctx, "site_fp", t0);
  }
  if (n) {
    timer_begin(ctx, "atom_force", &t0);
    hipLaunchKernelGGL((k_eam_atom_force<P, EV>), dim3(agrid.blocks), dim3(kAtomWG), wz.bytes, ctx->stream, n,
                       ctx->x4, prd_of(ctx), ctx->aneigh, ctx->maxneigh, ctx->anum, t, wz.m0, ctx->f, ctx->eatom,
                       ctx->vatom, ctx->amask, agrid.per, ctx->dflags + 14, ctx->dflags + 15);
    timer_end(ctx, "atom_force", t0);
  }
  timer_begin(ctx, "ip_force", &t0);
  // also the node forces: every integration point is on a node (the clusters
  // need that), so each writes its node's scatter term (no k_scatter launch)
  if (c.count)
    hipLaunchKernelGGL((k_cl_force<P, EV>), dim3(xcd_grid((long)c.count * 64, 64)), dim3(64), cl_lds(c), ctx->stream,
                       c, base, prd_of(ctx), t, ctx->erho, ctx->x4, ctx->ipf, ctx->eip, ctx->vip, ctx->ipmask,
                       ctx->nodef);
  timer_end(ctx, "ip_force", t0);
  CAC_HIP(ctx, hipGetLastError());
  return site_reactions<P>(ctx, true, base);
}

// One EAM force call (PairEAM::compute + ecompute, pair_eam.cpp:584-746,
// 927-1137) in the order the coupled lists need: site positions, atom
// density + F', node densities, site F', atom forces, integration-point
// forces + scatter.  Without coupling the atom and element halves are
// independent and the order changes nothing.  The element passes run over
// optional subsets (surrogate mode): site positions over `spos`, node
// densities over `erho_el`, site F' over `sfp`, forces and scatter over
// `force_el`.
template <class P, bool EV>
int eam_all(cac_ctx *ctx, const ElemArgs &base, const ElemArgs &spos, const ElemArgs &erho_el, const ElemArgs &sfp,
            const ElemArgs &force_el) {
  Tabs<P> t = tabs_of<P>(ctx);
  const int n = ctx->n;
  const bool el = ctx->ne > 0;
  const AtomGrid agrid = atom_grid(ctx, n);
  const TabWindow wr = table_window(ctx, 4 * sizeof(P));
  const TabWindow wz = t.rk ? knot_window(ctx) : table_window(ctx, 8 * sizeof(P));
  hipEvent_t t0;
  int rc;
  if ((rc = ensure_masks(ctx))) return rc;
  uint32_t *ipm = ctx->ip_on_nodes ? ctx->ipmask : nullptr;
  if (el) site_positions(ctx, spos);
  if (n) {
    CAC_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eam_atom_rho<P, EV>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)wr.bytes));
    CAC_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eam_atom_force<P, EV>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)wz.bytes));
    timer_begin(ctx, "atom_rho", &t0);
    if ((rc = atom_rho<P, EV>(ctx, t, agrid, wr))) return rc;
    timer_end(ctx, "atom_rho", t0);
  }
  if (el && host_nelem(erho_el)) {
    timer_begin(ctx, "ip_erho", &t0);
    hipLaunchKernelGGL((k_eam_erho<P>), dim3(xcd_grid(8L * host_nelem(erho_el) * kLanes, kBlock)), dim3(kBlock), 0,
                       ctx->stream, erho_el, prd_of(ctx), t, ctx->x4, ctx->erho, ipm);
    timer_end(ctx, "ip_erho", t0);
  }
  if (el && host_nsite(sfp)) {
    timer_begin(ctx, "site_fp", &t0);
    hipLaunchKernelGGL((k_site_fp<P>), dim3(xcd_grid(host_nsite(sfp), kBlock)), dim3(kBlock), 0, ctx->stream, sfp, t,
                       ctx->erho, ctx->xs4);
    timer_end(ctx, "site_fp", t0);
  }
  if (n) {
    timer_begin(ctx, "atom_force", &t0);
    hipLaunchKernelGGL((k_eam_atom_force<P, EV>), dim3(agrid.blocks), dim3(kAtomWG), wz.bytes, ctx->stream, n,
                       ctx->x4, prd_of(ctx), ctx->aneigh, ctx->maxneigh, ctx->anum, t, wz.m0, ctx->f, ctx->eatom,
                       ctx->vatom, ctx->amask, agrid.per, ctx->dflags + 14, ctx->dflags + 15);
    timer_end(ctx, "atom_force", t0);
  }
  const long nf = el ? host_nelem(force_el) : 0;
  if (nf) {
    timer_begin(ctx, "ip_force", &t0);
    hipLaunchKernelGGL((k_eam_ip_force<P, EV>), dim3(xcd_grid(nf * base.nint * kLanes, kBlock)), dim3(kBlock), 0,
                       ctx->stream, force_el, prd_of(ctx), t, ctx->erho, ctx->x4, ctx->ipf, ctx->eip, ctx->vip, ipm);
    timer_end(ctx, "ip_force", t0);
    scatter(ctx, force_el);
  }
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

template <class P, bool EV>
int eam_t(cac_ctx *ctx) {
  ElemArgs a = elem_args(ctx);
  const SurrState &S = ctx->surr;
  if (ctx->ne && S.enabled && S.ready && !EV) {
    int rc = eam_all<P, false>(ctx, a, with_sites(a, S.siteR2, S.nsR2), with_elems(a, S.elemR, S.nR),
                               with_sites(a, S.siteR, S.nsR), with_elems(a, S.elemF, S.nF));
    if (rc) return rc;
    if ((rc = surrogate_forces(ctx))) return rc;
    return site_reactions<P>(ctx, true, a);
  }
  if (ctx->ne && ctx->cl_ok) return eam_clusters<P, EV>(ctx, a);
  int rc = eam_all<P, EV>(ctx, a, a, a, a, a);
  if (rc) return rc;
  return site_reactions<P>(ctx, true, a);
}

// One LJ force call (PairLJCut::compute + ecompute, pair_lj_cut.cpp:227-288,
// 295-442): site positions, atom forces, integration-point forces, scatter.
template <class P, bool EV>
int lj_t(cac_ctx *ctx) {
  ElemArgs a = elem_args(ctx);
  const int n = ctx->n;
  hipEvent_t t0;
  if (ctx->ne) site_positions(ctx, a);
  if (n) {
    timer_begin(ctx, "atom_force", &t0);
    hipLaunchKernelGGL((k_lj_atom_force<P, EV>), dim3(xcd_grid((long)n, kBlock)), dim3(kBlock), 0,
                       ctx->stream, n, ctx->x4, prd_of(ctx), ctx->aneigh, ctx->maxneigh, ctx->anum, ctx->lj, ctx->f,
                       ctx->eatom, ctx->vatom);
    timer_end(ctx, "atom_force", t0);
  }
  if (ctx->ne) {
    timer_begin(ctx, "ip_force", &t0);
    hipLaunchKernelGGL((k_lj_ip_force<P, EV>), dim3(xcd_grid((long)a.nip * kLanes, kBlock)), dim3(kBlock), 0,
                       ctx->stream, a, prd_of(ctx), ctx->lj, ctx->x4, ctx->ipf, ctx->eip, ctx->vip);
    timer_end(ctx, "ip_force", t0);
    scatter(ctx, a);
  }
  CAC_HIP(ctx, hipGetLastError());
  return site_reactions<P>(ctx, false, a);
}

}  // namespace

#define CAC_DISPATCH(fn, ctx, ev, precision)                                      \
  ((precision) == CAC_PREC_MIXED ? ((ev) ? fn<float, true>(ctx) : fn<float, false>(ctx)) \
                                 : ((ev) ? fn<double, true>(ctx) : fn<double, false>(ctx)))

int launch_eam(cac_ctx *ctx, int eflag, int vflag, int precision) {
  return CAC_DISPATCH(eam_t, ctx, eflag || vflag, precision);
}

int launch_lj(cac_ctx *ctx, int eflag, int vflag, int precision) {
  return CAC_DISPATCH(lj_t, ctx, eflag || vflag, precision);
}

// Interpolated positions of every active site, the rows j >= n that coupled
// atom lists refer to (the force passes refresh only those they read).
int atom_list_site_rows(cac_ctx *ctx) {
  if (ctx->ne == 0 || ctx->nactive == 0) return CAC_OK;
  site_positions(ctx, elem_args(ctx));
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

}  // namespace cac
