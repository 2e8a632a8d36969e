// k_pair.hip -- EAM and LJ force passes over atoms and element integration
// points, and the integration-point -> node scatter.
//
// FP64 path: the reference's arithmetic term by term (pair_eam.cpp:584-746,
// 927-1137; pair_lj_cut.cpp:227-288, 295-442), compiled with
// -ffp-contract=off.  Results differ from the CPU oracle only through the
// order of the neighbor sums.
// MIXED path (Prec = float): r, the spline polynomials, F'(rho) and the pair
// force are evaluated in FP32; positions, site interpolation, rho, force,
// energy and virial sums stay FP64.
//
// Table layout on the device: one 16-wide row per r-knot, rhor c0..c6 in
// slots 0..6 and z2r c0..c6 in slots 8..14, so the force pass reads both
// splines of a pair from one 128-byte line (64 bytes in FP32); frho rows are
// 8 wide.
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

struct Prd {
  double v[3];
};

template <class P>
struct Tabs {
  const P *rz;  // [(nr+1)][16]
  const P *fr;  // [(nrho+1)][8]
  int nr, nrho;
  P rdr, rdrho;
  double cutforcesq, rhomax;
};

// r-table row and fraction, pair_eam.cpp:649-653
template <class P>
__device__ __forceinline__ const P *rrow(const Tabs<P> &t, P r, P *pout) {
  P p = r * t.rdr + P(1.0);
  int m = (int)p;
  m = m < t.nr - 1 ? m : t.nr - 1;
  p -= (P)m;
  p = p < P(1.0) ? p : P(1.0);
  *pout = p;
  return t.rz + 16 * (size_t)m;
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

template <class P>
__device__ __forceinline__ P sqrt_p(double rsq);
template <>
__device__ __forceinline__ double sqrt_p<double>(double rsq) { return sqrt(rsq); }
template <>
__device__ __forceinline__ float sqrt_p<float>(double rsq) { return sqrtf((float)rsq); }

// ------------------------------------------------------------- atoms, EAM
// rho pass + fp pass fused (pair_eam.cpp:629-680); fp stored in x4.w, which
// the force pass reads for ghosts too (forward_comm_pair, :683).
template <class P>
__global__ void k_eam_atom_rho(int n, double *__restrict__ x4, Prd prd, const uint32_t *__restrict__ neigh,
                               const int *__restrict__ num, Tabs<P> t, double *__restrict__ rho_out,
                               double *__restrict__ eatom) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x4[4 * i], yi = x4[4 * i + 1], zi = x4[4 * i + 2];
  const int jn = num[i];
  double rho = 0.0;
  for (int k = 0; k < jn; ++k) {
    uint32_t e = neigh[(size_t)k * n + i];
    int j = (int)(e & 0x7FFFFFFu), s[3];
    img_shift((int)(e >> 27), s);
    double delx = xi - shifted(x4[4 * j], s[0], prd.v[0]);
    double dely = yi - shifted(x4[4 * j + 1], s[1], prd.v[1]);
    double delz = zi - shifted(x4[4 * j + 2], s[2], prd.v[2]);
    double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < t.cutforcesq) {
      P p;
      const P *c = rrow<P>(t, sqrt_p<P>(rsq), &p);
      rho += (double)(((c[3] * p + c[4]) * p + c[5]) * p + c[6]);
    }
  }
  rho_out[i] = rho;
  double fp = (double)embed_fp<P>(t, rho);
  x4[4 * i + 3] = fp;
  if (eatom) eatom[i] = embed_energy<P>(t, rho, fp);
}

// force pass without Newton's third law, pair_eam.cpp:688-745
template <class P>
__global__ void k_eam_atom_force(int n, const double *__restrict__ x4, Prd prd, const uint32_t *__restrict__ neigh,
                                 const int *__restrict__ num, Tabs<P> t, double *__restrict__ f,
                                 double *__restrict__ eatom, double *__restrict__ vatom) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x4[4 * i], yi = x4[4 * i + 1], zi = x4[4 * i + 2];
  const P fpi = (P)x4[4 * i + 3];
  const int jn = num[i];
  double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
  double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
  for (int k = 0; k < jn; ++k) {
    uint32_t e = neigh[(size_t)k * n + i];
    int j = (int)(e & 0x7FFFFFFu), s[3];
    img_shift((int)(e >> 27), s);
    double delx = xi - shifted(x4[4 * j], s[0], prd.v[0]);
    double dely = yi - shifted(x4[4 * j + 1], s[1], prd.v[1]);
    double delz = zi - shifted(x4[4 * j + 2], s[2], prd.v[2]);
    double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < t.cutforcesq) {
      P r = sqrt_p<P>(rsq), p;
      const P *c = rrow<P>(t, r, &p);
      P rhoip = (c[0] * p + c[1]) * p + c[2];
      P z2p = (c[8] * p + c[9]) * p + c[10];
      P z2 = ((c[11] * p + c[12]) * p + c[13]) * p + c[14];
      P recip = P(1.0) / r;
      P phi = z2 * recip;
      P phip = z2p * recip - phi * recip;
      P psip = fpi * rhoip + (P)x4[4 * j + 3] * rhoip + phip;
      double fpair = (double)(-psip * recip);
      fx += delx * fpair;
      fy += dely * fpair;
      fz += delz * fpair;
      if (eatom) ep += 0.5 * (double)phi;
      if (vatom) {
        v0 += 0.5 * delx * delx * fpair;
        v1 += 0.5 * dely * dely * fpair;
        v2 += 0.5 * delz * delz * fpair;
        v3 += 0.5 * delx * dely * fpair;
        v4 += 0.5 * delx * delz * fpair;
        v5 += 0.5 * dely * delz * fpair;
      }
    }
  }
  f[i] = fx;
  f[n + i] = fy;
  f[2 * (size_t)n + i] = fz;
  if (eatom) eatom[i] += ep;
  if (vatom) {
    vatom[i] = v0;
    vatom[(size_t)n + i] = v1;
    vatom[2 * (size_t)n + i] = v2;
    vatom[3 * (size_t)n + i] = v3;
    vatom[4 * (size_t)n + i] = v4;
    vatom[5 * (size_t)n + i] = v5;
  }
}

// ------------------------------------------------------------- atoms, LJ
template <class P>
__global__ void k_lj_atom_force(int n, const double *__restrict__ x4, Prd prd, const uint32_t *__restrict__ neigh,
                                const int *__restrict__ num, LjParams lj, double *__restrict__ f,
                                double *__restrict__ eatom, double *__restrict__ vatom) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x4[4 * i], yi = x4[4 * i + 1], zi = x4[4 * i + 2];
  const P lj1 = (P)lj.lj1, lj2 = (P)lj.lj2, lj3 = (P)lj.lj3, lj4 = (P)lj.lj4;
  const int jn = num[i];
  double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
  double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
  for (int k = 0; k < jn; ++k) {
    uint32_t e = neigh[(size_t)k * n + i];
    int j = (int)(e & 0x7FFFFFFu), s[3];
    img_shift((int)(e >> 27), s);
    double delx = xi - shifted(x4[4 * j], s[0], prd.v[0]);
    double dely = yi - shifted(x4[4 * j + 1], s[1], prd.v[1]);
    double delz = zi - shifted(x4[4 * j + 2], s[2], prd.v[2]);
    double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < lj.cutsq) {
      P r2inv = P(1.0) / (P)rsq;
      P r6inv = r2inv * r2inv * r2inv;
      P forcelj = r6inv * (lj1 * r6inv - lj2);
      double fpair = (double)(P(1.0) * forcelj * r2inv);
      fx += delx * fpair;
      fy += dely * fpair;
      fz += delz * fpair;
      if (eatom) ep += 0.5 * (double)(r6inv * (lj3 * r6inv - lj4));
      if (vatom) {
        v0 += 0.5 * delx * delx * fpair;
        v1 += 0.5 * dely * dely * fpair;
        v2 += 0.5 * delz * delz * fpair;
        v3 += 0.5 * delx * dely * fpair;
        v4 += 0.5 * delx * delz * fpair;
        v5 += 0.5 * dely * delz * fpair;
      }
    }
  }
  f[i] = fx;
  f[n + i] = fy;
  f[2 * (size_t)n + i] = fz;
  if (eatom) eatom[i] = ep;
  if (vatom) {
    vatom[i] = v0;
    vatom[(size_t)n + i] = v1;
    vatom[2 * (size_t)n + i] = v2;
    vatom[3 * (size_t)n + i] = v3;
    vatom[4 * (size_t)n + i] = v4;
    vatom[5 * (size_t)n + i] = v5;
  }
}

// ------------------------------------------------------------- elements
struct ElemArgs {
  int ne, nint, nip;
  const double *nodex;     // [ne][8][3]
  const double *spa_site;  // [nsite][8]
  const double *spa_int;   // [nint][8]
  const double *weight;    // [nint]
  const uint64_t *ipneigh; // [maxnip][nip]
  const int *ipnum;        // [nip]
  int node_list[8];
};

__device__ __forceinline__ void load_shifted_nodes(const double *__restrict__ nodex, int j, int img, const Prd &prd,
                                                   double X[24]) {
  int s[3];
  img_shift(img, s);
  const double *src = nodex + 24 * (size_t)j;
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) X[3 * k + d] = shifted(src[3 * k + d], s[d], prd.v[d]);
}

// Node density, pair_eam.cpp:927-986: erho[i][k] from the list of the
// integration point at node k, measured from the node itself.
template <class P>
__global__ void k_eam_erho(ElemArgs a, Prd prd, Tabs<P> t, double *__restrict__ erho) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 8 * a.ne) return;
  const int i = q >> 3, k = q & 7;
  const double ix = a.nodex[24 * (size_t)i + 3 * k], iy = a.nodex[24 * (size_t)i + 3 * k + 1],
               iz = a.nodex[24 * (size_t)i + 3 * k + 2];
  const int p = i * a.nint + a.node_list[k];
  const int jn = a.ipnum[p];
  double acc = 0.0;
  for (int l = 0; l < jn; ++l) {
    uint64_t e = a.ipneigh[(size_t)l * a.nip + p];
    const int j = (int)(e & 0x7FFFFFFu), img = (int)((e >> 27) & 31), site = (int)(e >> 32);
    double X[24];
    load_shifted_nodes(a.nodex, j, img, prd, X);
    const double *N = a.spa_site + 8 * (size_t)site;
    double jx = 0.0, jy = 0.0, jz = 0.0;
    for (int kk = 0; kk < 8; ++kk) {
      jx += N[kk] * X[3 * kk];
      jy += N[kk] * X[3 * kk + 1];
      jz += N[kk] * X[3 * kk + 2];
    }
    double delx = ix - jx, dely = iy - jy, delz = iz - jz;
    double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < t.cutforcesq) {
      P pp;
      const P *c = rrow<P>(t, sqrt_p<P>(rsq), &pp);
      acc += (double)(((c[3] * pp + c[4]) * pp + c[5]) * pp + c[6]);
    }
  }
  erho[q] = acc;
}

// Integration-point force, pair_eam.cpp:1013-1129.  Writes f_g to ipf
// ([3][nip]) and the weighted energy / virial tallies.
template <class P>
__global__ void k_eam_ip_force(ElemArgs a, Prd prd, Tabs<P> t, const double *__restrict__ erho,
                               double *__restrict__ ipf, double *__restrict__ eip, double *__restrict__ vip) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.nip) return;
  const int i = p / a.nint, g = p % a.nint;
  const double *Ng = a.spa_int + 8 * g;
  double ix = 0.0, iy = 0.0, iz = 0.0, rhoi = 0.0;
  for (int kk = 0; kk < 8; ++kk) {
    ix += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk];
    iy += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk + 1];
    iz += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk + 2];
    rhoi += Ng[kk] * erho[8 * (size_t)i + kk];
  }
  const P fpi = embed_fp<P>(t, rhoi);
  const int jn = a.ipnum[p];
  double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
  double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
  for (int l = 0; l < jn; ++l) {
    uint64_t e = a.ipneigh[(size_t)l * a.nip + p];
    const int j = (int)(e & 0x7FFFFFFu), img = (int)((e >> 27) & 31), site = (int)(e >> 32);
    double X[24];
    load_shifted_nodes(a.nodex, j, img, prd, X);
    const double *N = a.spa_site + 8 * (size_t)site;
    const double *ej = erho + 8 * (size_t)j;
    double jx = 0.0, jy = 0.0, jz = 0.0, rhoj = 0.0;
    for (int kk = 0; kk < 8; ++kk) {
      jx += N[kk] * X[3 * kk];
      jy += N[kk] * X[3 * kk + 1];
      jz += N[kk] * X[3 * kk + 2];
      rhoj += N[kk] * ej[kk];
    }
    double delx = ix - jx, dely = iy - jy, delz = iz - jz;
    double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < t.cutforcesq) {
      P fpj = embed_fp<P>(t, rhoj);
      P r = sqrt_p<P>(rsq), pp;
      const P *c = rrow<P>(t, r, &pp);
      P rhoip = (c[0] * pp + c[1]) * pp + c[2];
      P z2p = (c[8] * pp + c[9]) * pp + c[10];
      P z2 = ((c[11] * pp + c[12]) * pp + c[13]) * pp + c[14];
      P recip = P(1.0) / r;
      P phi = z2 * recip;
      P phip = z2p * recip - phi * recip;
      P psip = (fpi + fpj) * rhoip + phip;
      double fpair = (double)(-psip * recip);
      fx += delx * fpair;
      fy += dely * fpair;
      fz += delz * fpair;
      if (eip) ep += 0.5 * (double)phi;
      if (vip) {
        v0 += 0.5 * delx * delx * fpair;
        v1 += 0.5 * dely * dely * fpair;
        v2 += 0.5 * delz * delz * fpair;
        v3 += 0.5 * delx * dely * fpair;
        v4 += 0.5 * delx * delz * fpair;
        v5 += 0.5 * dely * delz * fpair;
      }
    }
  }
  ipf[p] = fx;
  ipf[(size_t)a.nip + p] = fy;
  ipf[2 * (size_t)a.nip + p] = fz;
  const double w = a.weight[g];
  if (eip) eip[p] = w * (embed_energy<P>(t, rhoi, (double)fpi) + ep);
  if (vip) {
    vip[p] = w * v0;
    vip[(size_t)a.nip + p] = w * v1;
    vip[2 * (size_t)a.nip + p] = w * v2;
    vip[3 * (size_t)a.nip + p] = w * v3;
    vip[4 * (size_t)a.nip + p] = w * v4;
    vip[5 * (size_t)a.nip + p] = w * v5;
  }
}

// LJ integration-point force, pair_lj_cut.cpp:355-422.
template <class P>
__global__ void k_lj_ip_force(ElemArgs a, Prd prd, LjParams lj, double *__restrict__ ipf, double *__restrict__ eip,
                              double *__restrict__ vip) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.nip) return;
  const int i = p / a.nint, g = p % a.nint;
  const double *Ng = a.spa_int + 8 * g;
  double ix = 0.0, iy = 0.0, iz = 0.0;
  for (int kk = 0; kk < 8; ++kk) {
    ix += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk];
    iy += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk + 1];
    iz += Ng[kk] * a.nodex[24 * (size_t)i + 3 * kk + 2];
  }
  const P lj1 = (P)lj.lj1, lj2 = (P)lj.lj2, lj3 = (P)lj.lj3, lj4 = (P)lj.lj4;
  const int jn = a.ipnum[p];
  double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
  double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
  for (int l = 0; l < jn; ++l) {
    uint64_t e = a.ipneigh[(size_t)l * a.nip + p];
    const int j = (int)(e & 0x7FFFFFFu), img = (int)((e >> 27) & 31), site = (int)(e >> 32);
    double X[24];
    load_shifted_nodes(a.nodex, j, img, prd, X);
    const double *N = a.spa_site + 8 * (size_t)site;
    double jx = 0.0, jy = 0.0, jz = 0.0;
    for (int kk = 0; kk < 8; ++kk) {
      jx += N[kk] * X[3 * kk];
      jy += N[kk] * X[3 * kk + 1];
      jz += N[kk] * X[3 * kk + 2];
    }
    double delx = ix - jx, dely = iy - jy, delz = iz - jz;
    double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < lj.cutsq) {
      P r2inv = P(1.0) / (P)rsq;
      P r6inv = r2inv * r2inv * r2inv;
      P forcelj = r6inv * (lj1 * r6inv - lj2);
      double fpair = (double)(forcelj * r2inv);
      fx += delx * fpair;
      fy += dely * fpair;
      fz += delz * fpair;
      if (eip) ep += 0.5 * (double)(r6inv * (lj3 * r6inv - lj4));
      if (vip) {
        v0 += 0.5 * delx * delx * fpair;
        v1 += 0.5 * dely * dely * fpair;
        v2 += 0.5 * delz * delz * fpair;
        v3 += 0.5 * delx * dely * fpair;
        v4 += 0.5 * delx * delz * fpair;
        v5 += 0.5 * dely * delz * fpair;
      }
    }
  }
  ipf[p] = fx;
  ipf[(size_t)a.nip + p] = fy;
  ipf[2 * (size_t)a.nip + p] = fz;
  const double w = a.weight[g];
  if (eip) eip[p] = w * ep;
  if (vip) {
    vip[p] = w * v0;
    vip[(size_t)a.nip + p] = w * v1;
    vip[2 * (size_t)a.nip + p] = w * v2;
    vip[3 * (size_t)a.nip + p] = w * v3;
    vip[4 * (size_t)a.nip + p] = w * v4;
    vip[5 * (size_t)a.nip + p] = w * v5;
  }
}

// Scatter to nodes, pair_eam.cpp:1130-1134 / pair_lj_cut.cpp:424-428:
// nodef[i][k] = sum_g (w_g*N_k(g))*f_g in integration-point order, starting
// from the cleared force (Verlet::node_force_clear, verlet.cpp:131-144).
// One thread per node, no atomics, fixed order.
__global__ void k_scatter(ElemArgs a, const double *__restrict__ ipf, double *__restrict__ nodef) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 8 * a.ne) return;
  const int i = q >> 3, k = q & 7;
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
  t.rz = ctx->eam.rz;
  t.fr = ctx->eam.fr;
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
  t.rz = ctx->eam.rz_f;
  t.fr = ctx->eam.fr_f;
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
  a.nodex = ctx->nodex;
  a.spa_site = ctx->spa_site;
  a.spa_int = ctx->spa_int;
  a.weight = ctx->weight;
  a.ipneigh = ctx->ipneigh;
  a.ipnum = ctx->ipnum;
  for (int k = 0; k < 8; ++k) a.node_list[k] = ctx->node_list[k];
  return a;
}

template <class P>
int eam_atoms_t(cac_ctx *ctx, int eflag, int vflag) {
  const int n = ctx->n;
  Tabs<P> t = tabs_of<P>(ctx);
  hipEvent_t t0;
  timer_begin(ctx, "atom_rho", &t0);
  hipLaunchKernelGGL(k_eam_atom_rho<P>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, ctx->x4, prd_of(ctx),
                     ctx->aneigh, ctx->anum, t, ctx->rho, eflag ? ctx->eatom : nullptr);
  timer_end(ctx, "atom_rho", t0);
  timer_begin(ctx, "atom_force", &t0);
  hipLaunchKernelGGL(k_eam_atom_force<P>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, ctx->x4, prd_of(ctx),
                     ctx->aneigh, ctx->anum, t, ctx->f, eflag ? ctx->eatom : nullptr, vflag ? ctx->vatom : nullptr);
  timer_end(ctx, "atom_force", t0);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

template <class P>
int eam_elements_t(cac_ctx *ctx, int eflag, int vflag) {
  ElemArgs a = elem_args(ctx);
  Tabs<P> t = tabs_of<P>(ctx);
  hipEvent_t t0;
  timer_begin(ctx, "ip_erho", &t0);
  hipLaunchKernelGGL(k_eam_erho<P>, dim3(grid_for(8L * a.ne)), dim3(kBlock), 0, ctx->stream, a, prd_of(ctx), t,
                     ctx->erho);
  timer_end(ctx, "ip_erho", t0);
  timer_begin(ctx, "ip_force", &t0);
  hipLaunchKernelGGL(k_eam_ip_force<P>, dim3(grid_for(a.nip)), dim3(kBlock), 0, ctx->stream, a, prd_of(ctx), t,
                     ctx->erho, ctx->ipf, eflag ? ctx->eip : nullptr, vflag ? ctx->vip : nullptr);
  timer_end(ctx, "ip_force", t0);
  CAC_HIP(ctx, hipGetLastError());
  return launch_scatter(ctx);
}

template <class P>
int lj_atoms_t(cac_ctx *ctx, int eflag, int vflag) {
  const int n = ctx->n;
  hipEvent_t t0;
  timer_begin(ctx, "atom_force", &t0);
  hipLaunchKernelGGL(k_lj_atom_force<P>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, ctx->x4, prd_of(ctx),
                     ctx->aneigh, ctx->anum, ctx->lj, ctx->f, eflag ? ctx->eatom : nullptr,
                     vflag ? ctx->vatom : nullptr);
  timer_end(ctx, "atom_force", t0);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

template <class P>
int lj_elements_t(cac_ctx *ctx, int eflag, int vflag) {
  ElemArgs a = elem_args(ctx);
  hipEvent_t t0;
  timer_begin(ctx, "ip_force", &t0);
  hipLaunchKernelGGL(k_lj_ip_force<P>, dim3(grid_for(a.nip)), dim3(kBlock), 0, ctx->stream, a, prd_of(ctx), ctx->lj,
                     ctx->ipf, eflag ? ctx->eip : nullptr, vflag ? ctx->vip : nullptr);
  timer_end(ctx, "ip_force", t0);
  CAC_HIP(ctx, hipGetLastError());
  return launch_scatter(ctx);
}

}  // namespace

int launch_scatter(cac_ctx *ctx) {
  ElemArgs a = elem_args(ctx);
  hipEvent_t t0;
  timer_begin(ctx, "scatter", &t0);
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(8L * a.ne)), dim3(kBlock), 0, ctx->stream, a, ctx->ipf, ctx->nodef);
  timer_end(ctx, "scatter", t0);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

int launch_eam_atoms(cac_ctx *ctx, int eflag, int vflag, int precision) {
  if (ctx->n == 0) return CAC_OK;
  return precision == CAC_PREC_MIXED ? eam_atoms_t<float>(ctx, eflag, vflag) : eam_atoms_t<double>(ctx, eflag, vflag);
}

int launch_eam_elements(cac_ctx *ctx, int eflag, int vflag, int precision) {
  if (ctx->ne == 0) return CAC_OK;
  return precision == CAC_PREC_MIXED ? eam_elements_t<float>(ctx, eflag, vflag)
                                     : eam_elements_t<double>(ctx, eflag, vflag);
}

int launch_lj_atoms(cac_ctx *ctx, int eflag, int vflag, int precision) {
  if (ctx->n == 0) return CAC_OK;
  return precision == CAC_PREC_MIXED ? lj_atoms_t<float>(ctx, eflag, vflag) : lj_atoms_t<double>(ctx, eflag, vflag);
}

int launch_lj_elements(cac_ctx *ctx, int eflag, int vflag, int precision) {
  if (ctx->ne == 0) return CAC_OK;
  return precision == CAC_PREC_MIXED ? lj_elements_t<float>(ctx, eflag, vflag)
                                     : lj_elements_t<double>(ctx, eflag, vflag);
}

}  // namespace cac
