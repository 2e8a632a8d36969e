// Centro-symmetry parameter of every atom, Output::centro_atom
// (output.cpp:352-441), called after a neighbor build (verlet.cpp:93-94):
//
//   - the entries of atom i's full list with r^2 < cutneigh^2;
//   - fewer than nnn of them: 0;
//   - else the nnn nearest, the nnn(nnn-1)/2 pair values
//     |x_j + x_k - 2 x_i|^2 over them, and the sum of the nnn/2 smallest.
//
// One thread per atom.  The nnn nearest are kept sorted in registers while
// the list streams by (insertion with the slots unrolled, so every index is
// static), their positions are gathered again once, and the pair values go
// through a second sorted register array of nnn/2 slots.
//
// Order, so that the result is a function of the neighbor set alone: the
// selection ranks by (r^2, entry word), ties broken by the entry word
// (j | img << 27); the smallest pair values are summed in ascending order.
// The reference's quickselect (select2, select; output.cpp:458-554) leaves
// both to its partition order: a tie at the nnn-th distance is resolved
// arbitrarily there, and its sum of the nnn/2 values is a permutation of
// this one (equal to within rounding).
//
// Compiled without contraction, like the reference: (x_j + x_k) - 2 x_i and
// the squared norms round operation by operation.
#include "cac_internal.h"

namespace cac {
namespace {

constexpr int kCentroBlock = 256;
constexpr int kCU = 4;  // list entries fetched together

__device__ __forceinline__ void img_shift(int img, int s[3]) {
  s[0] = img % 3 - 1;
  s[1] = (img / 3) % 3 - 1;
  s[2] = img / 9 - 1;
}

__device__ __forceinline__ double shifted(double x, int s, double prd) {
  return s ? x + (double)s * prd : x;
}

struct Prd3 {
  double v[3];
};

// Position of list entry e (atom j, or active site row j >= n of x4): the
// row, plus s*prd for a periodic image (the coordinate a reference ghost
// carries, atom_vec_atomic.cpp:202-221).
__device__ __forceinline__ void entry_xyz(const double *__restrict__ x4, uint32_t e, const Prd3 &prd, double *px,
                                          double *py, double *pz) {
  const size_t j = e & 0x7FFFFFFu;
  const int img = (int)(e >> 27);
  const double2 xy = reinterpret_cast<const double2 *>(x4)[2 * j];
  const double z = x4[4 * j + 2];
  int s[3];
  img_shift(img, s);
  *px = shifted(xy.x, s[0], prd.v[0]);
  *py = shifted(xy.y, s[1], prd.v[1]);
  *pz = shifted(z, s[2], prd.v[2]);
}

__device__ __forceinline__ bool key_less(double a, uint32_t ea, double b, uint32_t eb) {
  return a < b || (a == b && ea < eb);
}

template <int NNN>
__global__ __launch_bounds__(kCentroBlock) void k_centro(int n, const double *__restrict__ x4, Prd3 prd,
                                                         const uint32_t *__restrict__ neigh, int M,
                                                         const int *__restrict__ num, double cutsq,
                                                         double *__restrict__ centro) {
  constexpr int H = NNN / 2;
  const int i = xcd_block() * kCentroBlock + (int)threadIdx.x;
  if (i >= n) return;
  const double xtmp = x4[4 * (size_t)i], ytmp = x4[4 * (size_t)i + 1], ztmp = x4[4 * (size_t)i + 2];
  double d[NNN];
  uint32_t ek[NNN];
#pragma unroll
  for (int s = 0; s < NNN; ++s) {
    d[s] = INFINITY;
    ek[s] = 0xFFFFFFFFu;
  }
  const int jn = num[i];
  const uint32_t *row = neigh + atom_slot(i, 0, M);
  int cnt = 0;
  // kCU entries at a time: their list words, then their position gathers,
  // are issued before any arithmetic, so several loads are in flight
  for (int k0 = 0; k0 < jn; k0 += kCU) {
    uint32_t eb[kCU];
    double xb[kCU], yb[kCU], zb[kCU];
#pragma unroll
    for (int u = 0; u < kCU; ++u) eb[u] = k0 + u < jn ? row[atom_off(k0 + u)] : 0xFFFFFFFFu;
#pragma unroll
    for (int u = 0; u < kCU; ++u)
      if (eb[u] != 0xFFFFFFFFu) entry_xyz(x4, eb[u], prd, &xb[u], &yb[u], &zb[u]);
#pragma unroll
    for (int u = 0; u < kCU; ++u) {
      if (eb[u] == 0xFFFFFFFFu) continue;
      const uint32_t e = eb[u];
      const double delx = xtmp - xb[u], dely = ytmp - yb[u], delz = ztmp - zb[u];
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (!(rsq < cutsq)) continue;
      ++cnt;
      if (!key_less(rsq, e, d[NNN - 1], ek[NNN - 1])) continue;
      // insert, keeping d ascending by (r^2, entry)
#pragma unroll
      for (int s = NNN - 1; s > 0; --s) {
        if (key_less(rsq, e, d[s - 1], ek[s - 1])) {
          d[s] = d[s - 1];
          ek[s] = ek[s - 1];
        } else if (key_less(rsq, e, d[s], ek[s])) {
          d[s] = rsq;
          ek[s] = e;
        }
      }
      if (key_less(rsq, e, d[0], ek[0])) {
        d[0] = rsq;
        ek[0] = e;
      }
    }
  }
  if (cnt < NNN) {
    centro[i] = 0.0;
    return;
  }
  double px[NNN], py[NNN], pz[NNN];
#pragma unroll
  for (int s = 0; s < NNN; ++s) entry_xyz(x4, ek[s], prd, &px[s], &py[s], &pz[s]);
  double best[H];
#pragma unroll
  for (int h = 0; h < H; ++h) best[h] = INFINITY;
#pragma unroll
  for (int j = 0; j < NNN; ++j) {
#pragma unroll
    for (int k = j + 1; k < NNN; ++k) {
      const double delx = px[j] + px[k] - 2.0 * xtmp;
      const double dely = py[j] + py[k] - 2.0 * ytmp;
      const double delz = pz[j] + pz[k] - 2.0 * ztmp;
      const double v = delx * delx + dely * dely + delz * delz;
      if (!(v < best[H - 1])) continue;
#pragma unroll
      for (int h = H - 1; h > 0; --h) {
        if (v < best[h - 1]) best[h] = best[h - 1];
        else if (v < best[h]) best[h] = v;
      }
      if (v < best[0]) best[0] = v;
    }
  }
  double value = 0.0;
#pragma unroll
  for (int h = 0; h < H; ++h) value += best[h];
  centro[i] = value;
}

// Every site of every element, Element::interpolate_atom_vec
// (element.cpp:1241-1265): x = sum_k N_k(site) X_k in node order from 0.0,
// elements in order, sites in the inter_ele order (x outermost); one thread
// per site, rows [ne*nsite][3].
__global__ __launch_bounds__(kCentroBlock) void k_interp_sites(int ne, int nsite, const double *__restrict__ nodex,
                                                               const double *__restrict__ spa_site,
                                                               double *__restrict__ out) {
  const long q = (long)blockIdx.x * kCentroBlock + threadIdx.x;
  if (q >= (long)ne * nsite) return;
  const long e = q / nsite;
  const int s = (int)(q - e * nsite);
  const double *X = nodex + 24 * e;
  const double *N = spa_site + 8 * (size_t)s;
  double x = 0.0, y = 0.0, z = 0.0;
  for (int k = 0; k < 8; ++k) {
    x += N[k] * X[3 * k];
    y += N[k] * X[3 * k + 1];
    z += N[k] * X[3 * k + 2];
  }
  out[3 * q] = x;
  out[3 * q + 1] = y;
  out[3 * q + 2] = z;
}

}  // namespace

int launch_interp_sites(cac_ctx *ctx, double *dout) {
  const long total = (long)ctx->ne * ctx->nsite;
  if (total == 0) return CAC_OK;
  hipLaunchKernelGGL(k_interp_sites, dim3((unsigned)((total + kCentroBlock - 1) / kCentroBlock)), dim3(kCentroBlock),
                     0, ctx->stream, ctx->ne, ctx->nsite, ctx->nodex, ctx->spa_site, dout);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

// Site rows of the coupled atom lists must hold the current interpolated
// positions; the caller has brought them up to date (atom_list_site_rows).
int launch_centro(cac_ctx *ctx, int nnn, double *dout) {
  const int n = ctx->n;
  if (n == 0) return CAC_OK;
  Prd3 prd;
  for (int d = 0; d < 3; ++d) prd.v[d] = ctx->prd[d];
  const double cutsq = ctx->cutneigh * ctx->cutneigh;  // cutneighmaxsq, output.cpp:374
  const unsigned grid = xcd_grid((long)n, kCentroBlock);
  hipEvent_t t0;
  timer_begin(ctx, "centro", &t0);
  if (nnn == 12)
    hipLaunchKernelGGL(k_centro<12>, dim3(grid), dim3(kCentroBlock), 0, ctx->stream, n, ctx->x4, prd, ctx->aneigh,
                       ctx->maxneigh, ctx->anum, cutsq, dout);
  else
    hipLaunchKernelGGL(k_centro<8>, dim3(grid), dim3(kCentroBlock), 0, ctx->stream, n, ctx->x4, prd, ctx->aneigh,
                       ctx->maxneigh, ctx->anum, cutsq, dout);
  timer_end(ctx, "centro", t0);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

}  // namespace cac
