// k_noalloc.hip -- the EAM atom passes over caller-owned device arrays,
// allocating nothing (cac_eam_atoms_noalloc).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops with fixed memory).  The inputs (positions and a
// full neighbor list) and the outputs (forces, optionally densities) are
// the caller's device arrays; the tables and the stream are the context's,
// set up before.  The one intermediate, F'(rho) of every atom (written by
// the density pass, read for the neighbors in the force pass), lives in
// `scratch`, a fixed device array of this file: at most kNoAllocMaxAtoms
// atoms per call.  Each device holds its own copy; its address is looked up
// once per device when a context is created there (noalloc_init) and
// passed to the kernels.  Every call writes scratch[0, n) before reading it,
// so nothing needs clearing; calls that overlap on one device share the
// array and race.
#include "cac_internal.h"

namespace cac {
namespace {

constexpr int kNoAllocMaxAtoms = 1 << 20;  // 8 MB of F' values per device
constexpr int kMaxDevices = 64;
constexpr int kBlockNA = 256;

static __device__ __align__(256) double scratch[kNoAllocMaxAtoms];

// device address of `scratch` on each device (0 until looked up there)
double *g_scratch[kMaxDevices] = {};

struct NaTabs {
  const double *rv, *rz, *fr;
  int nr, nrho;
  double rdr, rdrho, cutforcesq;
};

struct NaBox {
  double prd[3];
};

__device__ __forceinline__ double na_shift(double x, int img, int d, const NaBox &b) {
  const int s = (d == 0 ? img % 3 : d == 1 ? (img / 3) % 3 : img / 9) - 1;
  return s == 0 ? x : x + (double)s * b.prd[d];
}

// r-table knot and fraction, pair_eam.cpp:649-653
__device__ __forceinline__ int na_rknot(const NaTabs &t, double r, double *pout) {
  double p = r * t.rdr + 1.0;
  int m = (int)p;
  m = m < t.nr - 1 ? m : t.nr - 1;
  p -= (double)m;
  p = p < 1.0 ? p : 1.0;
  *pout = p;
  return m;
}

// PairEAM::compute rho and fp loops (pair_eam.cpp:629-680): rho_i over the
// entries within the force cutoff, then F'(rho_i) with the index clamp
// m in [1, nrho-1], into scratch (and rho into the caller's array if given).
__global__ __launch_bounds__(kBlockNA) void k_na_rho(int n, const double *__restrict__ x, int M,
                                                     const int *__restrict__ num, const uint32_t *__restrict__ nb,
                                                     NaTabs t, NaBox b, double *__restrict__ fp_out,
                                                     double *__restrict__ rho_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x[3 * (size_t)i], yi = x[3 * (size_t)i + 1], zi = x[3 * (size_t)i + 2];
  double rho = 0.0;
  const int jn = num[i];
  for (int k = 0; k < jn; ++k) {
    const uint32_t e = nb[(size_t)i * M + k];
    const size_t j = e & 0x7FFFFFFu;
    const int img = (int)(e >> 27);
    if (j >= (size_t)n || img > 26) continue;  // not an atom of this call: the list is the caller's
    const double delx = xi - na_shift(x[3 * j], img, 0, b);
    const double dely = yi - na_shift(x[3 * j + 1], img, 1, b);
    const double delz = zi - na_shift(x[3 * j + 2], img, 2, b);
    const double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < t.cutforcesq) {
      double p;
      const double *c = t.rv + 4 * (size_t)na_rknot(t, sqrt(rsq), &p);
      rho += ((c[0] * p + c[1]) * p + c[2]) * p + c[3];
    }
  }
  double p = rho * t.rdrho + 1.0;
  int m = (int)p;
  m = m < t.nrho - 1 ? m : t.nrho - 1;
  m = m > 1 ? m : 1;
  p -= (double)m;
  p = p < 1.0 ? p : 1.0;
  const double *c = t.fr + 8 * (size_t)m;
  fp_out[i] = (c[0] * p + c[1]) * p + c[2];
  if (rho_out) rho_out[i] = rho;
}

// PairEAM::compute force loop (pair_eam.cpp:688-745), no Newton's third
// law: psip = fp_i*rho' + fp_j*rho' + phi', f_i += del*(-psip/r).
__global__ __launch_bounds__(kBlockNA) void k_na_force(int n, const double *__restrict__ x, int M,
                                                       const int *__restrict__ num, const uint32_t *__restrict__ nb,
                                                       NaTabs t, NaBox b, const double *__restrict__ fp,
                                                       double *__restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x[3 * (size_t)i], yi = x[3 * (size_t)i + 1], zi = x[3 * (size_t)i + 2];
  const double fpi = fp[i];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  const int jn = num[i];
  for (int k = 0; k < jn; ++k) {
    const uint32_t e = nb[(size_t)i * M + k];
    const size_t j = e & 0x7FFFFFFu;
    const int img = (int)(e >> 27);
    if (j >= (size_t)n || img > 26) continue;  // not an atom of this call: the list is the caller's
    const double delx = xi - na_shift(x[3 * j], img, 0, b);
    const double dely = yi - na_shift(x[3 * j + 1], img, 1, b);
    const double delz = zi - na_shift(x[3 * j + 2], img, 2, b);
    const double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq < t.cutforcesq) {
      const double r = sqrt(rsq);
      double p;
      const double *c = t.rz + 8 * (size_t)na_rknot(t, r, &p);
      const double rhoip = ((3.0 * c[0] * p + 2.0 * c[1]) * p + c[2]) * t.rdr;
      const double z2p = ((3.0 * c[4] * p + 2.0 * c[5]) * p + c[6]) * t.rdr;
      const double z2 = ((c[4] * p + c[5]) * p + c[6]) * p + c[7];
      const double recip = 1.0 / r;
      const double phi = z2 * recip;
      const double phip = z2p * recip - phi * recip;
      const double psip = fpi * rhoip + fp[j] * rhoip + phip;
      const double fpair = -psip * recip;
      fx += delx * fpair;
      fy += dely * fpair;
      fz += delz * fpair;
    }
  }
  f[3 * (size_t)i] = fx;
  f[3 * (size_t)i + 1] = fy;
  f[3 * (size_t)i + 2] = fz;
}

}  // namespace

// Look up this device's address of `scratch` (the device is current).
int noalloc_init(cac_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, CAC_ERR_ARG, "device index out of range");
  if (g_scratch[ctx->device]) return CAC_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
  g_scratch[ctx->device] = static_cast<double *>(p);
  return CAC_OK;
}

}  // namespace cac

using namespace cac;

extern "C" int cac_eam_atoms_noalloc(cac_ctx *ctx, int n, const double *x, int maxneigh, const int *numneigh,
                                     const uint32_t *neigh, double *f, double *rho) {
  if (!ctx) return CAC_ERR_ARG;
  if (n > kNoAllocMaxAtoms)
    return fail(ctx, CAC_ERR_ARG, "cac_eam_atoms_noalloc: " + std::to_string(n) + " atoms exceed the limit of " +
                                      std::to_string(kNoAllocMaxAtoms) + " (fixed scratch of 8 MB per device)");
  if (n < 0 || (n > 0 && (!x || !numneigh || !neigh || !f)) || maxneigh < 0)
    return fail(ctx, CAC_ERR_ARG, "cac_eam_atoms_noalloc: bad arguments");
  if (ctx->pair != CAC_PAIR_EAM) return fail(ctx, CAC_ERR_ARG, "cac_eam_atoms_noalloc: no EAM tables set");
  if (!ctx->box_set) return fail(ctx, CAC_ERR_ARG, "cac_set_box not called");
  if (n == 0) return CAC_OK;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  double *fp = g_scratch[ctx->device];
  if (!fp) return fail(ctx, CAC_ERR_HIP, "cac_eam_atoms_noalloc: scratch address not set up on this device");
  NaTabs t;
  t.rv = ctx->eam.rv;
  t.rz = ctx->eam.rz;
  t.fr = ctx->eam.fr;
  t.nr = ctx->eam.nr;
  t.nrho = ctx->eam.nrho;
  t.rdr = ctx->eam.rdr;
  t.rdrho = ctx->eam.rdrho;
  t.cutforcesq = ctx->eam.cutforcesq;
  NaBox b;
  for (int d = 0; d < 3; ++d) b.prd[d] = ctx->prd[d];
  const unsigned grid = (unsigned)((n + kBlockNA - 1) / kBlockNA);
  hipLaunchKernelGGL(k_na_rho, dim3(grid), dim3(kBlockNA), 0, ctx->stream, n, x, maxneigh, numneigh, neigh, t, b, fp,
                     rho);
  hipLaunchKernelGGL(k_na_force, dim3(grid), dim3(kBlockNA), 0, ctx->stream, n, x, maxneigh, numneigh, neigh, t, b,
                     fp, f);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(ctx, e, "cac_eam_atoms_noalloc launch");
  return CAC_OK;
}
