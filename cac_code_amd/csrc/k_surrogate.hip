// k_surrogate.hip -- learned element surrogate on the force path.
//
// Features of element e (270):
//   [0, 24)    the nodal displacements in the element's frame,
//              (X_k - Xbar) - (X0_k - X0bar) (own centroid displacement removed);
//   [24, 102)  the centroid displacements of the 26 mesh neighbors relative to
//              its own, mi(Xbar_j - Xbar) - mi0(X0bar_j - X0bar);
//   [102, 270) for each corner k, the displacements (same relative form) of
//              the 7 neighbor nodes that meet at that corner: the integration
//              point at corner k sees mostly the sites next to those nodes, so
//              without them the nodal force under independent nodal motion is
//              not a function of the inputs.
// mi / mi0 are the minimum image in the current / reference box, so a
// homogeneous strain of the periodic box gives the same features on every
// element.  Each feature is divided by in_scale[f].  The prediction of the
// 24 nodal forces (divided by out_scale[c]) is A x + b (FP32, fitted by
// least squares) plus the MLP's output.
//
// Fallback to the plain integration-point path (classified at every
// neighbor rebuild): interface elements (a mesh neighbor missing, i.e. the
// atomistic region or a free surface) and elements whose Green-Lagrange
// strain from the averaged edge vectors exceeds strain_limit in any
// component.  The plain path then runs on subsets only:
//   F  = fallback elements                (force passes, scatter)
//   R  = F + candidate neighbors of F     (node densities, site F')
//   R2 = R + candidate neighbors of R     (site positions)
// and the MLP covers S = all elements not in F.
#include <algorithm>
#include <cmath>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "cac_internal.h"

namespace cac {

const float *mlp_forward_staged(cac_ctx *ctx);
uint16_t *mlp_input_buffer(cac_ctx *ctx, int n, int *ld);
int mlp_out_ld(cac_ctx *ctx);
bool mlp_shape_ok(cac_ctx *ctx, int nin, int nout);

namespace {

constexpr int kBlock = 256;
constexpr int kNFeat = 270, kNOut = 24;

// (neighbor slot, neighbor node) of the 7 nodes meeting corner k, k-major
struct CornerTab {
  signed char m[56], node[56];
};

inline unsigned grid_for(long n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

struct Prd3 {
  double v[3];
};

__device__ __forceinline__ double min_image(double d, double prd) {
  return d - prd * rint(d / prd);
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// current centroid of every element (its nodes are contiguous: elepbc moves
// them together)
__global__ void k_surr_centroid(int ne, const double *__restrict__ nodex, double *__restrict__ xbar) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  for (int d = 0; d < 3; ++d) {
    double acc = 0.0;
    for (int k = 0; k < 8; ++k) acc += nodex[24 * (size_t)e + 3 * k + d];
    xbar[3 * (size_t)e + d] = acc / 8;
  }
}

// feature group g of element e (groups of 3 consecutive features, 90 per
// element): g < 8 own node g, g < 34 neighbor centroid g - 8, else corner
// node q = g - 34.  r0 [ne][24] are the reference node offsets X0_k - X0bar,
// d0 [ne][78] the reference neighbor-centroid offsets, c0 [ne][168] the
// reference corner-node offsets; emit(c, v) receives feature c = 3g + d.
constexpr int kGroups = kNFeat / 3;

template <class Emit>
__device__ __forceinline__ void feature_group(int e, int g, const double *__restrict__ nodex,
                                              const double *__restrict__ r0, const double *__restrict__ d0,
                                              const double *__restrict__ c0, const double *__restrict__ xbar,
                                              const int *__restrict__ nbr, const CornerTab &ct, const Prd3 &prd,
                                              Emit emit) {
  const double *xb = xbar + 3 * (size_t)e;
  if (g < 8) {
    for (int d = 0; d < 3; ++d)
      emit(3 * g + d, (float)((nodex[24 * (size_t)e + 3 * g + d] - xb[d]) - r0[24 * (size_t)e + 3 * g + d]));
  } else if (g < 34) {
    const int m = g - 8, j = nbr[26 * (size_t)e + m];
    for (int d = 0; d < 3; ++d)
      emit(3 * g + d, j >= 0 ? (float)(min_image(xbar[3 * (size_t)j + d] - xb[d], prd.v[d]) -
                                       d0[78 * (size_t)e + 3 * m + d])
                             : 0.0f);
  } else {
    const int q = g - 34, j = nbr[26 * (size_t)e + ct.m[q]];
    for (int d = 0; d < 3; ++d)
      emit(3 * g + d, j >= 0 ? (float)(min_image(nodex[24 * (size_t)j + 3 * ct.node[q] + d] - xb[d], prd.v[d]) -
                                       c0[168 * (size_t)e + 3 * q + d])
                             : 0.0f);
  }
}

// normalized features of the surrogate elements: BF16 into the MLP input,
// FP32 into xf for the linear part; one wavefront per element
__global__ void k_surr_features(int ns, const int *__restrict__ elemS, const double *__restrict__ nodex,
                                const double *__restrict__ r0, const double *__restrict__ d0,
                                const double *__restrict__ c0, const double *__restrict__ xbar,
                                const int *__restrict__ nbr, CornerTab ct, Prd3 prd,
                                const float *__restrict__ inv_in, uint16_t *__restrict__ x, int ldx,
                                float *__restrict__ xf) {
  const int r = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);
  if (r >= ns) return;
  const int e = elemS[r];
  uint16_t *row = x + (size_t)r * ldx;
  float *rowf = xf + (size_t)r * kNFeat;
  for (int g = threadIdx.x & 63; g < kGroups; g += 64)
    feature_group(e, g, nodex, r0, d0, c0, xbar, nbr, ct, prd, [&](int c, float v) {
      v *= inv_in[c];
      row[c] = to_bf16(v);
      rowf[c] = v;
    });
}

// raw FP32 features of every element (dataset generation, tests)
__global__ void k_surr_features_all(int ne, const double *__restrict__ nodex, const double *__restrict__ r0,
                                    const double *__restrict__ d0, const double *__restrict__ c0,
                                    const double *__restrict__ xbar, const int *__restrict__ nbr, CornerTab ct,
                                    Prd3 prd, float *__restrict__ out) {
  const int e = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);
  if (e >= ne) return;
  float *row = out + kNFeat * (size_t)e;
  for (int g = threadIdx.x & 63; g < kGroups; g += 64)
    feature_group(e, g, nodex, r0, d0, c0, xbar, nbr, ct, prd, [&](int c, float v) { row[c] = v; });
}

// fallback class: 1 interface, 2 strained, 0 surrogate
__global__ void k_surr_classify(int ne, const double *__restrict__ nodex, const double *__restrict__ X0,
                                const int *__restrict__ nbr, double strain_limit, int *__restrict__ cls) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int c = 0;
  for (int m = 0; m < 26; ++m)
    if (nbr[26 * (size_t)e + m] < 0) c = 1;
  if (!c) {
    // averaged edge vectors along the three parent directions (node pairs
    // 0-1 3-2 4-5 7-6; 0-3 1-2 4-7 5-6; 0-4 1-5 2-6 3-7), current and
    // reference; F = E E0^-1; Green-Lagrange 1/2 (F^T F - I)
    const int pr[3][4][2] = {{{0, 1}, {3, 2}, {4, 5}, {7, 6}}, {{0, 3}, {1, 2}, {4, 7}, {5, 6}},
                             {{0, 4}, {1, 5}, {2, 6}, {3, 7}}};
    double E[3][3], E0[3][3];
    for (int a = 0; a < 3; ++a)
      for (int d = 0; d < 3; ++d) {
        double s = 0.0, s0 = 0.0;
        for (int q = 0; q < 4; ++q) {
          const size_t i0 = 24 * (size_t)e + 3 * pr[a][q][0] + d, i1 = 24 * (size_t)e + 3 * pr[a][q][1] + d;
          s += nodex[i1] - nodex[i0];
          s0 += X0[i1] - X0[i0];
        }
        E[d][a] = s / 4;
        E0[d][a] = s0 / 4;
      }
    // inverse of E0 (columns are the edge vectors)
    const double det = E0[0][0] * (E0[1][1] * E0[2][2] - E0[1][2] * E0[2][1]) -
                       E0[0][1] * (E0[1][0] * E0[2][2] - E0[1][2] * E0[2][0]) +
                       E0[0][2] * (E0[1][0] * E0[2][1] - E0[1][1] * E0[2][0]);
    double I[3][3];
    I[0][0] = (E0[1][1] * E0[2][2] - E0[1][2] * E0[2][1]) / det;
    I[0][1] = (E0[0][2] * E0[2][1] - E0[0][1] * E0[2][2]) / det;
    I[0][2] = (E0[0][1] * E0[1][2] - E0[0][2] * E0[1][1]) / det;
    I[1][0] = (E0[1][2] * E0[2][0] - E0[1][0] * E0[2][2]) / det;
    I[1][1] = (E0[0][0] * E0[2][2] - E0[0][2] * E0[2][0]) / det;
    I[1][2] = (E0[0][2] * E0[1][0] - E0[0][0] * E0[1][2]) / det;
    I[2][0] = (E0[1][0] * E0[2][1] - E0[1][1] * E0[2][0]) / det;
    I[2][1] = (E0[0][1] * E0[2][0] - E0[0][0] * E0[2][1]) / det;
    I[2][2] = (E0[0][0] * E0[1][1] - E0[0][1] * E0[1][0]) / det;
    double F[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double s = 0.0;
        for (int q = 0; q < 3; ++q) s += E[a][q] * I[q][b];
        F[a][b] = s;
      }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double s = 0.0;
        for (int q = 0; q < 3; ++q) s += F[q][a] * F[q][b];
        const double eps = 0.5 * (s - (a == b ? 1.0 : 0.0));
        if (fabs(eps) > strain_limit) c = 2;
      }
  }
  cls[e] = c;
}

// mark every candidate neighbor element of the elements in `src`
__global__ void k_mark_neighbors(int n, const int *__restrict__ src, int ne, const uint32_t *__restrict__ eneigh,
                                 const int *__restrict__ enumb, int *__restrict__ mark) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int i = src[o];
  mark[i] = 1;
  for (int l = 0; l < enumb[i]; ++l) mark[eneigh[(size_t)l * ne + i] & 0x7FFFFFFu] = 1;
}

// nodal forces of the surrogate elements: (A x + b + MLP(x)) * out_scale,
// the linear term in FP32 (one thread per element and output)
__global__ void k_surr_scatter(int ns, const int *__restrict__ elemS, const float *__restrict__ y, int ldy,
                               const float *__restrict__ xf, const float *__restrict__ A,
                               const float *__restrict__ b, const float *__restrict__ out_scale,
                               double *__restrict__ nodef) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)ns * kNOut) return;
  const int r = (int)(q / kNOut), c = (int)(q % kNOut);
  const float *x = xf + (size_t)r * kNFeat;
  const float *a = A + (size_t)c * kNFeat;
  float acc = b[c];
  for (int f = 0; f < kNFeat; ++f) acc = fmaf(a[f], x[f], acc);
  nodef[kNOut * (size_t)elemS[r] + c] = ((double)acc + (double)y[(size_t)r * ldy + c]) * (double)out_scale[c];
}

CornerTab corner_table() {
  // node corner flags (element.cpp:1375-1382 order) and the 26 neighbor
  // offsets (i outermost); neighbor d meets corner k iff each d_i is 0 or
  // 2*c_i - 1; its node there has c'_i = c_i where d_i = 0, else 1 - c_i
  static const int C[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  int offs[26][3], n = 0;
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c)
        if (a || b || c) {
          offs[n][0] = a;
          offs[n][1] = b;
          offs[n][2] = c;
          ++n;
        }
  CornerTab t;
  int q = 0;
  for (int k = 0; k < 8; ++k)
    for (int m = 0; m < 26; ++m) {
      bool meets = true;
      int cp[3];
      for (int i = 0; i < 3; ++i) {
        const int d = offs[m][i];
        if (d != 0 && d != 2 * C[k][i] - 1) meets = false;
        cp[i] = d == 0 ? C[k][i] : 1 - C[k][i];
      }
      if (!meets) continue;
      int node = 0;
      while (!(C[node][0] == cp[0] && C[node][1] == cp[1] && C[node][2] == cp[2])) ++node;
      t.m[q] = (signed char)m;
      t.node[q] = (signed char)node;
      ++q;
    }
  return t;
}

Prd3 prd_of(const cac_ctx *ctx) {
  Prd3 p;
  for (int d = 0; d < 3; ++d) p.v[d] = ctx->prd[d];
  return p;
}

int upload_list(cac_ctx *ctx, DevBuf &b, const std::vector<int> &v, int **out) {
  int rc = ensure(ctx, b, std::max<size_t>(v.size(), 1), out);
  if (rc) return rc;
  if (!v.empty()) CAC_HIP(ctx, hipMemcpy(*out, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
  return CAC_OK;
}

}  // namespace

// Classify elements and build the subset lists (after every rebuild while
// the surrogate is enabled).
int surrogate_classify(cac_ctx *ctx) {
  SurrState &S = ctx->surr;
  const int ne = ctx->ne;
  if (!S.enabled || ne == 0) return CAC_OK;
  int rc;
  int *cls, *mark;
  if ((rc = ensure(ctx, S.b_cls, (size_t)ne, &cls)) || (rc = ensure(ctx, S.b_mark, (size_t)ne, &mark))) return rc;
  hipLaunchKernelGGL(k_surr_classify, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->nodex, S.r0, S.nbr,
                     S.strain_limit, cls);
  std::vector<int> hcls(ne);
  CAC_HIP(ctx, hipMemcpyAsync(hcls.data(), cls, ne * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int> F, Sl;
  S.ninterface = S.nstrained = 0;
  for (int e = 0; e < ne; ++e) {
    if (hcls[e]) {
      F.push_back(e);
      (hcls[e] == 1 ? S.ninterface : S.nstrained)++;
    } else {
      Sl.push_back(e);
    }
  }
  S.nF = (int)F.size();
  S.nS = (int)Sl.size();
  if ((rc = upload_list(ctx, S.b_F, F, &S.elemF)) || (rc = upload_list(ctx, S.b_S, Sl, &S.elemS))) return rc;
  // R = F + neighbors, R2 = R + neighbors (device marking)
  auto expand = [&](const std::vector<int> &src, std::vector<int> &dst, DevBuf &b, int **dptr) -> int {
    int *dsrc;
    int rc2;
    if ((rc2 = upload_list(ctx, S.b_tmp, src, &dsrc))) return rc2;
    CAC_HIP(ctx, hipMemsetAsync(mark, 0, ne * sizeof(int), ctx->stream));
    if (!src.empty())
      hipLaunchKernelGGL(k_mark_neighbors, dim3(grid_for((long)src.size())), dim3(kBlock), 0, ctx->stream,
                         (int)src.size(), dsrc, ne, ctx->eneigh, ctx->enumb, mark);
    std::vector<int> hm(ne);
    CAC_HIP(ctx, hipMemcpyAsync(hm.data(), mark, ne * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dst.clear();
    for (int e = 0; e < ne; ++e)
      if (hm[e]) dst.push_back(e);
    return upload_list(ctx, b, dst, dptr);
  };
  std::vector<int> R, R2;
  if ((rc = expand(F, R, S.b_R, &S.elemR)) || (rc = expand(R, R2, S.b_R2, &S.elemR2))) return rc;
  S.nR = (int)R.size();
  S.nR2 = (int)R2.size();
  // active-site lists of R and R2 (sites are numbered element by element)
  std::vector<int> estart(ne + 1);
  CAC_HIP(ctx, hipMemcpy(estart.data(), ctx->estart, (ne + 1) * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> sR, sR2;
  for (int e : R)
    for (int a = estart[e]; a < estart[e + 1]; ++a) sR.push_back(a);
  for (int e : R2)
    for (int a = estart[e]; a < estart[e + 1]; ++a) sR2.push_back(a);
  S.nsR = (int)sR.size();
  S.nsR2 = (int)sR2.size();
  if ((rc = upload_list(ctx, S.b_sR, sR, &S.siteR)) || (rc = upload_list(ctx, S.b_sR2, sR2, &S.siteR2))) return rc;
  S.ready = true;
  return CAC_OK;
}

// Per step: features of S, MLP forward, nodal forces of S.
int surrogate_forces(cac_ctx *ctx) {
  SurrState &S = ctx->surr;
  if (S.nS == 0) return CAC_OK;
  int rc;
  if ((rc = ensure(ctx, S.b_ubar, 3 * (size_t)ctx->ne, &S.ubar))) return rc;
  hipEvent_t t0;
  timer_begin(ctx, "surr_feat", &t0);
  hipLaunchKernelGGL(k_surr_centroid, dim3(grid_for(ctx->ne)), dim3(kBlock), 0, ctx->stream, ctx->ne, ctx->nodex,
                     S.ubar);
  int ldx = 0;
  uint16_t *x = mlp_input_buffer(ctx, S.nS, &ldx);
  if (!x) return fail(ctx, CAC_ERR_HIP, "surrogate input buffer");
  float *xf;
  if ((rc = ensure(ctx, S.b_xf, (size_t)S.nS * kNFeat, &xf))) return rc;
  hipLaunchKernelGGL(k_surr_features, dim3((S.nS + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0, ctx->stream, S.nS, S.elemS, ctx->nodex,
                     S.r0, S.d0, S.c0, S.ubar, S.nbr, corner_table(), prd_of(ctx), S.inv_in, x, ldx, xf);
  timer_end(ctx, "surr_feat", t0);
  timer_begin(ctx, "surr_mlp", &t0);
  const float *y = mlp_forward_staged(ctx);
  timer_end(ctx, "surr_mlp", t0);
  hipLaunchKernelGGL(k_surr_scatter, dim3(grid_for((long)S.nS * kNOut)), dim3(kBlock), 0, ctx->stream, S.nS,
                     S.elemS, y, mlp_out_ld(ctx), xf, S.lin_A, S.lin_b, S.out_scale, ctx->nodef);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

void surrogate_destroy(cac_ctx *ctx) {
  SurrState &S = ctx->surr;
  DevBuf *bufs[] = {&S.b_X0, &S.b_d0, &S.b_c0, &S.b_A, &S.b_b, &S.b_xf, &S.b_nbr, &S.b_inv_in, &S.b_out, &S.b_cls, &S.b_mark, &S.b_F, &S.b_S, &S.b_R, &S.b_R2,
                    &S.b_sR, &S.b_sR2, &S.b_tmp, &S.b_ubar, &S.b_feat};
  for (DevBuf *b : bufs)
    if (b->p) hipFree(b->p);
}

}  // namespace cac

using namespace cac;

extern "C" {

int cac_surrogate_setup(cac_ctx *ctx, const double *X0, const int *nbr, const float *in_scale,
                        const float *out_scale, double strain_limit) {
  if (!ctx || !X0 || !nbr || !in_scale || !out_scale || !(strain_limit > 0)) return CAC_ERR_ARG;
  if (ctx->ne == 0) return fail(ctx, CAC_ERR_ARG, "upload elements before the surrogate");
  hipSetDevice(ctx->device);
  SurrState &S = ctx->surr;
  const int ne = ctx->ne;
  int rc;
  if ((rc = ensure(ctx, S.b_X0, 24 * (size_t)ne, &S.r0)) || (rc = ensure(ctx, S.b_d0, 78 * (size_t)ne, &S.d0)) ||
      (rc = ensure(ctx, S.b_c0, 168 * (size_t)ne, &S.c0)) || (rc = ensure(ctx, S.b_nbr, 26 * (size_t)ne, &S.nbr)) ||
      (rc = ensure(ctx, S.b_inv_in, kNFeat, &S.inv_in)) || (rc = ensure(ctx, S.b_out, kNOut, &S.out_scale)) ||
      (rc = ensure(ctx, S.b_A, (size_t)kNOut * kNFeat, &S.lin_A)) || (rc = ensure(ctx, S.b_b, kNOut, &S.lin_b)))
    return rc;
  CAC_HIP(ctx, hipMemset(S.lin_A, 0, (size_t)kNOut * kNFeat * sizeof(float)));
  CAC_HIP(ctx, hipMemset(S.lin_b, 0, kNOut * sizeof(float)));
  // reference offsets: nodes from their centroid, neighbor centroids from the
  // element's centroid (minimum image in the box set now, the reference box)
  std::vector<double> cen(3 * (size_t)ne), r0(24 * (size_t)ne), d0(78 * (size_t)ne, 0.0);
  for (int e = 0; e < ne; ++e)
    for (int d = 0; d < 3; ++d) {
      double acc = 0.0;
      for (int k = 0; k < 8; ++k) acc += X0[24 * (size_t)e + 3 * k + d];
      cen[3 * (size_t)e + d] = acc / 8;
      for (int k = 0; k < 8; ++k) r0[24 * (size_t)e + 3 * k + d] = X0[24 * (size_t)e + 3 * k + d] - acc / 8;
    }
  for (int e = 0; e < ne; ++e)
    for (int m = 0; m < 26; ++m) {
      const int j = nbr[26 * (size_t)e + m];
      if (j < -1 || j >= ne) return fail(ctx, CAC_ERR_ARG, "neighbor table entry out of range");
      if (j < 0) continue;
      for (int d = 0; d < 3; ++d) {
        const double dd = cen[3 * (size_t)j + d] - cen[3 * (size_t)e + d];
        d0[78 * (size_t)e + 3 * m + d] = dd - ctx->prd[d] * std::rint(dd / ctx->prd[d]);
      }
    }
  const CornerTab ct = corner_table();
  std::vector<double> c0(168 * (size_t)ne, 0.0);
  for (int e = 0; e < ne; ++e)
    for (int q = 0; q < 56; ++q) {
      const int j = nbr[26 * (size_t)e + ct.m[q]];
      if (j < 0) continue;
      for (int d = 0; d < 3; ++d) {
        const double dd = X0[24 * (size_t)j + 3 * ct.node[q] + d] - cen[3 * (size_t)e + d];
        c0[168 * (size_t)e + 3 * q + d] = dd - ctx->prd[d] * std::rint(dd / ctx->prd[d]);
      }
    }
  std::vector<float> inv(kNFeat);
  for (int c = 0; c < kNFeat; ++c) {
    if (!(in_scale[c] > 0)) return fail(ctx, CAC_ERR_ARG, "input scales must be positive");
    inv[c] = 1.0f / in_scale[c];
  }
  CAC_HIP(ctx, hipMemcpy(S.r0, r0.data(), r0.size() * sizeof(double), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.d0, d0.data(), d0.size() * sizeof(double), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.c0, c0.data(), c0.size() * sizeof(double), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.nbr, nbr, 26 * (size_t)ne * sizeof(int), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.inv_in, inv.data(), kNFeat * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.out_scale, out_scale, kNOut * sizeof(float), hipMemcpyHostToDevice));
  S.strain_limit = strain_limit;
  S.setup = true;
  S.ready = false;
  return CAC_OK;
}

int cac_surrogate_set_linear(cac_ctx *ctx, const float *A, const float *b) {
  if (!ctx || !A || !b) return CAC_ERR_ARG;
  SurrState &S = ctx->surr;
  if (!S.setup) return fail(ctx, CAC_ERR_ARG, "cac_surrogate_setup not called");
  hipSetDevice(ctx->device);
  CAC_HIP(ctx, hipMemcpy(S.lin_A, A, (size_t)kNOut * kNFeat * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.lin_b, b, kNOut * sizeof(float), hipMemcpyHostToDevice));
  return CAC_OK;
}

int cac_surrogate_enable(cac_ctx *ctx, int on) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  SurrState &S = ctx->surr;
  if (on) {
    if (!S.setup) return fail(ctx, CAC_ERR_ARG, "cac_surrogate_setup not called");
    if (!ctx->mlp) return fail(ctx, CAC_ERR_ARG, "cac_mlp_create not called");
    if (!mlp_shape_ok(ctx, kNFeat, kNOut)) return fail(ctx, CAC_ERR_ARG, "the MLP must map 270 features to 24 forces");
    if (ctx->pair != CAC_PAIR_EAM) return fail(ctx, CAC_ERR_ARG, "the surrogate is trained for pair_style eam");
  }
  S.enabled = on != 0;
  S.ready = false;
  if (S.enabled && ctx->built) return surrogate_classify(ctx);
  return CAC_OK;
}

int cac_surrogate_features(cac_ctx *ctx, float *feats, int *cls) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  SurrState &S = ctx->surr;
  if (!S.setup) return fail(ctx, CAC_ERR_ARG, "cac_surrogate_setup not called");
  const int ne = ctx->ne;
  int rc;
  float *f;
  int *c;
  if ((rc = ensure(ctx, S.b_ubar, 3 * (size_t)ne, &S.ubar)) || (rc = ensure(ctx, S.b_feat, kNFeat * (size_t)ne, &f)) ||
      (rc = ensure(ctx, S.b_cls, (size_t)ne, &c)))
    return rc;
  hipLaunchKernelGGL(k_surr_centroid, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->nodex, S.ubar);
  hipLaunchKernelGGL(k_surr_features_all, dim3((ne + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0, ctx->stream, ne, ctx->nodex, S.r0,
                     S.d0, S.c0, S.ubar, S.nbr, corner_table(), prd_of(ctx), f);
  hipLaunchKernelGGL(k_surr_classify, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->nodex, S.r0, S.nbr,
                     S.strain_limit, c);
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (feats) CAC_HIP(ctx, hipMemcpy(feats, f, kNFeat * (size_t)ne * sizeof(float), hipMemcpyDeviceToHost));
  if (cls) CAC_HIP(ctx, hipMemcpy(cls, c, (size_t)ne * sizeof(int), hipMemcpyDeviceToHost));
  return CAC_OK;
}

int cac_surrogate_info(cac_ctx *ctx, int *ninterface, int *nstrained, int *nsurrogate) {
  if (!ctx) return CAC_ERR_ARG;
  const SurrState &S = ctx->surr;
  if (ninterface) *ninterface = S.ninterface;
  if (nstrained) *nstrained = S.nstrained;
  if (nsurrogate) *nsurrogate = S.nS;
  return CAC_OK;
}

}  // extern "C"
