// k_surrogate.hip -- learned element surrogate on the force path.
//
// Frame of element e: its centroid Xbar and the affine map F_e that best fits
// its nodes, F_e = (sum_k (X_k - Xbar) r0_k^T) G0^-1 with r0_k = X0_k - X0bar
// the reference offsets and G0 = sum_k r0_k r0_k^T.  Features (279, groups of
// three), all deviations from that affine map, so a homogeneous deformation
// of the periodic box gives zero deviations on every element:
//   [0, 24)     own nodes:            (X_k - Xbar) - F_e r0_k
//   [24, 102)   26 neighbor centroids: mi(Xbar_j - Xbar) - F_e d0_m
//   [102, 270)  for each corner k, the 7 neighbor nodes that meet there:
//               mi(X_{j,n} - Xbar) - F_e c0_q
//   [270, 279)  strain F_e - I, row-major
// with mi the minimum image in the current box and d0, c0 the reference
// offsets (minimum image in the reference box).  Each feature is divided by
// in_scale[f] (x below).  Prediction of nodal force c = 3k + d, divided by
// out_scale[c]:
//   b_c + A_c . x + sum_{i<=j} Q_c[ij] z_i z_j + MLP(x)_c
// where z (33) is corner k's local input (own node k, its 7 corner
// neighbors, the strain), A, Q, b are FP32 least-squares coefficients and
// the MLP (BF16 MFMA, k_mlp.hip) is trained on what they leave.
//
// Fallback to the plain integration-point path (classified at every
// neighbor rebuild): interface elements (a mesh neighbor missing, i.e. the
// atomistic region or a free surface) and elements whose Green-Lagrange
// strain from the averaged edge vectors exceeds strain_limit in any
// component.  The plain path then runs on subsets only:
//   F  = fallback elements                (force passes, scatter)
//   R  = F + candidate neighbors of F     (node densities, site F')
//   R2 = R + candidate neighbors of R     (site positions)
// and the surrogate covers S = all elements not in F.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "cac_internal.h"

namespace cac {

const float *mlp_forward_staged(cac_ctx *ctx, int n);
uint16_t *mlp_input_buffer(cac_ctx *ctx, int n, int *ld);
int mlp_out_ld(cac_ctx *ctx);
bool mlp_shape_ok(cac_ctx *ctx, int nin, int nout);

namespace {

constexpr int kBlock = 256;
constexpr int kNFeat = CAC_SURR_NFEAT, kNOut = CAC_SURR_NOUT, kNLoc = 33, kNQuad = CAC_SURR_NQUAD;
constexpr int kGroups = kNFeat / 3;
constexpr int kNFeatCfg = CAC_SURR_NFEAT_CONFIG, kGroupsCfg = kNFeatCfg / 3;
static_assert(kGroupsCfg == 34, "configured inputs: 8 own nodes and 26 neighbor centroids");
static_assert(kNQuad == kNLoc * (kNLoc + 1) / 2, "quadratic term count");
constexpr int kPredElems = 64;  // elements per prediction block (one per lane)

// (neighbor slot, neighbor node) of the 7 nodes meeting corner k, k-major
struct CornerTab {
  signed char m[56], node[56];
};

inline unsigned grid_for(long n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

struct Prd3 {
  double v[3];
};

// nearest periodic image of a separation d along an axis of period prd
// (1/prd precomputed): only the choice of image matters, and d*(1/prd)
// picks the same one as d/prd except at exactly half a period
__device__ __forceinline__ double min_image(double d, double prd, double rprd) {
  return d - prd * rint(d * rprd);
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Reference geometry of an element, deduplicated into classes (all interior
// elements of a structured mesh share one): row of kRefLen doubles
//   [0, 24)    r0   node offsets X0_k - X0bar
//   [24, 102)  d0   neighbor-centroid offsets (minimum image, reference box)
//   [102, 270) c0   corner-neighbor node offsets (same)
//   [270, 279) G0^-1, G0 = sum_k r0_k r0_k^T
constexpr int kRefLen = 279, kRefD0 = 24, kRefC0 = 102, kRefG = 270;

// current centroid of every element (its nodes are contiguous: elepbc moves
// them together); one thread per element and direction
__global__ void k_surr_centroid(int ne, const double *__restrict__ nodex, double *__restrict__ xbar) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * ne) return;
  const int e = q / 3, d = q % 3;
  double acc = 0.0;
  for (int k = 0; k < 8; ++k) acc += nodex[24 * (size_t)e + 3 * k + d];
  xbar[q] = acc / 8;
}

struct FeatArgs {
  const double *nodex, *xbar, *ref;
  const int *nbr, *rcls;
  CornerTab ct;
  Prd3 prd, rprd;
};

// Per surrogate row, built at classification: element, reference class and
// the 26 mesh neighbors, so the feature kernel's data loads depend on one
// index load only.
constexpr int kMeta = 28;

// affine map F (row-major) of element e from its current nodes:
// F = (sum_k (X_k - Xbar) r0_k^T) G0^-1; every lane of the wavefront
// computes it from the same (broadcast) loads
__device__ __forceinline__ void element_frame(int e, const FeatArgs &a, const double *ref, double F[9]) {
  const double *X = a.nodex + 24 * (size_t)e;
  const double *xb = a.xbar + 3 * (size_t)e;
  double M[9] = {};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double u0 = X[3 * k] - xb[0], u1 = X[3 * k + 1] - xb[1], u2 = X[3 * k + 2] - xb[2];
    const double s0 = ref[3 * k], s1 = ref[3 * k + 1], s2 = ref[3 * k + 2];
    M[0] += u0 * s0; M[1] += u0 * s1; M[2] += u0 * s2;
    M[3] += u1 * s0; M[4] += u1 * s1; M[5] += u1 * s2;
    M[6] += u2 * s0; M[7] += u2 * s1; M[8] += u2 * s2;
  }
  const double *G = ref + kRefG;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) F[3 * i + j] = M[3 * i] * G[j] + M[3 * i + 1] * G[3 + j] + M[3 * i + 2] * G[6 + j];
}

// feature group g of element e: g < 8 own node g, g < 34 neighbor centroid
// g - 8, g < 90 corner node q = g - 34, else strain row g - 90; emit(c, v)
// receives feature c = 3g + d
template <class Emit>
__device__ __forceinline__ void feature_group(int e, int g, const FeatArgs &a, const int *nb, const double *ref,
                                              const double F[9], Emit emit) {
  const double *xb = a.xbar + 3 * (size_t)e;
  double v[3];
  const double *r = nullptr;
  bool have = true;
  if (g < 8) {
    for (int d = 0; d < 3; ++d) v[d] = a.nodex[24 * (size_t)e + 3 * g + d] - xb[d];
    r = ref + 3 * g;
  } else if (g < 34) {
    const int m = g - 8, j = nb[m];
    have = j >= 0;
    if (have)
      for (int d = 0; d < 3; ++d) v[d] = min_image(a.xbar[3 * (size_t)j + d] - xb[d], a.prd.v[d], a.rprd.v[d]);
    r = ref + kRefD0 + 3 * m;
  } else if (g < 90) {
    const int q = g - 34, j = nb[a.ct.m[q]];
    have = j >= 0;
    if (have)
      for (int d = 0; d < 3; ++d)
        v[d] = min_image(a.nodex[24 * (size_t)j + 3 * a.ct.node[q] + d] - xb[d], a.prd.v[d], a.rprd.v[d]);
    r = ref + kRefC0 + 3 * q;
  } else {
    const int i = g - 90;
    for (int d = 0; d < 3; ++d) emit(3 * g + d, (float)(F[3 * i + d] - (i == d ? 1.0 : 0.0)));
    return;
  }
  for (int d = 0; d < 3; ++d)
    emit(3 * g + d, have ? (float)(v[d] - (F[3 * d] * r[0] + F[3 * d + 1] * r[1] + F[3 * d + 2] * r[2])) : 0.0f);
}

// normalized features of the surrogate elements: BF16 into the MLP input,
// FP32 into xf for the least-squares part; one wavefront per element.  The
// nodes of the element and of its 26 mesh neighbors (27 contiguous 192-byte
// rows) are first copied to LDS with coalesced loads and their centroids
// formed there (the node mean in node order, as k_surr_centroid forms it), so
// every feature group then reads LDS: one round of global loads per element
// instead of chains of dependent gathers.  Same values as feature_group.
constexpr int kFeatWaves = kBlock / 64;

// NG = kGroups: all 279 features, BF16 into the MLP input and FP32 into xf;
// NG = kGroupsCfg: the first 102 (own nodes and neighbor centroids), BF16
// only (the configured network, which has no least-squares part).
template <int NG>
__global__ __launch_bounds__(kBlock) void k_surr_features(int ns, const int *__restrict__ meta, FeatArgs a,
                                                          const float *__restrict__ inv_in, uint16_t *__restrict__ x,
                                                          int ldx, float *__restrict__ xf) {
  __shared__ double sn[kFeatWaves][27 * 24];
  __shared__ double sc[kFeatWaves][27 * 3];
  __shared__ signed char ctm[56], ctn[56];
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  if (threadIdx.x < 56) {
    ctm[threadIdx.x] = a.ct.m[threadIdx.x];
    ctn[threadIdx.x] = a.ct.node[threadIdx.x];
  }
  // wave-uniform row (the meta row and reference class are read with scalar
  // loads); rows past ns stage nothing and store nothing
  const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * kFeatWaves + w);
  const bool live = r < ns;
  const int *mr = meta + kMeta * (size_t)(live ? r : 0);
  const int e = mr[0];
  const int *nb = mr + 2;
  double *N = sn[w], *C = sc[w];
  if (live)
    for (int q = lane; q < 27 * 24; q += 64) {
      const int sl = q / 24;
      const int j = sl == 0 ? e : nb[sl - 1];
      N[q] = j >= 0 ? a.nodex[24 * (size_t)j + (q - 24 * sl)] : 0.0;
    }
  __syncthreads();
  if (live)
    for (int q = lane; q < 27 * 3; q += 64) {
      const int sl = q / 3, d = q - 3 * sl;
      double acc = 0.0;
      for (int k = 0; k < 8; ++k) acc += N[24 * sl + 3 * k + d];
      C[q] = acc / 8;
    }
  __syncthreads();
  if (!live) return;
  const double *ref = a.ref + kRefLen * (size_t)mr[1];
  // affine map of the element (element_frame with the staged nodes)
  double F[9];
  {
    double M[9] = {};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double u0 = N[3 * k] - C[0], u1 = N[3 * k + 1] - C[1], u2 = N[3 * k + 2] - C[2];
      const double s0 = ref[3 * k], s1 = ref[3 * k + 1], s2 = ref[3 * k + 2];
      M[0] += u0 * s0; M[1] += u0 * s1; M[2] += u0 * s2;
      M[3] += u1 * s0; M[4] += u1 * s1; M[5] += u1 * s2;
      M[6] += u2 * s0; M[7] += u2 * s1; M[8] += u2 * s2;
    }
    const double *G = ref + kRefG;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        F[3 * i + j] = M[3 * i] * G[j] + M[3 * i + 1] * G[3 + j] + M[3 * i + 2] * G[6 + j];
  }
  uint16_t *row = x + (size_t)r * ldx;
  float *rowf = NG == kGroups ? xf + (size_t)r * kNFeat : nullptr;
  auto emit = [&](int c, float v) {
    v *= inv_in[c];
    row[c] = to_bf16(v);
    if (NG == kGroups) rowf[c] = v;
  };
  for (int g = lane; g < NG; g += 64) {
    double v[3];
    const double *rr = nullptr;
    bool have = true;
    if (g < 8) {
      for (int d = 0; d < 3; ++d) v[d] = N[3 * g + d] - C[d];
      rr = ref + 3 * g;
    } else if (g < 34) {
      const int m = g - 8;
      have = nb[m] >= 0;
      if (have)
        for (int d = 0; d < 3; ++d) v[d] = min_image(C[3 * (1 + m) + d] - C[d], a.prd.v[d], a.rprd.v[d]);
      rr = ref + kRefD0 + 3 * m;
    } else if (g < 90) {
      const int q = g - 34, m = ctm[q];
      have = nb[m] >= 0;
      if (have)
        for (int d = 0; d < 3; ++d)
          v[d] = min_image(N[24 * (1 + m) + 3 * ctn[q] + d] - C[d], a.prd.v[d], a.rprd.v[d]);
      rr = ref + kRefC0 + 3 * q;
    } else {
      const int i = g - 90;
      for (int d = 0; d < 3; ++d) emit(3 * g + d, (float)(F[3 * i + d] - (i == d ? 1.0 : 0.0)));
      continue;
    }
    for (int d = 0; d < 3; ++d)
      emit(3 * g + d, have ? (float)(v[d] - (F[3 * d] * rr[0] + F[3 * d + 1] * rr[1] + F[3 * d + 2] * rr[2])) : 0.0f);
  }
}

// raw FP32 features of every element (dataset generation, tests)
__global__ void k_surr_features_all(int ne, FeatArgs a, float *__restrict__ out) {
  const int e = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);
  if (e >= ne) return;
  float *row = out + kNFeat * (size_t)e;
  const double *ref = a.ref + kRefLen * (size_t)a.rcls[e];
  double F[9];
  element_frame(e, a, ref, F);
  for (int g = threadIdx.x & 63; g < kGroups; g += 64)
    feature_group(e, g, a, a.nbr + 26 * (size_t)e, ref, F, [&](int c, float v) { row[c] = v; });
}

// local input i of corner k: own node k, its 7 corner neighbors, the strain
__device__ __forceinline__ int local_feature(int k, int i) {
  return i < 3 ? 3 * k + i : (i < 24 ? 102 + 21 * k + (i - 3) : 270 + (i - 24));
}

// nodal forces of the surrogate elements.  Block: 64 elements (one per
// lane) x 8 nodes (one per wavefront).  The elements' features are staged in
// LDS; the node index is wave-uniform, so the least-squares coefficients are
// read with wave-uniform addresses from global memory (scalar loads: the
// multiply-adds take them straight from scalar registers).  (Staging the
// coefficients in LDS as well measured slower: the 149 KB then allow one
// block per CU.)
constexpr int kLdAq = 280, kLdQq = 564;  // padded coefficient rows (floats)

__global__ void __launch_bounds__(512) k_surr_predict(int ns, const int *__restrict__ elemS,
                                                      const float *__restrict__ xf, const float *__restrict__ y,
                                                      int ldy, const float *__restrict__ A,
                                                      const float *__restrict__ Q, const float *__restrict__ b,
                                                      const float *__restrict__ out_scale,
                                                      double *__restrict__ nodef) {
  __shared__ __align__(16) float xs[kPredElems * kNFeat];
  const int r0 = blockIdx.x * kPredElems;
  const int nrow = min(kPredElems, ns - r0);
  {
    // 64 rows of 279 floats start on a 16-byte boundary; the last block's
    // tail is copied by scalars
    constexpr int kX4 = kPredElems * kNFeat / 4;
    constexpr int kPer = (kX4 + 511) / 512;
    const int nx4 = nrow * kNFeat / 4;
    const float4 *sx = reinterpret_cast<const float4 *>(xf + (size_t)r0 * kNFeat);
    float4 v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = threadIdx.x + 512 * u;
      if (i < nx4) v[u] = sx[i];
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = threadIdx.x + 512 * u;
      if (i < nx4) reinterpret_cast<float4 *>(xs)[i] = v[u];
    }
    for (int i = 4 * nx4 + (int)threadIdx.x; i < nrow * kNFeat; i += 512) xs[i] = xf[(size_t)r0 * kNFeat + i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int k = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (lane >= nrow) return;
  const float *x = xs + lane * kNFeat;
  const float *A0 = A + 3 * k * kLdAq, *A1 = A0 + kLdAq, *A2 = A1 + kLdAq;
  float a0 = b[3 * k], a1 = b[3 * k + 1], a2 = b[3 * k + 2];
#pragma unroll 8
  for (int f = 0; f < kNFeat; ++f) {
    const float xv = x[f];
    a0 = fmaf(A0[f], xv, a0);
    a1 = fmaf(A1[f], xv, a1);
    a2 = fmaf(A2[f], xv, a2);
  }
  float z[kNLoc];
#pragma unroll
  for (int i = 0; i < kNLoc; ++i) z[i] = x[local_feature(k, i)];
  const float *Q0 = Q + 3 * k * kLdQq, *Q1 = Q0 + kLdQq, *Q2 = Q1 + kLdQq;
  int t = 0;
#pragma unroll
  for (int i = 0; i < kNLoc; ++i)
#pragma unroll
    for (int j = i; j < kNLoc; ++j, ++t) {
      const float p = z[i] * z[j];
      a0 = fmaf(Q0[t], p, a0);
      a1 = fmaf(Q1[t], p, a1);
      a2 = fmaf(Q2[t], p, a2);
    }
  const int r = r0 + lane;
  const float *yr = y + (size_t)r * ldy + 3 * k;
  double *out = nodef + kNOut * (size_t)elemS[r] + 3 * k;
  out[0] = ((double)a0 + (double)yr[0]) * (double)out_scale[3 * k];
  out[1] = ((double)a1 + (double)yr[1]) * (double)out_scale[3 * k + 1];
  out[2] = ((double)a2 + (double)yr[2]) * (double)out_scale[3 * k + 2];
}

// nodal forces of the surrogate elements from the configured network
// alone: force c of row r = MLP(x)_c * out_scale[c]
__global__ void k_surr_output(int ns, const int *__restrict__ elemS, const float *__restrict__ y, int ldy,
                              const float *__restrict__ out_scale, double *__restrict__ nodef) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ns * kNOut) return;
  const int r = q / kNOut, c = q - kNOut * r;
  nodef[kNOut * (size_t)elemS[r] + c] = (double)y[(size_t)r * ldy + c] * (double)out_scale[c];
}

// fallback class: 1 interface (a mesh neighbor is not an element, or, with
// atom-element coupling, an atom is in one of the element's lists or one of
// its sites in an atom's list), 2 strained, 0 surrogate
__global__ void k_surr_classify(int ne, const double *__restrict__ nodex, const double *__restrict__ ref,
                                const int *__restrict__ rcls, const int *__restrict__ nbr,
                                const int *__restrict__ touch, double strain_limit, int *__restrict__ cls) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const double *r0 = ref + kRefLen * (size_t)rcls[e];
  int c = (touch && touch[e]) ? 1 : 0;
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
          const int k0 = 3 * pr[a][q][0] + d, k1 = 3 * pr[a][q][1] + d;
          s += nodex[24 * (size_t)e + k1] - nodex[24 * (size_t)e + k0];
          s0 += r0[k1] - r0[k0];
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

// current centroids of all elements, then the feature arguments
int frames(cac_ctx *ctx, FeatArgs *a, bool centroids = true) {
  SurrState &S = ctx->surr;
  const int ne = ctx->ne;
  int rc;
  if ((rc = ensure(ctx, S.b_ubar, 3 * (size_t)ne, &S.ubar))) return rc;
  if (centroids)
    hipLaunchKernelGGL(k_surr_centroid, dim3(grid_for(3L * ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->nodex,
                       S.ubar);
  Prd3 rp;
  for (int d = 0; d < 3; ++d) rp.v[d] = 1.0 / ctx->prd[d];
  *a = FeatArgs{ctx->nodex, S.ubar, S.ref, S.nbr, S.rcls, corner_table(), prd_of(ctx), rp};
  return CAC_OK;
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
  hipLaunchKernelGGL(k_surr_classify, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->nodex, S.ref, S.rcls,
                     S.nbr, ctx->coupled ? ctx->etouch : nullptr, S.strain_limit, cls);
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
  {
    std::vector<int> meta((size_t)kMeta * std::max<size_t>(Sl.size(), 1), 0);
    for (size_t r = 0; r < Sl.size(); ++r) {
      const int e = Sl[r];
      meta[kMeta * r] = e;
      meta[kMeta * r + 1] = S.hrcls[e];
      std::copy(S.hnbr.begin() + 26 * (size_t)e, S.hnbr.begin() + 26 * (size_t)(e + 1), meta.begin() + kMeta * r + 2);
    }
    if ((rc = upload_list(ctx, S.b_meta, meta, &S.meta))) return rc;
  }
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
  hipEvent_t t0;
  timer_begin(ctx, "surr_feat", &t0);
  FeatArgs fa;
  if ((rc = frames(ctx, &fa, false))) return rc;
  int ldx = 0;
  uint16_t *x = mlp_input_buffer(ctx, S.nS, &ldx);
  if (!x) return fail(ctx, CAC_ERR_HIP, "surrogate input buffer");
  const bool full = mlp_shape_ok(ctx, kNFeat, kNOut);  // else the configured 102-input network
  float *xf = nullptr;
  if (full) {
    if ((rc = ensure(ctx, S.b_xf, (size_t)S.nS * kNFeat, &xf))) return rc;
    hipLaunchKernelGGL(k_surr_features<kGroups>, dim3(grid_for(S.nS, kFeatWaves)), dim3(kBlock), 0, ctx->stream,
                       S.nS, S.meta, fa, S.inv_in, x, ldx, xf);
  } else {
    hipLaunchKernelGGL(k_surr_features<kGroupsCfg>, dim3(grid_for(S.nS, kFeatWaves)), dim3(kBlock), 0,
                       ctx->stream, S.nS, S.meta, fa, S.inv_in, x, ldx, xf);
  }
  timer_end(ctx, "surr_feat", t0);
  timer_begin(ctx, "surr_mlp", &t0);
  const float *y = mlp_forward_staged(ctx, S.nS);
  timer_end(ctx, "surr_mlp", t0);
  timer_begin(ctx, "surr_fit", &t0);
  if (full)
    hipLaunchKernelGGL(k_surr_predict, dim3(grid_for(S.nS, kPredElems)), dim3(8 * 64), 0, ctx->stream, S.nS,
                       S.elemS, xf, y, mlp_out_ld(ctx), S.fit_A, S.fit_Q, S.fit_b, S.out_scale, ctx->nodef);
  else
    hipLaunchKernelGGL(k_surr_output, dim3(grid_for((long)S.nS * kNOut)), dim3(kBlock), 0, ctx->stream, S.nS,
                       S.elemS, y, mlp_out_ld(ctx), S.out_scale, ctx->nodef);
  timer_end(ctx, "surr_fit", t0);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

void surrogate_destroy(cac_ctx *ctx) {
  SurrState &S = ctx->surr;
  DevBuf *bufs[] = {&S.b_ref, &S.b_rcls, &S.b_A,    &S.b_Q,  &S.b_b,   &S.b_xf, &S.b_nbr, &S.b_inv_in,
                    &S.b_out, &S.b_cls,  &S.b_mark, &S.b_F,  &S.b_S,   &S.b_R,  &S.b_R2,  &S.b_sR,
                    &S.b_sR2, &S.b_tmp,  &S.b_ubar, &S.b_feat, &S.b_meta};
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
  if ((rc = ensure(ctx, S.b_nbr, 26 * (size_t)ne, &S.nbr)) || (rc = ensure(ctx, S.b_rcls, (size_t)ne, &S.rcls)) ||
      (rc = ensure(ctx, S.b_inv_in, kNFeat, &S.inv_in)) || (rc = ensure(ctx, S.b_out, kNOut, &S.out_scale)) ||
      (rc = ensure(ctx, S.b_A, (size_t)kNOut * kLdAq, &S.fit_A)) ||
      (rc = ensure(ctx, S.b_Q, (size_t)kNOut * kLdQq, &S.fit_Q)) || (rc = ensure(ctx, S.b_b, kNOut, &S.fit_b)))
    return rc;
  CAC_HIP(ctx, hipMemset(S.fit_A, 0, (size_t)kNOut * kLdAq * sizeof(float)));
  CAC_HIP(ctx, hipMemset(S.fit_Q, 0, (size_t)kNOut * kLdQq * sizeof(float)));
  CAC_HIP(ctx, hipMemset(S.fit_b, 0, kNOut * sizeof(float)));
  // reference rows (minimum image in the box set now, the reference box),
  // deduplicated into classes: rows equal to 1e-9 A share one class
  const CornerTab ct = corner_table();
  std::vector<double> cen(3 * (size_t)ne);
  for (int e = 0; e < ne; ++e)
    for (int d = 0; d < 3; ++d) {
      double acc = 0.0;
      for (int k = 0; k < 8; ++k) acc += X0[24 * (size_t)e + 3 * k + d];
      cen[3 * (size_t)e + d] = acc / 8;
    }
  auto mi = [&](double dd, int d) { return dd - ctx->prd[d] * std::rint(dd / ctx->prd[d]); };
  std::vector<double> table;
  std::vector<int> rcls(ne);
  std::map<std::vector<long long>, int> index;
  std::vector<double> row(kRefLen);
  std::vector<long long> key(kRefLen);
  for (int e = 0; e < ne; ++e) {
    std::fill(row.begin(), row.end(), 0.0);
    for (int k = 0; k < 8; ++k)
      for (int d = 0; d < 3; ++d) row[3 * k + d] = X0[24 * (size_t)e + 3 * k + d] - cen[3 * (size_t)e + d];
    for (int m = 0; m < 26; ++m) {
      const int j = nbr[26 * (size_t)e + m];
      if (j < -1 || j >= ne) return fail(ctx, CAC_ERR_ARG, "neighbor table entry out of range");
      if (j >= 0)
        for (int d = 0; d < 3; ++d)
          row[kRefD0 + 3 * m + d] = mi(cen[3 * (size_t)j + d] - cen[3 * (size_t)e + d], d);
    }
    for (int q = 0; q < 56; ++q) {
      const int j = nbr[26 * (size_t)e + ct.m[q]];
      if (j >= 0)
        for (int d = 0; d < 3; ++d)
          row[kRefC0 + 3 * q + d] = mi(X0[24 * (size_t)j + 3 * ct.node[q] + d] - cen[3 * (size_t)e + d], d);
    }
    // G0^-1, G0 = sum_k r0_k r0_k^T (positive definite for a non-degenerate element)
    double G[3][3] = {};
    for (int k = 0; k < 8; ++k)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) G[i][j] += row[3 * k + i] * row[3 * k + j];
    const double det = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) -
                       G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                       G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    if (!(std::fabs(det) > 0)) return fail(ctx, CAC_ERR_ARG, "degenerate reference element");
    double *I = row.data() + kRefG;
    I[0] = (G[1][1] * G[2][2] - G[1][2] * G[2][1]) / det;
    I[1] = (G[0][2] * G[2][1] - G[0][1] * G[2][2]) / det;
    I[2] = (G[0][1] * G[1][2] - G[0][2] * G[1][1]) / det;
    I[3] = (G[1][2] * G[2][0] - G[1][0] * G[2][2]) / det;
    I[4] = (G[0][0] * G[2][2] - G[0][2] * G[2][0]) / det;
    I[5] = (G[0][2] * G[1][0] - G[0][0] * G[1][2]) / det;
    I[6] = (G[1][0] * G[2][1] - G[1][1] * G[2][0]) / det;
    I[7] = (G[0][1] * G[2][0] - G[0][0] * G[2][1]) / det;
    I[8] = (G[0][0] * G[1][1] - G[0][1] * G[1][0]) / det;
    for (int c = 0; c < kRefG; ++c) key[c] = std::llround(row[c] * 1e9);
    for (int c = kRefG; c < kRefLen; ++c) key[c] = std::llround(row[c] * 1e15);
    auto it = index.find(key);
    if (it == index.end()) {
      it = index.emplace(key, (int)(table.size() / kRefLen)).first;
      table.insert(table.end(), row.begin(), row.end());
    }
    rcls[e] = it->second;
  }
  S.nref = (int)(table.size() / kRefLen);
  double *dref;
  if ((rc = ensure(ctx, S.b_ref, table.size(), &dref))) return rc;
  S.ref = dref;
  std::vector<float> inv(kNFeat);
  for (int c = 0; c < kNFeat; ++c) {
    if (!(in_scale[c] > 0)) return fail(ctx, CAC_ERR_ARG, "input scales must be positive");
    inv[c] = 1.0f / in_scale[c];
  }
  CAC_HIP(ctx, hipMemcpy(S.ref, table.data(), table.size() * sizeof(double), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.rcls, rcls.data(), rcls.size() * sizeof(int), hipMemcpyHostToDevice));
  S.hrcls = rcls;
  S.hnbr.assign(nbr, nbr + 26 * (size_t)ne);
  CAC_HIP(ctx, hipMemcpy(S.nbr, nbr, 26 * (size_t)ne * sizeof(int), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.inv_in, inv.data(), kNFeat * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.out_scale, out_scale, kNOut * sizeof(float), hipMemcpyHostToDevice));
  S.strain_limit = strain_limit;
  S.setup = true;
  S.ready = false;
  return CAC_OK;
}

int cac_surrogate_set_fit(cac_ctx *ctx, const float *A, const float *Q, const float *b) {
  if (!ctx || !A || !Q || !b) return CAC_ERR_ARG;
  SurrState &S = ctx->surr;
  if (!S.setup) return fail(ctx, CAC_ERR_ARG, "cac_surrogate_setup not called");
  hipSetDevice(ctx->device);
  // rows padded to 16 bytes (kLdAq, kLdQq floats) for the staging loads
  std::vector<float> ap((size_t)kNOut * kLdAq, 0.0f), qp((size_t)kNOut * kLdQq, 0.0f);
  for (int c = 0; c < kNOut; ++c) {
    std::copy(A + (size_t)c * kNFeat, A + (size_t)(c + 1) * kNFeat, ap.begin() + (size_t)c * kLdAq);
    std::copy(Q + (size_t)c * kNQuad, Q + (size_t)(c + 1) * kNQuad, qp.begin() + (size_t)c * kLdQq);
  }
  CAC_HIP(ctx, hipMemcpy(S.fit_A, ap.data(), ap.size() * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.fit_Q, qp.data(), qp.size() * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(S.fit_b, b, kNOut * sizeof(float), hipMemcpyHostToDevice));
  return CAC_OK;
}

int cac_surrogate_enable(cac_ctx *ctx, int on) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  SurrState &S = ctx->surr;
  if (on) {
    if (!S.setup) return fail(ctx, CAC_ERR_ARG, "cac_surrogate_setup not called");
    if (!ctx->mlp) return fail(ctx, CAC_ERR_ARG, "cac_mlp_create not called");
    if (!mlp_shape_ok(ctx, kNFeat, kNOut) && !mlp_shape_ok(ctx, kNFeatCfg, kNOut))
      return fail(ctx, CAC_ERR_ARG, "the MLP must map 279 (or the configured 102) features to 24 forces");
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
  if ((rc = ensure(ctx, S.b_feat, kNFeat * (size_t)ne, &f)) || (rc = ensure(ctx, S.b_cls, (size_t)ne, &c))) return rc;
  FeatArgs fa;
  if ((rc = frames(ctx, &fa))) return rc;
  hipLaunchKernelGGL(k_surr_features_all, dim3(grid_for(ne, kBlock / 64)), dim3(kBlock), 0, ctx->stream, ne, fa, f);
  hipLaunchKernelGGL(k_surr_classify, dim3(grid_for(ne)), dim3(kBlock), 0, ctx->stream, ne, ctx->nodex, S.ref, S.rcls,
                     S.nbr, (ctx->coupled && ctx->built) ? ctx->etouch : nullptr, S.strain_limit, c);
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
