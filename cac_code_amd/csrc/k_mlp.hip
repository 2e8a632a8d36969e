// k_mlp.hip -- BF16 MFMA multilayer perceptron for the learned element
// surrogate: forward, backward and Adam, all on the device.
//
// Network: in -> H -> ... -> H -> out, SiLU after every hidden layer, BF16
// weights and activations as GEMM operands, FP32 accumulation (MFMA
// v_mfma_f32_32x32x16_bf16 on gfx950), FP32 master weights for Adam.
//
// One GEMM shape serves every product: C[m][n] = sum_k A[m][k] * B[n][k]
// with both operands contiguous in k (C = A * B^T).
//   forward      Z_l  = H_{l-1} W_l^T          A = H_{l-1} [batch][in], B = W_l  [out][in]
//   backward     dH   = dZ_l W_l               A = dZ_l [batch][out],   B = Wt_l [in][out]
//   weight grad  dW_l = dZ_l^T H_{l-1}         A = dZt_l [out][batch],  B = Ht   [in][batch]
// The epilogues write the transposed copies the next product needs.
// Every dimension is padded to a multiple of 32; padded weights, inputs and
// output gradients are zero, so they stay zero.
//
// MFMA 32x32x16 lane layout: lane l supplies A[l%32][8*(l/32) + j] and
// B[8*(l/32) + j][l%32] for j = 0..7 and receives
// C[8*(v/4) + 4*(l/32) + v%4][l%32] in accumulator v = 0..15.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "cac_internal.h"

namespace cac {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16 round to nearest even (explicit, so the CPU oracle can round
// the same way bit for bit).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

struct MlpLayer {
  int in, out;      // true sizes
  int inp, outp;    // padded to 32 (tile edge; the reduction needs 16)
  float *w, *b;     // FP32 master [outp][inp], [outp]
  float *mw, *vw, *mb, *vb;  // Adam moments
  float *gw, *gb;   // gradients
  uint16_t *wb, *wt;         // BF16 [outp][inp], transposed [inp][outp]
};

struct MlpState {
  int nin = 0, nhid = 0, nlayer = 0, nout = 0;  // nlayer = number of hidden layers
  std::vector<MlpLayer> L;
  std::vector<DevBuf> bufs;
  int cap = 0;                       // batch rows allocated (multiple of 32)
  uint16_t *x = nullptr;             // [cap][inp0] input, BF16
  std::vector<float *> Z;            // [cap][outp] pre-activations (hidden)
  std::vector<uint16_t *> H, Ht;     // post-activations [cap][outp], [outp][cap]
  float *y = nullptr;                // [cap][outp_last]
  uint16_t *dz = nullptr, *dzt = nullptr, *dh = nullptr;  // scratch [cap][maxw], [maxw][cap]
  uint16_t *dz2 = nullptr, *dzt2 = nullptr;
  uint16_t *xt = nullptr;            // [inp0][cap]
  float *loss = nullptr;             // [64] partial sums
  long step = 0;
  // resident dataset (FP32, true widths)
  int nd = 0;
  float *dx = nullptr, *dt = nullptr;
  int *perm = nullptr;
  DevBuf b_x, b_y, b_dz, b_dzt, b_dz2, b_dzt2, b_xt, b_loss, b_dx, b_dt, b_perm, b_bx, b_bt;
  std::vector<DevBuf> b_Z, b_H, b_Ht;
};

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

enum Epi { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_BWD = 2, EPI_WGRAD = 3, EPI_INFER_HIDDEN = 4 };

struct EpiArgs {
  const float *bias;     // [N]
  float *zf;             // EPI_HIDDEN: Z out [M][N]; EPI_OUT: Y [M][N]; EPI_WGRAD: dW [M][N]
  uint16_t *hb;          // BF16 out [M][ldh]
  uint16_t *ht;          // BF16 transposed out [N][ldt]
  const float *zprev;    // EPI_BWD: pre-activation of the layer below [M][N]
  int ldh, ldt;
};

// One 32x32 tile of C per wavefront.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_gemm_tn(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
                                                     const uint16_t *__restrict__ B, int ldb, EpiArgs e) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tiles_n = N >> 5;
  const int tile = blockIdx.x * (kBlock / 64) + wave;
  if (tile >= (M >> 5) * tiles_n) return;
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int kofs = 8 * (lane >> 5);
  const uint16_t *ap = A + (size_t)(tm * 32 + (lane & 31)) * lda + kofs;
  const uint16_t *bp = B + (size_t)(tn * 32 + (lane & 31)) * ldb + kofs;
  f32x16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(ap + k0);
    const bf16x8 b = *reinterpret_cast<const bf16x8 *>(bp + k0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const int col = tn * 32 + (lane & 31);
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int row = tm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    const float c = acc[v];
    if (MODE == EPI_HIDDEN || MODE == EPI_INFER_HIDDEN) {
      const float z = c + e.bias[col];
      const float h = z * sigm(z);
      if (MODE == EPI_HIDDEN) {
        e.zf[(size_t)row * N + col] = z;
        e.ht[(size_t)col * e.ldt + row] = f2bf(h);
      }
      e.hb[(size_t)row * e.ldh + col] = f2bf(h);
    } else if (MODE == EPI_OUT) {
      e.zf[(size_t)row * N + col] = c + e.bias[col];
    } else if (MODE == EPI_BWD) {
      const float z = e.zprev[(size_t)row * N + col];
      const float s = sigm(z);
      const float dz = c * (s * (1.0f + z * (1.0f - s)));
      e.hb[(size_t)row * e.ldh + col] = f2bf(dz);
      e.ht[(size_t)col * e.ldt + row] = f2bf(dz);
    } else {  // EPI_WGRAD
      e.zf[(size_t)row * N + col] = c;
    }
  }
}

// Fused inference for the surrogate shape (3 hidden layers of 256, output
// padded to 32, input padded to at most kFusedIn): one block of 256 threads
// carries 64 rows through every layer with the activations in LDS, so only
// the BF16 input rows, the weights (L2-resident) and the FP32 output touch
// memory.  Wave w owns output columns [64w, 64w + 64) of each hidden layer
// (2 x 2 MFMA tiles); waves 0 and 1 own the two 32-row tiles of the output
// layer.  Row strides are padded by 8 BF16 (16 B) against LDS bank
// conflicts of the 16-byte fragment reads.
constexpr int kFusedRows = 64, kFusedH = 256, kFusedIn = 288, kFusedOut = 32;
constexpr int kLdA = kFusedIn + 8, kLdB = kFusedH + 8;

struct FusedW {
  const uint16_t *w[4];
  const float *b[4];
};

__device__ __forceinline__ void fused_layer(const uint16_t *src, int lds, int K, const uint16_t *__restrict__ W,
                                            const float *__restrict__ bias, uint16_t *dst, int ldd, int wave,
                                            int lane) {
  const int kofs = 8 * (lane >> 5), r = lane & 31;
  const uint16_t *a0p = src + r * lds + kofs, *a1p = src + (32 + r) * lds + kofs;
  const int c0 = 64 * wave;
  const uint16_t *b0p = W + (size_t)(c0 + r) * K + kofs, *b1p = W + (size_t)(c0 + 32 + r) * K + kofs;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
  // weight fragments are prefetched two k-steps ahead (K is a multiple of 32)
  bf16x8 b0 = *reinterpret_cast<const bf16x8 *>(b0p), b1 = *reinterpret_cast<const bf16x8 *>(b1p);
  bf16x8 g0 = *reinterpret_cast<const bf16x8 *>(b0p + 16), g1 = *reinterpret_cast<const bf16x8 *>(b1p + 16);
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 n0 = b0, n1 = b1, m0 = g0, m1 = g1;
    if (k0 + 32 < K) {
      n0 = *reinterpret_cast<const bf16x8 *>(b0p + k0 + 32);
      n1 = *reinterpret_cast<const bf16x8 *>(b1p + k0 + 32);
      m0 = *reinterpret_cast<const bf16x8 *>(b0p + k0 + 48);
      m1 = *reinterpret_cast<const bf16x8 *>(b1p + k0 + 48);
    }
    bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(a0p + k0);
    bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(a1p + k0);
    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
    a0 = *reinterpret_cast<const bf16x8 *>(a0p + k0 + 16);
    a1 = *reinterpret_cast<const bf16x8 *>(a1p + k0 + 16);
    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, g0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, g1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, g0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, g1, acc[1][1], 0, 0, 0);
    b0 = n0; b1 = n1; g0 = m0; g1 = m1;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = c0 + 32 * j + r;
      const float bc = bias[col];
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const float z = acc[i][j][v] + bc;
        dst[row * ldd + col] = f2bf(z * sigm(z));
      }
    }
}

__global__ __launch_bounds__(kBlock) void k_mlp_infer_fused(int n, int inp, const uint16_t *__restrict__ x,
                                                             FusedW fw, float *__restrict__ y) {
  __shared__ uint16_t bufA[kFusedRows * kLdA];
  __shared__ uint16_t bufB[kFusedRows * kLdB];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = blockIdx.x * kFusedRows;
  // stage the input rows (16-byte chunks; rows past n are zero)
  const int chunks = inp / 8;
  for (int i = threadIdx.x; i < kFusedRows * chunks; i += kBlock) {
    const int rr = i / chunks, c = i % chunks;
    bf16x8 v;
    if (row0 + rr < n) {
      v = *reinterpret_cast<const bf16x8 *>(x + (size_t)(row0 + rr) * inp + 8 * c);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
    }
    *reinterpret_cast<bf16x8 *>(bufA + rr * kLdA + 8 * c) = v;
  }
  __syncthreads();
  fused_layer(bufA, kLdA, inp, fw.w[0], fw.b[0], bufB, kLdB, wave, lane);
  __syncthreads();
  fused_layer(bufB, kLdB, kFusedH, fw.w[1], fw.b[1], bufA, kLdA, wave, lane);
  __syncthreads();
  fused_layer(bufA, kLdA, kFusedH, fw.w[2], fw.b[2], bufB, kLdB, wave, lane);
  __syncthreads();
  if (wave < 2) {
    const int kofs = 8 * (lane >> 5), r = lane & 31;
    const uint16_t *ap = bufB + (32 * wave + r) * kLdB + kofs;
    const uint16_t *bp = fw.w[3] + (size_t)r * kFusedH + kofs;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    for (int k0 = 0; k0 < kFusedH; k0 += 16)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const bf16x8 *>(ap + k0),
                                                    *reinterpret_cast<const bf16x8 *>(bp + k0), acc, 0, 0, 0);
    const float bc = fw.b[3][r];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int row = row0 + 32 * wave + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
      if (row < n) y[(size_t)row * kFusedOut + r] = acc[v] + bc;
    }
  }
}

// Mean-squared-error gradient of the output layer: dY = 2 (Y - T) / (n*out)
// over the true rows and columns, zero in the padding; per-block partial
// loss sums in fixed order.
__global__ void k_loss(int n, int out, int outp, int cap, const float *__restrict__ y, const float *__restrict__ t,
                       int ldtt, const int *__restrict__ rows, uint16_t *__restrict__ dz, uint16_t *__restrict__ dzt,
                       float *__restrict__ partial) {
  __shared__ float s[kBlock];
  const long total = (long)cap * outp;
  const float scale = 2.0f / ((float)n * (float)out);
  float acc = 0.0f;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
    const int r = (int)(q / outp), c = (int)(q % outp);
    float g = 0.0f;
    if (r < n && c < out) {
      const int src = rows ? rows[r] : r;
      const float d = y[(size_t)r * outp + c] - t[(size_t)src * ldtt + c];
      acc += d * d;
      g = scale * d;
    }
    dz[(size_t)r * outp + c] = f2bf(g);
    dzt[(size_t)c * cap + r] = f2bf(g);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// db[c] = sum over batch rows of dZ[r][c] (from the transposed copy).
__global__ void k_bias_grad(int outp, int cap, const uint16_t *__restrict__ dzt, float *__restrict__ gb) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= outp) return;
  float acc = 0.0f;
  for (int r = 0; r < cap; ++r) acc += bf2f(dzt[(size_t)c * cap + r]);
  gb[c] = acc;
}

// Gather batch rows of the resident dataset into the padded BF16 input and
// its transpose.
__global__ void k_gather_input(int n, int nin, int inp, int cap, const float *__restrict__ src,
                               const int *__restrict__ rows, uint16_t *__restrict__ x, uint16_t *__restrict__ xt) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)cap * inp) return;
  const int r = (int)(q / inp), c = (int)(q % inp);
  float v = 0.0f;
  if (r < n && c < nin) v = src[(size_t)(rows ? rows[r] : r) * nin + c];
  const uint16_t h = f2bf(v);
  x[(size_t)r * inp + c] = h;
  if (xt) xt[(size_t)c * cap + r] = h;
}

// Adam (Kingma & Ba 2015) on one parameter array, then the BF16 copies.
__global__ void k_adam(int rows, int cols, float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2,
                       uint16_t *__restrict__ pb, uint16_t *__restrict__ pt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows * cols) return;
  const float gg = g[q];
  const float mm = b1 * m[q] + (1.0f - b1) * gg;
  const float vv = b2 * v[q] + (1.0f - b2) * gg * gg;
  m[q] = mm;
  v[q] = vv;
  const float w = p[q] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  p[q] = w;
  if (pb) {
    const uint16_t h = f2bf(w);
    pb[q] = h;
    if (pt) pt[(size_t)(q % cols) * rows + q / cols] = h;
  }
}

__global__ void k_to_bf16(int rows, int cols, const float *__restrict__ p, uint16_t *__restrict__ pb,
                          uint16_t *__restrict__ pt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows * cols) return;
  const uint16_t h = f2bf(p[q]);
  pb[q] = h;
  pt[(size_t)(q % cols) * rows + q / cols] = h;
}

inline unsigned grid_for(long n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

template <int MODE>
void gemm(cac_ctx *ctx, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
          const EpiArgs &e) {
  const long tiles = (long)(M / 32) * (N / 32);
  hipLaunchKernelGGL(k_gemm_tn<MODE>, dim3(grid_for(tiles, kBlock / 64)), dim3(kBlock), 0, ctx->stream, M, N, K, A,
                     lda, B, ldb, e);
}

int pad_to(int v, int m) { return ((v + m - 1) / m) * m; }

template <class T>
int alloc(cac_ctx *ctx, MlpState *s, DevBuf &b, size_t n, T **out) {
  int rc = ensure(ctx, b, n, out);
  if (rc) return rc;
  CAC_HIP(ctx, hipMemsetAsync(*out, 0, n * sizeof(T), ctx->stream));
  return CAC_OK;
}

int ensure_batch(cac_ctx *ctx, int n) {
  MlpState *s = ctx->mlp;
  const int cap = pad_to(n > 0 ? n : 1, 32);
  if (cap <= s->cap) return CAC_OK;
  s->cap = cap;
  int maxw = s->L[0].inp;
  for (auto &l : s->L) maxw = std::max(maxw, std::max(l.inp, l.outp));
  int rc;
  if ((rc = alloc(ctx, s, s->b_x, (size_t)cap * s->L[0].inp, &s->x)) ||
      (rc = alloc(ctx, s, s->b_xt, (size_t)cap * s->L[0].inp, &s->xt)) ||
      (rc = alloc(ctx, s, s->b_y, (size_t)cap * s->L.back().outp, &s->y)) ||
      (rc = alloc(ctx, s, s->b_dz, (size_t)cap * maxw, &s->dz)) ||
      (rc = alloc(ctx, s, s->b_dzt, (size_t)cap * maxw, &s->dzt)) ||
      (rc = alloc(ctx, s, s->b_dz2, (size_t)cap * maxw, &s->dz2)) ||
      (rc = alloc(ctx, s, s->b_dzt2, (size_t)cap * maxw, &s->dzt2)) ||
      (rc = alloc(ctx, s, s->b_loss, 64, &s->loss)))
    return rc;
  s->Z.resize(s->nlayer);
  s->H.resize(s->nlayer);
  s->Ht.resize(s->nlayer);
  s->b_Z.resize(s->nlayer);
  s->b_H.resize(s->nlayer);
  s->b_Ht.resize(s->nlayer);
  for (int l = 0; l < s->nlayer; ++l) {
    const int w = s->L[l].outp;
    if ((rc = alloc(ctx, s, s->b_Z[l], (size_t)cap * w, &s->Z[l])) ||
        (rc = alloc(ctx, s, s->b_H[l], (size_t)cap * w, &s->H[l])) ||
        (rc = alloc(ctx, s, s->b_Ht[l], (size_t)cap * w, &s->Ht[l])))
      return rc;
  }
  return CAC_OK;
}

// Forward over `cap` rows of s->x; train = keep Z and transposed H.
void forward(cac_ctx *ctx, bool train) {
  MlpState *s = ctx->mlp;
  const int cap = s->cap;
  const uint16_t *in = s->x;
  int ldin = s->L[0].inp;
  for (int l = 0; l <= s->nlayer; ++l) {
    MlpLayer &L = s->L[l];
    EpiArgs e{};
    e.bias = L.b;
    if (l < s->nlayer) {
      e.zf = s->Z[l];
      e.hb = s->H[l];
      e.ht = s->Ht[l];
      e.ldh = L.outp;
      e.ldt = cap;
      if (train)
        gemm<EPI_HIDDEN>(ctx, cap, L.outp, L.inp, in, ldin, L.wb, L.inp, e);
      else
        gemm<EPI_INFER_HIDDEN>(ctx, cap, L.outp, L.inp, in, ldin, L.wb, L.inp, e);
      in = s->H[l];
      ldin = L.outp;
    } else {
      e.zf = s->y;
      gemm<EPI_OUT>(ctx, cap, L.outp, L.inp, in, ldin, L.wb, L.inp, e);
    }
  }
}

// Backward from dY already in s->dz / s->dzt (output layer): weight and bias
// gradients of every layer.
void backward(cac_ctx *ctx) {
  MlpState *s = ctx->mlp;
  const int cap = s->cap;
  uint16_t *dz = s->dz, *dzt = s->dzt, *dz2 = s->dz2, *dzt2 = s->dzt2;
  for (int l = s->nlayer; l >= 0; --l) {
    MlpLayer &L = s->L[l];
    const uint16_t *hin_t = l > 0 ? s->Ht[l - 1] : s->xt;
    EpiArgs e{};
    e.zf = L.gw;
    gemm<EPI_WGRAD>(ctx, L.outp, L.inp, cap, dzt, cap, hin_t, cap, e);
    hipLaunchKernelGGL(k_bias_grad, dim3(grid_for(L.outp)), dim3(kBlock), 0, ctx->stream, L.outp, cap, dzt, L.gb);
    if (l > 0) {
      EpiArgs b{};
      b.zprev = s->Z[l - 1];
      b.hb = dz2;
      b.ht = dzt2;
      b.ldh = L.inp;
      b.ldt = cap;
      gemm<EPI_BWD>(ctx, cap, L.inp, L.outp, dz, L.outp, L.wt, L.outp, b);
      std::swap(dz, dz2);
      std::swap(dzt, dzt2);
    }
  }
}

void adam(cac_ctx *ctx, float lr) {
  MlpState *s = ctx->mlp;
  s->step++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)s->step), c2 = 1.0f - powf(b2, (float)s->step);
  for (auto &L : s->L) {
    hipLaunchKernelGGL(k_adam, dim3(grid_for((long)L.outp * L.inp)), dim3(kBlock), 0, ctx->stream, L.outp, L.inp, L.w,
                       L.mw, L.vw, L.gw, lr, b1, b2, eps, c1, c2, L.wb, L.wt);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(L.outp)), dim3(kBlock), 0, ctx->stream, L.outp, 1, L.b, L.mb, L.vb, L.gb,
                       lr, b1, b2, eps, c1, c2, (uint16_t *)nullptr, (uint16_t *)nullptr);
  }
}

int loss_and_grad(cac_ctx *ctx, int n, const float *t, int ldt, const int *rows, float *loss_out) {
  MlpState *s = ctx->mlp;
  MlpLayer &L = s->L.back();
  hipLaunchKernelGGL(k_loss, dim3(64), dim3(kBlock), 0, ctx->stream, n, L.out, L.outp, s->cap, s->y, t, ldt, rows,
                     s->dz, s->dzt, s->loss);
  if (loss_out) {
    float part[64];
    CAC_HIP(ctx, hipMemcpyAsync(part, s->loss, sizeof(part), hipMemcpyDeviceToHost, ctx->stream));
    CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double acc = 0.0;
    for (int b = 0; b < 64; ++b) acc += part[b];
    *loss_out = (float)(acc / ((double)n * L.out));
  }
  return CAC_OK;
}

int ready(cac_ctx *ctx) {
  if (!ctx) return CAC_ERR_ARG;
  if (!ctx->mlp) return fail(ctx, CAC_ERR_ARG, "cac_mlp_create not called");
  hipSetDevice(ctx->device);
  return CAC_OK;
}

int upload_input(cac_ctx *ctx, int n, const float *x) {
  MlpState *s = ctx->mlp;
  int rc;
  if ((rc = ensure_batch(ctx, n))) return rc;
  float *dx;  // a batch buffer of its own: the resident dataset stays intact
  if ((rc = ensure(ctx, s->b_bx, (size_t)std::max(n, 1) * s->nin, &dx))) return rc;
  CAC_HIP(ctx, hipMemcpyAsync(dx, x, (size_t)n * s->nin * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_gather_input, dim3(grid_for((long)s->cap * s->L[0].inp)), dim3(kBlock), 0, ctx->stream, n,
                     s->nin, s->L[0].inp, s->cap, dx, (const int *)nullptr, s->x, s->xt);
  return CAC_OK;
}

}  // namespace

void mlp_destroy(cac_ctx *ctx) {
  MlpState *s = ctx->mlp;
  if (!s) return;
  for (auto &b : s->bufs)
    if (b.p) hipFree(b.p);
  DevBuf *own[] = {&s->b_x, &s->b_y, &s->b_dz, &s->b_dzt, &s->b_dz2, &s->b_dzt2, &s->b_xt, &s->b_loss, &s->b_dx,
                   &s->b_dt, &s->b_perm, &s->b_bx, &s->b_bt};
  for (DevBuf *b : own)
    if (b->p) hipFree(b->p);
  for (auto *v : {&s->b_Z, &s->b_H, &s->b_Ht})
    for (auto &b : *v)
      if (b.p) hipFree(b.p);
  delete s;
  ctx->mlp = nullptr;
}

bool fused_shape(const MlpState *s) {
  return s->nlayer == 3 && s->nhid == kFusedH && s->L[0].inp <= kFusedIn && s->L.back().outp == kFusedOut;
}

// Device-side inference on the first n rows staged in ctx->mlp->x (the
// surrogate force path): returns the output buffer [n][outp].
const float *mlp_forward_staged(cac_ctx *ctx, int n) {
  MlpState *s = ctx->mlp;
  if (fused_shape(s)) {
    FusedW fw;
    for (int l = 0; l < 4; ++l) {
      fw.w[l] = s->L[l].wb;
      fw.b[l] = s->L[l].b;
    }
    hipLaunchKernelGGL(k_mlp_infer_fused, dim3(grid_for(n, kFusedRows)), dim3(kBlock), 0, ctx->stream, n,
                       s->L[0].inp, s->x, fw, s->y);
  } else {
    forward(ctx, false);
  }
  return s->y;
}

uint16_t *mlp_input_buffer(cac_ctx *ctx, int n, int *ld) {
  if (ensure_batch(ctx, n)) return nullptr;
  *ld = ctx->mlp->L[0].inp;
  return ctx->mlp->x;
}

int mlp_out_ld(cac_ctx *ctx) { return ctx->mlp->L.back().outp; }

bool mlp_shape_ok(cac_ctx *ctx, int nin, int nout) {
  return ctx->mlp && ctx->mlp->nin == nin && ctx->mlp->nout == nout;
}

}  // namespace cac

using namespace cac;

extern "C" {

int cac_mlp_create(cac_ctx *ctx, int nin, int nhidden, int nlayers, int nout, uint64_t seed) {
  if (!ctx || nin <= 0 || nhidden <= 0 || nlayers < 1 || nout <= 0 || nhidden % 32) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  mlp_destroy(ctx);
  MlpState *s = new MlpState();
  ctx->mlp = s;
  s->nin = nin;
  s->nhid = nhidden;
  s->nlayer = nlayers;
  s->nout = nout;
  s->bufs.resize(10 * (nlayers + 1));
  std::mt19937_64 rng(seed);
  int in = nin;
  for (int l = 0; l <= nlayers; ++l) {
    MlpLayer L{};
    L.in = in;
    L.out = l < nlayers ? nhidden : nout;
    L.inp = pad_to(L.in, 32);
    L.outp = pad_to(L.out, 32);
    const size_t nw = (size_t)L.outp * L.inp;
    DevBuf *b = &s->bufs[10 * l];
    int rc;
    if ((rc = alloc(ctx, s, b[0], nw, &L.w)) || (rc = alloc(ctx, s, b[1], (size_t)L.outp, &L.b)) ||
        (rc = alloc(ctx, s, b[2], nw, &L.mw)) || (rc = alloc(ctx, s, b[3], nw, &L.vw)) ||
        (rc = alloc(ctx, s, b[4], (size_t)L.outp, &L.mb)) || (rc = alloc(ctx, s, b[5], (size_t)L.outp, &L.vb)) ||
        (rc = alloc(ctx, s, b[6], nw, &L.gw)) || (rc = alloc(ctx, s, b[7], (size_t)L.outp, &L.gb)) ||
        (rc = alloc(ctx, s, b[8], nw, &L.wb)) || (rc = alloc(ctx, s, b[9], nw, &L.wt)))
      return rc;
    // Kaiming-uniform init on the true block, zero padding
    std::vector<float> w(nw, 0.0f), bias(L.outp, 0.0f);
    const float bound = sqrtf(6.0f / (float)L.in) * (l < nlayers ? 1.0f : 0.25f);
    std::uniform_real_distribution<float> U(-bound, bound);
    for (int o = 0; o < L.out; ++o)
      for (int i = 0; i < L.in; ++i) w[(size_t)o * L.inp + i] = U(rng);
    CAC_HIP(ctx, hipMemcpy(L.w, w.data(), nw * sizeof(float), hipMemcpyHostToDevice));
    CAC_HIP(ctx, hipMemcpy(L.b, bias.data(), L.outp * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for((long)nw)), dim3(kBlock), 0, ctx->stream, L.outp, L.inp, L.w, L.wb,
                       L.wt);
    s->L.push_back(L);
    in = L.out;
  }
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int cac_mlp_set_weights(cac_ctx *ctx, int layer, const float *w, const float *b) {
  int rc = ready(ctx);
  if (rc) return rc;
  MlpState *s = ctx->mlp;
  if (layer < 0 || layer > s->nlayer || !w || !b) return CAC_ERR_ARG;
  MlpLayer &L = s->L[layer];
  std::vector<float> wp((size_t)L.outp * L.inp, 0.0f), bp(L.outp, 0.0f);
  for (int o = 0; o < L.out; ++o) {
    for (int i = 0; i < L.in; ++i) wp[(size_t)o * L.inp + i] = w[(size_t)o * L.in + i];
    bp[o] = b[o];
  }
  CAC_HIP(ctx, hipMemcpy(L.w, wp.data(), wp.size() * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(L.b, bp.data(), bp.size() * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for((long)wp.size())), dim3(kBlock), 0, ctx->stream, L.outp, L.inp, L.w,
                     L.wb, L.wt);
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int cac_mlp_get_weights(cac_ctx *ctx, int layer, float *w, float *b) {
  int rc = ready(ctx);
  if (rc) return rc;
  MlpState *s = ctx->mlp;
  if (layer < 0 || layer > s->nlayer || !w || !b) return CAC_ERR_ARG;
  MlpLayer &L = s->L[layer];
  std::vector<float> wp((size_t)L.outp * L.inp), bp(L.outp);
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  CAC_HIP(ctx, hipMemcpy(wp.data(), L.w, wp.size() * sizeof(float), hipMemcpyDeviceToHost));
  CAC_HIP(ctx, hipMemcpy(bp.data(), L.b, bp.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int o = 0; o < L.out; ++o) {
    for (int i = 0; i < L.in; ++i) w[(size_t)o * L.in + i] = wp[(size_t)o * L.inp + i];
    b[o] = bp[o];
  }
  return CAC_OK;
}

int cac_mlp_forward(cac_ctx *ctx, int n, const float *x, float *y) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (n <= 0 || !x || !y) return CAC_ERR_ARG;
  if ((rc = upload_input(ctx, n, x))) return rc;
  forward(ctx, false);
  MlpState *s = ctx->mlp;
  MlpLayer &L = s->L.back();
  std::vector<float> out((size_t)s->cap * L.outp);
  CAC_HIP(ctx, hipMemcpyAsync(out.data(), s->y, out.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < L.out; ++c) y[(size_t)r * L.out + c] = out[(size_t)r * L.outp + c];
  return CAC_OK;
}

// Loss and gradients of one batch (no update): grads[l] receives dW_l
// ([out][in]) followed by db_l ([out]) for each layer l.
int cac_mlp_grads(cac_ctx *ctx, int n, const float *x, const float *t, float *loss, float **grads) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (n <= 0 || !x || !t) return CAC_ERR_ARG;
  MlpState *s = ctx->mlp;
  if ((rc = upload_input(ctx, n, x))) return rc;
  float *dt;
  if ((rc = ensure(ctx, s->b_bt, (size_t)n * s->nout, &dt))) return rc;
  CAC_HIP(ctx, hipMemcpyAsync(dt, t, (size_t)n * s->nout * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  forward(ctx, true);
  if ((rc = loss_and_grad(ctx, n, dt, s->nout, nullptr, loss))) return rc;
  backward(ctx);
  if (grads) {
    CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int l = 0; l <= s->nlayer; ++l) {
      if (!grads[l]) continue;
      MlpLayer &L = s->L[l];
      std::vector<float> gw((size_t)L.outp * L.inp), gb(L.outp);
      CAC_HIP(ctx, hipMemcpy(gw.data(), L.gw, gw.size() * sizeof(float), hipMemcpyDeviceToHost));
      CAC_HIP(ctx, hipMemcpy(gb.data(), L.gb, gb.size() * sizeof(float), hipMemcpyDeviceToHost));
      float *o = grads[l];
      for (int r = 0; r < L.out; ++r)
        for (int i = 0; i < L.in; ++i) o[(size_t)r * L.in + i] = gw[(size_t)r * L.inp + i];
      for (int r = 0; r < L.out; ++r) o[(size_t)L.out * L.in + r] = gb[r];
    }
  }
  return CAC_OK;
}

// One Adam step on a host batch.
int cac_mlp_train_step(cac_ctx *ctx, int n, const float *x, const float *t, float lr, float *loss) {
  int rc = cac_mlp_grads(ctx, n, x, t, loss, nullptr);
  if (rc) return rc;
  adam(ctx, lr);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

// Keep a training set resident on the device: x [n][nin], t [n][nout].
int cac_mlp_set_dataset(cac_ctx *ctx, int n, const float *x, const float *t) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (n <= 0 || !x || !t) return CAC_ERR_ARG;
  MlpState *s = ctx->mlp;
  if ((rc = ensure(ctx, s->b_dx, (size_t)n * s->nin, &s->dx)) ||
      (rc = ensure(ctx, s->b_dt, (size_t)n * s->nout, &s->dt)) || (rc = ensure(ctx, s->b_perm, (size_t)n, &s->perm)))
    return rc;
  CAC_HIP(ctx, hipMemcpy(s->dx, x, (size_t)n * s->nin * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(s->dt, t, (size_t)n * s->nout * sizeof(float), hipMemcpyHostToDevice));
  s->nd = n;
  return CAC_OK;
}

// Epochs of minibatch Adam over the resident set (shuffled each epoch with
// a seeded permutation); *loss gets the mean loss of the last epoch.
int cac_mlp_train(cac_ctx *ctx, int epochs, int batch, float lr, uint64_t seed, float *loss) {
  int rc = ready(ctx);
  if (rc) return rc;
  MlpState *s = ctx->mlp;
  if (s->nd <= 0) return fail(ctx, CAC_ERR_ARG, "no resident dataset (cac_mlp_set_dataset)");
  if (epochs <= 0 || batch <= 0) return CAC_ERR_ARG;
  batch = std::min(batch, s->nd);
  if ((rc = ensure_batch(ctx, batch))) return rc;
  std::mt19937_64 rng(seed);
  std::vector<int> perm(s->nd);
  double last = 0.0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < s->nd; ++i) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    CAC_HIP(ctx, hipMemcpyAsync(s->perm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    double acc = 0.0;
    int nb = 0;
    for (int b0 = 0; b0 + batch <= s->nd; b0 += batch) {
      hipLaunchKernelGGL(k_gather_input, dim3(grid_for((long)s->cap * s->L[0].inp)), dim3(kBlock), 0, ctx->stream,
                         batch, s->nin, s->L[0].inp, s->cap, s->dx, s->perm + b0, s->x, s->xt);
      forward(ctx, true);
      const bool last_epoch = (ep == epochs - 1);
      float l = 0.0f;
      if ((rc = loss_and_grad(ctx, batch, s->dt, s->nout, s->perm + b0, last_epoch ? &l : nullptr))) return rc;
      backward(ctx);
      adam(ctx, lr);
      acc += l;
      ++nb;
    }
    last = nb ? acc / nb : 0.0;
    CAC_HIP(ctx, hipGetLastError());
  }
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (loss) *loss = (float)last;
  return CAC_OK;
}

}  // extern "C"
