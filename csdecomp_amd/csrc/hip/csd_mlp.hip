// csd_mlp.hip -- the signed-distance surrogate: a small MLP
//   x = normalised q (in_dim <= 32, zero-padded to 32)
//   a_l = bf16(SiLU(W_l a_{l-1} + b_l)), l = 1..4, width 256
//   y = w_5 . a_4 + b_5
// with BF16 weights and activations and FP32 accumulation on the gfx950
// 32x32x16 BF16 MFMA, trained with Adam on FP32 master weights, plus the
// exact FP32 signed-distance labels it is trained on. No counterpart in the
// reference (its checker returns booleans only).
//
// MFMA fragment layout (v_mfma_f32_32x32x16_bf16, one wave): lane l holds
// A[l % 32][8 (l / 32) + j] and B[8 (l / 32) + j][l % 32], j < 8; the
// result D[r][l % 32] for r = 8 (i / 4) + 4 (l / 32) + i % 4, i < 16.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "csd_collide.h"
#include "csd_mlp_internal.h"

namespace csd {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

constexpr int kW = 256;      // hidden width
constexpr int kIn = 32;      // padded input width
constexpr int kHid = 4;      // hidden layers
constexpr int kRowsWave = 32;
constexpr int kBlock = 256;  // 4 waves

// float -> bf16, round to nearest even (NaN kept quiet)
__host__ __device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// float -> bf16 with the gfx950 conversion instruction (round to nearest
// even, as f2bf, for the finite activations of the inference kernel)
__device__ __forceinline__ uint16_t f2bf_hw(float x) {
  return __builtin_bit_cast(uint16_t, (__bf16)x);
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + expf(-z)); }
// SiLU for activations that are rounded to bf16 right after (8-bit
// mantissa): hardware exp2 and reciprocal (about 1 ulp in fp32) instead of
// the accurate expf and IEEE division, 5 instructions instead of ~25. For
// z -> -inf, exp2 overflows to inf and the reciprocal to 0 (SiLU -> -0).
__device__ __forceinline__ float silu_bf(float z) {
  const float e = __builtin_amdgcn_exp2f(-1.44269504088896341f * z);
  return z * __builtin_amdgcn_rcpf(1.f + e);
}
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.f / (1.f + expf(-z));
  return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ f32x16 mfma(const bf16x8 &a, const bf16x8 &b, const f32x16 &c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {  // 16-byte aligned
  return *reinterpret_cast<const bf16x8 *>(p);
}

// ---------------------------------------------------------------------------
// Parameter layout (FP32 master, also the gradient layout):
//   W1 [256][32], b1 [256], W2 [256][256], b2, W3, b3, W4, b4, w5 [256], b5
// BF16 copies: Wb_l [256][K_l] (row = output neuron) and, for l >= 2, the
// transposes Wt_l [K_l][256]; w5b [256].
// ---------------------------------------------------------------------------
struct MlpLayout {
  static constexpr size_t kW1 = 0, kB1 = kW1 + (size_t)kW * kIn;
  static constexpr size_t kW2 = kB1 + kW, kB2 = kW2 + (size_t)kW * kW;
  static constexpr size_t kW3 = kB2 + kW, kB3 = kW3 + (size_t)kW * kW;
  static constexpr size_t kW4 = kB3 + kW, kB4 = kW4 + (size_t)kW * kW;
  static constexpr size_t kW5 = kB4 + kW, kB5 = kW5 + kW;
  static constexpr size_t kTotal = kB5 + 1;
  static constexpr size_t w(int l) { return l == 1 ? kW1 : l == 2 ? kW2 : l == 3 ? kW3 : kW4; }
  static constexpr size_t b(int l) { return l == 1 ? kB1 : l == 2 ? kB2 : l == 3 ? kB3 : kB4; }
  static constexpr int K(int l) { return l == 1 ? kIn : kW; }
};

struct MlpDev {
  const float *p;        // FP32 master parameters
  const uint16_t *wb[kHid + 1];  // [1..4] bf16 weights [256][K_l]
  const uint16_t *w5b;   // [256]
  const float *lo, *scale;  // input normalisation x = (q - lo) * scale - 1
  int in_dim;
};

// ---------------------------------------------------------------------------
// Fused inference: one wave computes 32 rows through all layers. Layer
// outputs stay in registers; each 32x32 output tile goes through a small
// per-wave LDS tile to become the next layer's A fragments. Weight chunks
// (256 rows x 64 k; layer 1: its 32) are staged in LDS once per block for
// its 4 waves; two blocks share a CU (2 waves per SIMD).
// ---------------------------------------------------------------------------
constexpr int kChunk = 64;             // k per weight chunk (layer 1: its 32)
constexpr int kChunkPad = kChunk + 8;  // 144-byte rows: 16-byte aligned, fewer conflicts
constexpr int kTilePad = 32 + 8;
// 256 rows x cw bf16 of W (row stride K) from column k0 into LDS rows of
// kChunkPad: 32 * cw 16-byte pieces, cw / 8 per thread
__device__ __forceinline__ void load_chunk(uint16_t *sw, const uint16_t *W, int K, int k0,
                                           int cw) {
  const int per_row = cw / 8;
#pragma unroll
  for (int r = 0; r < cw / 8; ++r) {
    const int piece = threadIdx.x + r * kBlock;
    const int row = piece / per_row, part = piece % per_row;
    *reinterpret_cast<uint4 *>(sw + row * kChunkPad + part * 8) =
        *reinterpret_cast<const uint4 *>(W + (size_t)row * K + k0 + part * 8);
  }
}
// raw inputs of row r: q[r * sn + k * sj]
struct StridedInput {
  const float *q;
  int64_t sn, sj;
  int in_dim;
  __device__ __forceinline__ void load(int64_t r, float (&x)[kIn]) const {
#pragma unroll
    for (int k = 0; k < kIn; ++k) x[k] = k < in_dim ? q[r * sn + (int64_t)k * sj] : 0.f;
  }
};

// raw inputs of row (slot, i, l): a point between a particle's projection
// onto its segment and the particle (projection as in the bisection)
struct SegmentInput {
  SegmentPoints s;
  __device__ __forceinline__ void load(int64_t row, float (&x)[kIn]) const {
    const int64_t per = (int64_t)s.cand * s.L;
    const int slot = (int)(row / per);
    const int rem = (int)(row - (int64_t)slot * per);
    const int i = rem / s.L, l = rem - i * s.L;
    const int r = s.active[slot];
    const int n = min(s.u[r], s.cand);
    const int ii = n > 0 ? min(i, n - 1) : 0;
    const float *p = s.part + ((size_t)r * s.U + ii) * s.dim;
    const float *q1 = s.ls + (size_t)r * s.dim, *q2 = s.le + (size_t)r * s.dim;
    float a = 0.f, bb = 0.f;
    for (int d = 0; d < s.dim; ++d) {
      a = fmaf(p[d] - q1[d], q2[d] - q1[d], a);
      bb = fmaf(q2[d] - q1[d], q2[d] - q1[d], bb);
    }
    const float t = fminf(fmaxf(a / bb, 0.f), 1.f);
    const float f = (float)(l + 1) / (float)(s.L + 1);
#pragma unroll
    for (int k = 0; k < kIn; ++k) {
      float v = 0.f;
      if (k < s.dim) {
        const float pr = (1.f - t) * q1[k] + t * q2[k];
        v = n > 0 ? fmaf(f, p[k] - pr, pr) : q1[k];
      }
      x[k] = v;
    }
  }
};

template <class In>
__global__ __launch_bounds__(kBlock, 2) void k_mlp_infer(MlpDev M, In in, int64_t n,
                                                      float *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint16_t sw[kW * kChunkPad];
  __shared__ __attribute__((aligned(16))) uint16_t st[4][32 * kTilePad];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * (4 * kRowsWave) + wave * kRowsWave;
  const int hr = lane >> 5, c = lane & 31;

  // input fragments: row c, k = 8 hr + j (k-step 0) and 16 + 8 hr + j (k-step 1)
  bf16x8 af[kW / 16];
  {
    const int64_t r = min(row0 + c, n - 1);
    float raw[kIn];
    in.load(r, raw);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      uint16_t h[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        // lane-dependent feature index as a select between two static ones
        const int k0 = 16 * s + j, k1 = 16 * s + 8 + j;
        const float x0 = k0 < M.in_dim ? (raw[k0] - M.lo[k0]) * M.scale[k0] - 1.f : 0.f;
        const float x1 = k1 < M.in_dim ? (raw[k1] - M.lo[k1]) * M.scale[k1] - 1.f : 0.f;
        h[j] = f2bf(hr ? x1 : x0);
      }
      af[s] = *reinterpret_cast<const bf16x8 *>(h);
    }
  }
  float part[16];
#pragma unroll
  for (int l = 1; l <= kHid; ++l) {
    const int K = l == 1 ? kIn : kW;
    f32x16 acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = f32x16{};
    const int cw = K < kChunk ? K : kChunk;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += cw) {
      __syncthreads();
      load_chunk(sw, M.wb[l], K, k0, cw);
      __syncthreads();
#pragma unroll
      for (int s = 0; s < cw / 16; ++s) {
        const bf16x8 a = af[(k0 >> 4) + s];
#pragma unroll
        for (int t = 0; t < 8; ++t)
          acc[t] = mfma(a, ld8(sw + (t * 32 + c) * kChunkPad + 16 * s + 8 * hr), acc[t]);
      }
    }
    const float *bias = M.p + MlpLayout::b(l);
    if (l < kHid) {
      // bias + SiLU, bf16, then transpose each tile into the next A fragments
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float bc = bias[t * 32 + c];
        uint16_t *tile = st[wave];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int r = 8 * (i >> 2) + 4 * hr + (i & 3);
          tile[r * kTilePad + c] = f2bf_hw(silu_bf(acc[t][i] + bc));
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): tile written (single wave)
        __builtin_amdgcn_wave_barrier();
        af[2 * t] = ld8(tile + c * kTilePad + 8 * hr);
        af[2 * t + 1] = ld8(tile + c * kTilePad + 16 + 8 * hr);
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
      }
    } else {
      // head: y = w5 . bf16(a4) + b5, reduced over the 32 lanes holding a row
#pragma unroll
      for (int i = 0; i < 16; ++i) part[i] = 0.f;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float bc = bias[t * 32 + c];
        const float wc = bf2f(M.w5b[t * 32 + c]);
#pragma unroll
        for (int i = 0; i < 16; ++i)
          part[i] = fmaf(bf2f(f2bf_hw(silu_bf(acc[t][i] + bc))), wc, part[i]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) part[i] += __shfl_xor(part[i], o, 32);
  }
  if (c == 0) {
    const float b5 = M.p[MlpLayout::kB5];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int64_t r = row0 + 8 * (i >> 2) + 4 * hr + (i & 3);
      if (r < n) out[r] = part[i] + b5;
    }
  }
}

// ---------------------------------------------------------------------------
// Training: tiled NT GEMM  C[m][n] = sum_k A[m][k] B[n][k]  (bf16 in, fp32
// out; A [M][lda], B [N][ldb], k contiguous). Block tile 128 x 128, four
// waves of 64 x 64, K in LDS chunks of 32. blockIdx.z splits K: partial sums
// go to C + z * M * ldc. Needs K, kper multiples of 32 and lda, ldb
// multiples of 8.
// ---------------------------------------------------------------------------
constexpr int kGT = 128, kGK = 32, kGP = kGK + 8;

__global__ __launch_bounds__(kBlock) void k_gemm_nt(const uint16_t *__restrict__ A, int lda,
                                                    const uint16_t *__restrict__ B, int ldb,
                                                    float *__restrict__ C, int ldc, int M, int N,
                                                    int K, int kper) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kGT * kGP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kGT * kGP];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
  const int kbeg = blockIdx.z * kper, kend = min(K, kbeg + kper);
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int hr = lane >> 5, c = lane & 31;
  const int lr = t >> 1, lk = (t & 1) * 16;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x16{};
  for (int k0 = kbeg; k0 < kend; k0 += kGK) {
    uint4 a0 = {0, 0, 0, 0}, a1 = a0, b0 = a0, b1 = a0;
    if (m0 + lr < M) {
      const uint4 *p = reinterpret_cast<const uint4 *>(A + (size_t)(m0 + lr) * lda + k0 + lk);
      a0 = p[0];
      a1 = p[1];
    }
    if (n0 + lr < N) {
      const uint4 *p = reinterpret_cast<const uint4 *>(B + (size_t)(n0 + lr) * ldb + k0 + lk);
      b0 = p[0];
      b1 = p[1];
    }
    __syncthreads();
    *reinterpret_cast<uint4 *>(As + lr * kGP + lk) = a0;
    *reinterpret_cast<uint4 *>(As + lr * kGP + lk + 8) = a1;
    *reinterpret_cast<uint4 *>(Bs + lr * kGP + lk) = b0;
    *reinterpret_cast<uint4 *>(Bs + lr * kGP + lk + 8) = b1;
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGK; ks += 16) {
      bf16x8 af[2], bfr[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = ld8(As + (wm + 32 * i + c) * kGP + ks + 8 * hr);
#pragma unroll
      for (int j = 0; j < 2; ++j) bfr[j] = ld8(Bs + (wn + 32 * j + c) * kGP + ks + 8 * hr);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma(af[i], bfr[j], acc[i][j]);
    }
  }
  float *Cz = C + (size_t)blockIdx.z * M * ldc;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = m0 + wm + 32 * i + 8 * (e >> 2) + 4 * hr + (e & 3);
        const int col = n0 + wn + 32 * j + c;
        if (row < M && col < N) Cz[(size_t)row * ldc + col] = acc[i][j][e];
      }
}

// x0 [B][32] bf16 and its transpose [32][B]
__global__ void k_prep_input(const float *__restrict__ q, int B, int in_dim,
                             const float *__restrict__ lo, const float *__restrict__ scale,
                             uint16_t *__restrict__ x0, uint16_t *__restrict__ x0t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * kIn) return;
  const int r = i / kIn, k = i % kIn;
  const float x = k < in_dim ? (q[(size_t)r * in_dim + k] - lo[k]) * scale[k] - 1.f : 0.f;
  const uint16_t h = f2bf(x);
  x0[i] = h;
  x0t[(size_t)k * B + r] = h;
}

// z [B][256] += bias; a = bf16(SiLU(z)) as [B][256] and [256][B]
__global__ void k_bias_silu(float *__restrict__ z, const float *__restrict__ bias, int B,
                            uint16_t *__restrict__ a, uint16_t *__restrict__ at) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * kW) return;
  const int r = i / kW, k = i % kW;
  const float v = z[i] + bias[k];
  z[i] = v;
  const uint16_t h = f2bf(silu(v));
  a[i] = h;
  at[(size_t)k * B + r] = h;
}

// head forward + backward, one row per thread: y = w5 . a4 + b5, loss term,
// dy = 2 (y - t) / B, delta4 = dy w5 SiLU'(z4) (fp32 [B][256]).
__global__ void k_head(const uint16_t *__restrict__ a4, const float *__restrict__ z4,
                       const uint16_t *__restrict__ w5b, const float *__restrict__ p,
                       const float *__restrict__ target, int B, float *__restrict__ y,
                       float *__restrict__ dy, float *__restrict__ d4) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  float s = 0.f;
  for (int k = 0; k < kW; ++k) s = fmaf(bf2f(a4[(size_t)r * kW + k]), bf2f(w5b[k]), s);
  const float yy = s + p[MlpLayout::kB5];
  y[r] = yy;
  const float g = 2.f * (yy - target[r]) / (float)B;
  dy[r] = g;
  for (int k = 0; k < kW; ++k)
    d4[(size_t)r * kW + k] = g * bf2f(w5b[k]) * dsilu(z4[(size_t)r * kW + k]);
}

// delta [B][256] (fp32, in place: dA -> dA * SiLU'(z) when z != null), to
// bf16 [B][256] and [256][B]
__global__ void k_delta(float *__restrict__ d, const float *__restrict__ z, int B,
                        uint16_t *__restrict__ db, uint16_t *__restrict__ dbt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * kW) return;
  const int r = i / kW, k = i % kW;
  float v = d[i];
  if (z) {
    v *= dsilu(z[i]);
    d[i] = v;
  }
  const uint16_t h = f2bf(v);
  db[i] = h;
  dbt[(size_t)k * B + r] = h;
}

// column sums of X [rows][cols] (cols <= 256): partial[blockIdx.x][c]
__global__ void k_colsum(const float *__restrict__ X, int rows, int cols, int rows_per_block,
                         float *__restrict__ partial) {
  const int c = threadIdx.x;
  if (c >= cols) return;
  const int r0 = blockIdx.x * rows_per_block, r1 = min(rows, r0 + rows_per_block);
  float s = 0.f;
  for (int r = r0; r < r1; ++r) s += X[(size_t)r * cols + c];
  partial[(size_t)blockIdx.x * cols + c] = s;
}

// out[i] = sum_s parts[s][i] (fixed order)
__global__ void k_sum_parts(const float *__restrict__ parts, int S, int64_t len,
                            float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (int k = 0; k < S; ++k) s += parts[(size_t)k * len + i];
  out[i] = s;
}

// dW5 [256] and db5 partial sums over rows
__global__ void k_head_grads(const uint16_t *__restrict__ a4, const float *__restrict__ dy,
                             int B, int rows_per_block, float *__restrict__ partial) {
  const int c = threadIdx.x;  // 0..256 (256 = bias)
  const int r0 = blockIdx.x * rows_per_block, r1 = min(B, r0 + rows_per_block);
  float s = 0.f;
  if (c < kW)
    for (int r = r0; r < r1; ++r) s = fmaf(dy[r], bf2f(a4[(size_t)r * kW + c]), s);
  else
    for (int r = r0; r < r1; ++r) s += dy[r];
  partial[(size_t)blockIdx.x * (kW + 1) + c] = s;
}

// loss = mean (y - t)^2, one block
__global__ void k_loss(const float *__restrict__ y, const float *__restrict__ t, int B,
                       float *__restrict__ loss) {
  __shared__ float sh[kBlock];
  float s = 0.f;
  for (int i = threadIdx.x; i < B; i += kBlock) {
    const float e = y[i] - t[i];
    s = fmaf(e, e, s);
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = sh[0] / (float)B;
}

// Adam (FP32 master), bias-corrected: m, v moments; bc1 = 1 - b1^t, bc2 = 1 - b2^t
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// bf16 copies of the weights: W_l [256][K_l], transposes for l >= 2, w5
__global__ void k_refresh(const float *__restrict__ p, uint16_t *__restrict__ wb1,
                          uint16_t *__restrict__ wb2, uint16_t *__restrict__ wb3,
                          uint16_t *__restrict__ wb4, uint16_t *__restrict__ wt2,
                          uint16_t *__restrict__ wt3, uint16_t *__restrict__ wt4,
                          uint16_t *__restrict__ w5b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kW * kIn) wb1[i] = f2bf(p[MlpLayout::kW1 + i]);
  if (i < kW * kW) {
    const int o = i / kW, k = i % kW;
    const uint16_t h2 = f2bf(p[MlpLayout::kW2 + i]), h3 = f2bf(p[MlpLayout::kW3 + i]),
                   h4 = f2bf(p[MlpLayout::kW4 + i]);
    wb2[i] = h2;
    wb3[i] = h3;
    wb4[i] = h4;
    wt2[(size_t)k * kW + o] = h2;
    wt3[(size_t)k * kW + o] = h3;
    wt4[(size_t)k * kW + o] = h4;
  }
  if (i < kW) w5b[i] = f2bf(p[MlpLayout::kW5 + i]);
}

// exact signed-distance labels
template <int VAR>
__global__ __launch_bounds__(256) void k_signed_distance(PlantImage P, const float *__restrict__ q,
                                                         int64_t n, int64_t sn, int64_t sj,
                                                         float *__restrict__ out) {
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n) return;
  const QMem qa{q, gi * sn, sj};
  float d;
  if constexpr (VAR == kFast8x4) d = min_signed_distance_fast<8, 4>(P, qa);
  else if constexpr (VAR == kFast8x8) d = min_signed_distance_fast<8, 8>(P, qa);
  else if constexpr (VAR == kFast16x4) d = min_signed_distance_fast<16, 4>(P, qa);
  else d = min_signed_distance_fast<16, 8>(P, qa);
  out[gi] = d;
}

int fail(int code, const char *msg) {
  set_error(msg);
  return code;
}

}  // namespace
}  // namespace csd

using namespace csd;

// ---------------------------------------------------------------------------
// model object
// ---------------------------------------------------------------------------
struct csd_mlp {
  int in_dim = 0;
  void *arena = nullptr;  // parameters, bf16 copies, Adam state, normalisation
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *lo = nullptr, *scale = nullptr;
  uint16_t *wb[kHid + 1] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint16_t *wt[kHid + 1] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint16_t *w5b = nullptr;
  int64_t step = 0;
  // EI-ZO workspaces currently guided by this model
  mutable std::atomic<int> attached{0};
  // training workspace, sized for the largest batch seen
  int cap = 0;
  void *ws = nullptr;
  size_t ws_bytes = 0;
  MlpDev dev() const {
    MlpDev d;
    d.p = p;
    for (int l = 0; l <= kHid; ++l) d.wb[l] = wb[l];
    d.w5b = w5b;
    d.lo = lo;
    d.scale = scale;
    d.in_dim = in_dim;
    return d;
  }
};

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int refresh(csd_mlp *m, hipStream_t st) {
  hipLaunchKernelGGL(k_refresh, dim3(kW * kW / 256), dim3(256), 0, st, m->p, m->wb[1], m->wb[2],
                     m->wb[3], m->wb[4], m->wt[2], m->wt[3], m->wt[4], m->w5b);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

// training workspace for batch B (multiple of 128)
struct TrainWs {
  uint16_t *x0, *x0t;                 // [B][32], [32][B]
  float *z[kHid + 1];                 // [B][256] fp32 (pre-activation, + bias)
  uint16_t *a[kHid + 1], *at[kHid + 1];  // bf16 activations and transposes
  float *y, *dy, *d;                  // [B], [B], [B][256] fp32 delta
  uint16_t *db, *dbt;                 // bf16 delta and transpose
  float *parts;                       // split-K / column partial sums
  float *loss;
};
constexpr int kSplit = 32;        // split-K for the weight gradients
constexpr int kColBlocks = 64;    // row blocks of the column sums

size_t train_ws_bytes(int B) {
  size_t s = 0;
  s += 2 * align256((size_t)B * kIn * 2);
  s += kHid * align256((size_t)B * kW * 4);
  s += 2 * kHid * align256((size_t)B * kW * 2);
  s += 2 * align256((size_t)B * 4) + align256((size_t)B * kW * 4);
  s += 2 * align256((size_t)B * kW * 2);
  s += align256((size_t)kSplit * kW * kW * 4 + (size_t)kColBlocks * (kW + 1) * 4);
  s += align256(4);
  return s;
}

TrainWs carve(void *base, int B) {
  char *p = (char *)base;
  auto take = [&](size_t bytes) {
    char *r = p;
    p += align256(bytes);
    return (void *)r;
  };
  TrainWs w;
  w.x0 = (uint16_t *)take((size_t)B * kIn * 2);
  w.x0t = (uint16_t *)take((size_t)B * kIn * 2);
  w.z[0] = nullptr;
  w.a[0] = w.x0;
  w.at[0] = w.x0t;
  for (int l = 1; l <= kHid; ++l) w.z[l] = (float *)take((size_t)B * kW * 4);
  for (int l = 1; l <= kHid; ++l) {
    w.a[l] = (uint16_t *)take((size_t)B * kW * 2);
    w.at[l] = (uint16_t *)take((size_t)B * kW * 2);
  }
  w.y = (float *)take((size_t)B * 4);
  w.dy = (float *)take((size_t)B * 4);
  w.d = (float *)take((size_t)B * kW * 4);
  w.db = (uint16_t *)take((size_t)B * kW * 2);
  w.dbt = (uint16_t *)take((size_t)B * kW * 2);
  w.parts = (float *)take((size_t)kSplit * kW * kW * 4 + (size_t)kColBlocks * (kW + 1) * 4);
  w.loss = (float *)take(4);
  return w;
}

// launches the NT GEMM with K split `split` ways; *parts = partial sums written
int gemm_nt(const uint16_t *A, int lda, const uint16_t *B, int ldb, float *C, int ldc, int M,
            int N, int K, int split, hipStream_t st, int *parts = nullptr) {
  const int kper = ((K / split + kGK - 1) / kGK) * kGK;
  dim3 g((N + kGT - 1) / kGT, (M + kGT - 1) / kGT, (K + kper - 1) / kper);
  hipLaunchKernelGGL(k_gemm_nt, g, dim3(kBlock), 0, st, A, lda, B, ldb, C, ldc, M, N, K, kper);
  CSD_HIP(hipGetLastError());
  if (parts) *parts = (int)g.z;
  return CSD_OK;
}

// forward + backward for one batch; gradients into m->g, loss into w.loss
int forward_backward(csd_mlp *m, const float *q, const float *target, int B, hipStream_t st) {
  if (B > m->cap) {
    if (m->ws) hipFree(m->ws);
    m->ws = nullptr;
    m->cap = 0;
    const size_t bytes = train_ws_bytes(B);
    hipError_t e = hipMalloc(&m->ws, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc (MLP training workspace)");
    m->cap = B;
    m->ws_bytes = bytes;
  }
  TrainWs w = carve(m->ws, B);
  const int T = 256;
  hipLaunchKernelGGL(k_prep_input, dim3((B * kIn + T - 1) / T), dim3(T), 0, st, q, B, m->in_dim,
                     m->lo, m->scale, w.x0, w.x0t);
  for (int l = 1; l <= kHid; ++l) {
    const int K = MlpLayout::K(l);
    int rc = gemm_nt(w.a[l - 1], K, m->wb[l], K, w.z[l], kW, B, kW, K, 1, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bias_silu, dim3(B * kW / T), dim3(T), 0, st, w.z[l],
                       m->p + MlpLayout::b(l), B, w.a[l], w.at[l]);
  }
  hipLaunchKernelGGL(k_head, dim3((B + T - 1) / T), dim3(T), 0, st, w.a[kHid], w.z[kHid],
                     m->w5b, m->p, target, B, w.y, w.dy, w.d);
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(kBlock), 0, st, w.y, target, B, w.loss);
  // head gradients
  const int rpb = (B + kColBlocks - 1) / kColBlocks;
  float *colp = w.parts + (size_t)kSplit * kW * kW;
  hipLaunchKernelGGL(k_head_grads, dim3(kColBlocks), dim3(kW + 1), 0, st, w.a[kHid], w.dy, B,
                     rpb, colp);
  hipLaunchKernelGGL(k_sum_parts, dim3(2), dim3(T), 0, st, colp, kColBlocks, (int64_t)(kW + 1),
                     m->g + MlpLayout::kW5);
  for (int l = kHid; l >= 1; --l) {
    // delta_l: fp32 in w.d (already multiplied by SiLU' for l = 4)
    hipLaunchKernelGGL(k_delta, dim3(B * kW / T), dim3(T), 0, st, w.d,
                       l == kHid ? (const float *)nullptr : w.z[l], B, w.db, w.dbt);
    // db_l
    hipLaunchKernelGGL(k_colsum, dim3(kColBlocks), dim3(kW), 0, st, w.d, B, kW, rpb, colp);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kW), 0, st, colp, kColBlocks, (int64_t)kW,
                       m->g + MlpLayout::b(l));
    // dW_l [256][K_l] = delta_l^T a_{l-1}, split over the batch
    const int K = MlpLayout::K(l);
    int S = 0;
    int rc = gemm_nt(w.dbt, B, w.at[l - 1], B, w.parts, K, kW, K, B, kSplit, st, &S);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_parts, dim3((kW * K + T - 1) / T), dim3(T), 0, st, w.parts, S,
                       (int64_t)kW * K, m->g + MlpLayout::w(l));
    if (l > 1) {
      // dA_{l-1} [B][256] = delta_l W_l  (B operand: W_l^T rows)
      rc = gemm_nt(w.db, kW, m->wt[l], kW, w.d, kW, B, kW, kW, 1, st);
      if (rc) return rc;
    }
  }
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

int check_batch(const csd_mlp *m, int B) {
  if (!m) return fail(CSD_ERR_INVALID_ARGUMENT, "null model");
  if (B <= 0 || B % 128) return fail(CSD_ERR_INVALID_ARGUMENT, "batch must be a positive multiple of 128");
  if (B > (1 << 22)) return fail(CSD_ERR_CAPACITY, "batch above 4194304");
  return CSD_OK;
}
}  // namespace

extern "C" {

int csd_mlp_create(int32_t in_dim, const float *lower, const float *upper, uint64_t seed,
                   csd_mlp **out) {
  if (!out || !lower || !upper) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (in_dim < 1 || in_dim > kIn) return fail(CSD_ERR_CAPACITY, "surrogate input dimension must be 1..32");
  for (int k = 0; k < in_dim; ++k)
    if (!(upper[k] > lower[k])) return fail(CSD_ERR_INVALID_ARGUMENT, "upper must exceed lower");
  const size_t P = MlpLayout::kTotal;
  size_t bytes = 4 * align256(P * 4) + 2 * align256(kIn * 4);
  const size_t wsz[kHid + 1] = {0, (size_t)kW * kIn, (size_t)kW * kW, (size_t)kW * kW,
                                (size_t)kW * kW};
  for (int l = 1; l <= kHid; ++l) bytes += 2 * align256(wsz[l] * 2);
  bytes += align256(kW * 2);
  csd_mlp *m = new csd_mlp();
  m->in_dim = in_dim;
  hipError_t e = hipMalloc(&m->arena, bytes);
  if (e != hipSuccess) {
    delete m;
    return hip_fail(e, "hipMalloc (MLP)");
  }
  char *p = (char *)m->arena;
  auto take = [&](size_t b) {
    char *r = p;
    p += align256(b);
    return (void *)r;
  };
  m->p = (float *)take(P * 4);
  m->g = (float *)take(P * 4);
  m->m = (float *)take(P * 4);
  m->v = (float *)take(P * 4);
  m->lo = (float *)take(kIn * 4);
  m->scale = (float *)take(kIn * 4);
  for (int l = 1; l <= kHid; ++l) {
    m->wb[l] = (uint16_t *)take(wsz[l] * 2);
    m->wt[l] = (uint16_t *)take(wsz[l] * 2);
  }
  m->w5b = (uint16_t *)take(kW * 2);
  // initial parameters: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and
  // biases (fan_in of layer 1 = in_dim; padded inputs get zero weights)
  std::vector<float> h(P, 0.f), lo(kIn, 0.f), sc(kIn, 0.f);
  uint64_t s = seed;
  auto uni = [&](float bound) {
    const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
    return (float)((2.0 * u - 1.0) * bound);
  };
  for (int l = 1; l <= kHid + 1; ++l) {
    const int fan = l == 1 ? in_dim : kW;
    const float bound = 1.f / sqrtf((float)fan);
    if (l <= kHid) {
      const int K = MlpLayout::K(l);
      for (int o = 0; o < kW; ++o)
        for (int k = 0; k < K; ++k)
          h[MlpLayout::w(l) + (size_t)o * K + k] = (l == 1 && k >= in_dim) ? 0.f : uni(bound);
      for (int o = 0; o < kW; ++o) h[MlpLayout::b(l) + o] = uni(bound);
    } else {
      for (int k = 0; k < kW; ++k) h[MlpLayout::kW5 + k] = uni(bound);
      h[MlpLayout::kB5] = uni(bound);
    }
  }
  for (int k = 0; k < in_dim; ++k) {
    lo[k] = lower[k];
    sc[k] = 2.f / (upper[k] - lower[k]);
  }
  std::vector<float> zeros(P, 0.f);
  e = hipMemcpy(m->p, h.data(), P * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->g, zeros.data(), P * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->m, zeros.data(), P * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->v, zeros.data(), P * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->lo, lo.data(), kIn * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->scale, sc.data(), kIn * 4, hipMemcpyHostToDevice);
  int rc = e == hipSuccess ? refresh(m, nullptr) : hip_fail(e, "hipMemcpy (MLP init)");
  if (rc == CSD_OK) rc = hipDeviceSynchronize() == hipSuccess ? CSD_OK : CSD_ERR_HIP;
  if (rc != CSD_OK) {
    hipFree(m->arena);
    delete m;
    return rc;
  }
  *out = m;
  return CSD_OK;
}

int csd_mlp_destroy(csd_mlp *m) {
  if (!m) return CSD_OK;
  if (m->attached.load() > 0)
    return fail(CSD_ERR_INVALID_ARGUMENT,
                "the surrogate still guides an EI-ZO workspace: clear the guidance or destroy "
                "the workspace first");
  // kernels launched asynchronously may still read the model
  hipDeviceSynchronize();
  if (m->ws) hipFree(m->ws);
  if (m->arena) hipFree(m->arena);
  delete m;
  return CSD_OK;
}

int csd_mlp_num_params(const csd_mlp *m, int64_t *count) {
  if (!m || !count) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  *count = (int64_t)MlpLayout::kTotal;
  return CSD_OK;
}

int csd_mlp_get_params(const csd_mlp *m, float *host_params) {
  if (!m || !host_params) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  CSD_HIP(hipMemcpy(host_params, m->p, MlpLayout::kTotal * 4, hipMemcpyDeviceToHost));
  return CSD_OK;
}

int csd_mlp_set_params(csd_mlp *m, const float *host_params) {
  if (!m || !host_params) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  CSD_HIP(hipMemcpy(m->p, host_params, MlpLayout::kTotal * 4, hipMemcpyHostToDevice));
  int rc = refresh(m, nullptr);
  if (rc) return rc;
  CSD_HIP(hipDeviceSynchronize());
  return CSD_OK;
}

int csd_mlp_predict(const csd_mlp *m, const float *q, int64_t n, int64_t stride_n,
                    int64_t stride_j, float *out, void *stream) {
  if (!m) return fail(CSD_ERR_INVALID_ARGUMENT, "null model");
  if (n < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative count");
  if (n == 0) return CSD_OK;
  if (!q || !out) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  const int64_t blocks = (n + 4 * kRowsWave - 1) / (4 * kRowsWave);
  if (blocks > 0x7fffffffLL) return fail(CSD_ERR_CAPACITY, "too many rows");
  hipLaunchKernelGGL((k_mlp_infer<StridedInput>), dim3((unsigned)blocks), dim3(kBlock), 0,
                     (hipStream_t)stream, m->dev(), StridedInput{q, stride_n, stride_j, m->in_dim},
                     n, out);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

int csd_mlp_gradients(csd_mlp *m, const float *q, const float *target, int32_t batch,
                      float *grads, float *loss, void *stream) {
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if (!q || !target) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  hipStream_t st = (hipStream_t)stream;
  rc = forward_backward(m, q, target, batch, st);
  if (rc) return rc;
  TrainWs w = carve(m->ws, batch);
  if (grads) CSD_HIP(hipMemcpyAsync(grads, m->g, MlpLayout::kTotal * 4, hipMemcpyDeviceToDevice, st));
  if (loss) CSD_HIP(hipMemcpyAsync(loss, w.loss, 4, hipMemcpyDeviceToDevice, st));
  return CSD_OK;
}

int csd_mlp_train_step(csd_mlp *m, const float *q, const float *target, int32_t batch, float lr,
                       float *loss, void *stream) {
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if (!q || !target) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  if (!(lr > 0.f)) return fail(CSD_ERR_INVALID_ARGUMENT, "learning rate must be positive");
  hipStream_t st = (hipStream_t)stream;
  rc = forward_backward(m, q, target, batch, st);
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m->step += 1;
  const float bc1 = (float)(1.0 - pow((double)b1, (double)m->step));
  const float bc2 = (float)(1.0 - pow((double)b2, (double)m->step));
  const int64_t P = MlpLayout::kTotal;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, m->p, m->g, m->m,
                     m->v, P, lr, b1, b2, eps, bc1, bc2);
  rc = refresh(m, st);
  if (rc) return rc;
  if (loss) {
    TrainWs w = carve(m->ws, batch);
    CSD_HIP(hipMemcpyAsync(loss, w.loss, 4, hipMemcpyDeviceToDevice, st));
  }
  return CSD_OK;
}

}  // extern "C"

void csd::mlp_attach(const csd_mlp *m, int delta) {
  if (m) m->attached.fetch_add(delta);
}

int csd::mlp_predict_segment_points(const csd_mlp *m, const SegmentPoints &s, int num_slots,
                                    float *out, hipStream_t st) {
  if (!m) return fail(CSD_ERR_INVALID_ARGUMENT, "null surrogate");
  if (s.dim != m->in_dim) return fail(CSD_ERR_INVALID_ARGUMENT, "surrogate input dimension differs from the plant's");
  const int64_t n = (int64_t)num_slots * s.cand * s.L;
  if (n == 0) return CSD_OK;
  const int64_t blocks = (n + 4 * kRowsWave - 1) / (4 * kRowsWave);
  if (blocks > 0x7fffffffLL) return fail(CSD_ERR_CAPACITY, "too many rows");
  hipLaunchKernelGGL((k_mlp_infer<SegmentInput>), dim3((unsigned)blocks), dim3(kBlock), 0, st,
                     m->dev(), SegmentInput{s}, n, out);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

extern "C" {

int csd_signed_distance(const csd_plant *plant, const float *q, int64_t n, int64_t stride_n,
                        int64_t stride_j, float *out, void *stream) {
  if (!plant) return fail(CSD_ERR_INVALID_ARGUMENT, "null plant");
  if (n < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative count");
  if (n == 0) return CSD_OK;
  if (!q || !out) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  const PlantImage &P = plant->img;
  if (!P.fast)
    return fail(CSD_ERR_UNSUPPORTED,
                "signed distance needs every moving collision geometry to be a sphere");
  const int var = fast_variant(P);
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffLL) return fail(CSD_ERR_CAPACITY, "too many configurations");
  hipStream_t st = (hipStream_t)stream;
  const dim3 g((unsigned)blocks), b(256);
  switch (var) {
    case kFast8x4:
      hipLaunchKernelGGL((k_signed_distance<kFast8x4>), g, b, 0, st, P, q, n, stride_n, stride_j, out);
      break;
    case kFast8x8:
      hipLaunchKernelGGL((k_signed_distance<kFast8x8>), g, b, 0, st, P, q, n, stride_n, stride_j, out);
      break;
    case kFast16x4:
      hipLaunchKernelGGL((k_signed_distance<kFast16x4>), g, b, 0, st, P, q, n, stride_n, stride_j, out);
      break;
    default:
      hipLaunchKernelGGL((k_signed_distance<kFast16x8>), g, b, 0, st, P, q, n, stride_n, stride_j, out);
      break;
  }
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

}  // extern "C"
