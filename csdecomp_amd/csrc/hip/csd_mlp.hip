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
  __device__ __forceinline__ bool live(int64_t) const { return true; }
  __device__ __forceinline__ void load(int64_t r, float (&x)[kIn]) const {
#pragma unroll
    for (int k = 0; k < kIn; ++k) x[k] = k < in_dim ? q[r * sn + (int64_t)k * sj] : 0.f;
  }
};

// raw inputs of row (slot, i, l): a point between a particle's projection
// onto its segment and the particle (projection as in the bisection)
struct SegmentInput {
  SegmentPoints s;
  // a block whose first row lies in a slot past the live count has nothing to do
  __device__ __forceinline__ bool live(int64_t row0) const {
    return !s.live_slots || row0 / ((int64_t)s.cand * s.L) < (int64_t)*s.live_slots;
  }
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

// raw inputs of row r: sample r % 256 of staging segment r / 256, stored as
// [dim][256] (the EI-ZO sampler's layout)
struct StagedInput {
  const float *stage;
  int dim;
  const uint32_t *live_slots;  // null: every slot
  int64_t rows_per_slot;
  __device__ __forceinline__ bool live(int64_t row0) const {
    return !live_slots || row0 / rows_per_slot < (int64_t)*live_slots;
  }
  __device__ __forceinline__ void load(int64_t r, float (&x)[kIn]) const {
    const float *seg = stage + (size_t)(r >> 8) * 256 * dim + (r & 255);
#pragma unroll
    for (int k = 0; k < kIn; ++k) x[k] = k < dim ? seg[(size_t)k * 256] : 0.f;
  }
};

template <class In>
__global__ __launch_bounds__(kBlock, 2) void k_mlp_infer(MlpDev M, In in, int64_t n,
                                                      float *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint16_t sw[kW * kChunkPad];
  __shared__ __attribute__((aligned(16))) uint16_t st[4][32 * kTilePad];
  if (!in.live((int64_t)blockIdx.x * (4 * kRowsWave))) return;  // block-uniform
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
// Training, 3 launches per step at any batch B (a multiple of 128):
//   k_train_rows    one block per 32 rows, everything that is row-local:
//                   x0 = bf16(normalised q); layers l = 1..4 with
//                   z = a_{l-1} W_l^T + b_l, a_l = bf16(SiLU(z)) and
//                   s_l = bf16(SiLU'(z)); y = w5 . a_4 + b5, the loss,
//                   dy = 2 (y - t) / B and delta_4 = dy w5 SiLU'(z_4); then
//                   delta_{l-1} = (delta_l W_l) * s_{l-1} for l = 4..2.
//                   Activations, deltas and s_l stay in LDS; the transposes
//                   a_l^T [K][B] (l = 0..3) and delta_l^T [256][B]
//                   (l = 1..4) go to HBM for the weight gradients; the
//                   sums of db_l, dw5, db5 and the loss go to accumulators.
//   k_train_wgrad   dW_l = delta_l^T a_{l-1} for l = 1..4, each block a tile
//                   of one dW_l over a share of the batch (partial sums in
//                   a fixed order, no atomics)
//   k_train_update  the gradient (partials summed in a fixed order), then
//                   Adam and the bf16 weight copies (or only the gradient,
//                   for csd_mlp_gradients); clears the accumulators
// GEMM operands are bf16 [rows][k] with k contiguous. The weights stream
// through one LDS buffer in K chunks of 128 (layer 1: its 32), the next
// chunk's global loads in flight in registers while the current one feeds
// the MFMAs; the chunks of consecutive layers form one stream, so the next
// layer's first chunk loads during the current layer's epilogue. At these
// sizes the step is bound by load latency and launch count, not by MFMA.
// ---------------------------------------------------------------------------
constexpr int kKC = 128;  // K chunk (bf16 elements)
constexpr int kRB = 32;   // rows per block of the row kernel

// small gradient accumulators (summed with atomics): b1..b4, w5, b5, loss
constexpr int kAccB = 0, kAccW5 = kHid * kW, kAccB5 = kAccW5 + kW, kAccLoss = kAccB5 + 1;
constexpr int kAccLen = kAccLoss + 1;

struct TrainDev {
  int B;
  uint16_t *at[kHid];       // a_l^T [K_l][B], l = 0..3 (a_0 = x0, K_0 = 32)
  uint16_t *dt[kHid + 1];   // delta_l^T [256][B], l = 1..4
  float *wparts[kHid + 1];  // dW_l partial sums [S_l][256][K_l]
  int split[kHid + 1];      // S_l
  float *acc;               // [kAccLen], zero between steps
};

// Arguments of the row kernel (plain pointers: a runtime index into an
// array of the argument block would put it in scratch memory).
struct RowArgs {
  int B;
  const float *q, *target;
  uint16_t *at0, *at1, *at2, *at3;
  uint16_t *dt1, *dt2, *dt3, *dt4;
  float *acc;
};

// K-chunk staging: rows [r0, r0 + NA) of A and [c0, c0 + NB) of Bop (each
// [rows][ld] bf16), columns k0 .. k0 + KC, as 16-byte pieces; thread t takes
// pieces t, t + 256, ... and holds them in registers until they are stored
// to LDS rows of pitch KC + 8 (16-byte aligned rows; the 8 extra elements
// shift consecutive rows by four banks). NA = 0 stages Bop alone.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int NA, int NB, int KC, int NT = 256>
struct ChunkLoader {
  static constexpr int kPerRow = KC / 8;
  static constexpr int kPieces = (NA + NB) * kPerRow;
  static constexpr int kPer = (kPieces + NT - 1) / NT;
  static constexpr int kPitch = KC + 8;
  static_assert(kPer <= 16, "at most 16 pieces per thread");
  u32x4 v[kPer];
  __device__ __forceinline__ void load(const uint16_t *A, int lda, int r0, const uint16_t *Bm,
                                       int ldb, int c0, int k0) {
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int pc = (int)threadIdx.x + NT * j;
      if ((j + 1) * NT <= kPieces || pc < kPieces) {
        const int row = pc / kPerRow, part = pc % kPerRow;
        const uint16_t *src = row < NA ? A + (size_t)(r0 + row) * lda
                                       : Bm + (size_t)(c0 + row - NA) * ldb;
        v[j] = *reinterpret_cast<const u32x4 *>(src + k0 + part * 8);
      }
    }
  }
  __device__ __forceinline__ void store(uint16_t *sA) const {  // sA: NA + NB rows
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int pc = (int)threadIdx.x + NT * j;
      if ((j + 1) * NT <= kPieces || pc < kPieces)
        *reinterpret_cast<u32x4 *>(sA + (pc / kPerRow) * kPitch + (pc % kPerRow) * 8) = v[j];
    }
  }
};

// LDS of the row kernel (bf16 elements unless noted)
constexpr int kApitch = kW + 8;                  // activations / deltas [32][264]
constexpr int kWpitch = kKC + 8;                 // weight chunk [256][136]
constexpr int kLdsA = kRB * kApitch;             // 8,448
constexpr int kLdsW = kW * kWpitch;              // 34,816
constexpr int kLdsS = kRB * kW;                  // one s_l [32][256]: 8,192

// The row kernel's waves: 8 of them, each holding one 32 x 32 tile (32
// columns) of a layer's 32 x 256 output, two per SIMD.
constexpr int kRowTiles = 1;                 // 32 x 32 MFMA tiles per wave
constexpr int kRowWaves = 8 / kRowTiles;
constexpr int kRowThreads = 64 * kRowWaves;

// One 128-deep chunk of a layer's row GEMM: C[32][256] += sA[.][k0 .. k0+128]
// sW^T, wave w holding columns 32 TPW w .. + 32 TPW as TPW 32x32 MFMA tiles.
template <int KC, int PITCH, int TPW>
__device__ __forceinline__ void chunk_mfma(const uint16_t *sA, int k0, const uint16_t *sW,
                                           f32x16 (&acc)[TPW]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hr = lane >> 5, c = lane & 31;
#pragma unroll
  for (int ks = 0; ks < KC; ks += 16) {
    const bf16x8 af = ld8(sA + c * kApitch + k0 + ks + 8 * hr);
#pragma unroll
    for (int j = 0; j < TPW; ++j)
      acc[j] = mfma(af, ld8(sW + (32 * TPW * wave + 32 * j + c) * PITCH + ks + 8 * hr), acc[j]);
  }
}

// 4 bf16 values of consecutive rows to the transposed layout [col][B]
__device__ __forceinline__ void store_t4(uint16_t *t, int B, int col, int row,
                                         const uint16_t (&h)[4]) {
  uint2 v;
  v.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
  v.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
  *reinterpret_cast<uint2 *>(t + (size_t)col * B + row) = v;
}

// sigmoid from the hardware exp2 and reciprocal (~1 ulp): the activation is
// rounded to bf16 at once and SiLU' only scales the deltas (for z -> -inf,
// exp2 -> inf and sigmoid -> 0)
__device__ __forceinline__ float sigmoid_hw(float z) {
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}

// The chunk stream of the row kernel: W1 (one chunk of 32), then two chunks
// of 128 for each of W2, W3, W4 (forward) and W4^T, W3^T, W2^T (deltas).
// Chunk n = 1..12 is columns ((n - 1) % 2) * 128 .. + 128 of matrix (n - 1) / 2.
__device__ __forceinline__ const uint16_t *stream_matrix(const MlpDev &M, const uint16_t *wt2,
                                                         const uint16_t *wt3,
                                                         const uint16_t *wt4, int n) {
  const int m = (n - 1) >> 1;
  return m == 0 ? M.wb[2] : m == 1 ? M.wb[3] : m == 2 ? M.wb[4] : m == 3 ? wt4 : m == 4 ? wt3
                                                                                      : wt2;
}

struct RowLds {
  uint16_t *sA;  // [32][kApitch]: the next GEMM's A operand
  uint16_t *sS;  // s_1, s_2, s_3, each [32][256]
  float (*s_row)[kRB];
  float *s_dy;
};

// The epilogue after chunk n (n = 0 or even): n = 0, 2, 4 forward layers
// 1..3; 6 layer 4 and the head; 8, 10, 12 delta_3, delta_2, delta_1. bf16
// rounding by the hardware conversion (round to nearest even, as f2bf, for
// these finite values).
template <int TPW>
__device__ __forceinline__ void row_epilogue(int n, const RowArgs &T, const MlpDev &M,
                                             const RowLds &L, f32x16 (&acc)[TPW]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hr = lane >> 5, c = lane & 31;
  const int B = T.B, r0 = blockIdx.x * kRB;
  const int l = n / 2 + 1;  // forward layer, or 4 + (backward step)
  if (l <= 3) {
    const float *bias = M.p + MlpLayout::b(l);
    uint16_t *atl = l == 1 ? T.at1 : l == 2 ? T.at2 : T.at3;
    uint16_t *sl = L.sS + (l - 1) * kLdsS;
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int col = 32 * TPW * wave + 32 * j + c;
      const float bc = bias[col];
#pragma unroll
      for (int qd = 0; qd < 4; ++qd) {
        uint16_t h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int rr = 8 * qd + 4 * hr + e;
          const float z = acc[j][4 * qd + e] + bc;
          const float sg = sigmoid_hw(z);
          h[e] = f2bf_hw(z * sg);
          L.sA[rr * kApitch + col] = h[e];
          sl[rr * kW + col] = f2bf_hw(sg * (1.f + z * (1.f - sg)));
        }
        store_t4(atl, B, col, r0 + 8 * qd + 4 * hr, h);
      }
    }
  } else if (l == 4) {
    // layer 4, y = w5 . a4 + b5 per row, loss, delta_4 = dy w5 SiLU'(z4)
    const float *bias = M.p + MlpLayout::b(4);
    float av[TPW][16], sv[TPW][16];
    float part[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) part[i] = 0.f;
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int col = 32 * TPW * wave + 32 * j + c;
      const float bc = bias[col];
      const float w = bf2f(M.w5b[col]);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const float z = acc[j][i] + bc;
        const float sg = sigmoid_hw(z);
        av[j][i] = bf2f(f2bf_hw(z * sg));
        sv[j][i] = sg * (1.f + z * (1.f - sg));
        part[i] = fmaf(av[j][i], w, part[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
#pragma unroll
      for (int o = 16; o > 0; o >>= 1) part[i] += __shfl_xor(part[i], o, 32);
    }
    if (c == 0) {
#pragma unroll
      for (int i = 0; i < 16; ++i) L.s_row[wave][8 * (i >> 2) + 4 * hr + (i & 3)] = part[i];
    }
    __syncthreads();
    if (threadIdx.x < kRB) {  // lanes 0..31 of wave 0, one row each
      const int r = threadIdx.x;
      float y = L.s_row[0][r];
#pragma unroll
      for (int w = 1; w < kRowWaves; ++w) y += L.s_row[w][r];
      y += M.p[MlpLayout::kB5];
      const float e = y - T.target[r0 + r];
      const float dy = 2.f * e / (float)B;
      L.s_dy[r] = dy;
      float sdy = dy, se = e * e;
#pragma unroll
      for (int o = 16; o > 0; o >>= 1) {
        sdy += __shfl_xor(sdy, o, 32);
        se += __shfl_xor(se, o, 32);
      }
      if (r == 0) {
        atomicAdd(T.acc + kAccB5, sdy);
        atomicAdd(T.acc + kAccLoss, se);
      }
    }
    __syncthreads();
    // delta_4 (bf16) into sA and delta_4^T; column sums db4 and dw5
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int col = 32 * TPW * wave + 32 * j + c;
      const float w = bf2f(M.w5b[col]);
      float sd = 0.f, sw = 0.f;
#pragma unroll
      for (int qd = 0; qd < 4; ++qd) {
        uint16_t h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int i = 4 * qd + e, rr = 8 * qd + 4 * hr + e;
          const float dy = L.s_dy[rr];
          const float dv = dy * w * sv[j][i];
          sd += dv;
          sw = fmaf(dy, av[j][i], sw);
          h[e] = f2bf_hw(dv);
          L.sA[rr * kApitch + col] = h[e];
        }
        store_t4(T.dt4, B, col, r0 + 8 * qd + 4 * hr, h);
      }
      sd += __shfl_xor(sd, 32);
      sw += __shfl_xor(sw, 32);
      if (hr == 0) {
        atomicAdd(T.acc + kAccB + (kHid - 1) * kW + col, sd);
        atomicAdd(T.acc + kAccW5 + col, sw);
      }
    }
  } else {
    // delta_{m-1} = (delta_m W_m) * s_{m-1}, m = 4, 3, 2 for n = 8, 10, 12
    const int m = 8 - n / 2;
    const uint16_t *sl = L.sS + (m - 2) * kLdsS;
    uint16_t *dtl = m == 4 ? T.dt3 : m == 3 ? T.dt2 : T.dt1;
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int col = 32 * TPW * wave + 32 * j + c;
      float sd = 0.f;
#pragma unroll
      for (int qd = 0; qd < 4; ++qd) {
        uint16_t h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int rr = 8 * qd + 4 * hr + e;
          const float dv = acc[j][4 * qd + e] * bf2f(sl[rr * kW + col]);
          sd += dv;
          h[e] = f2bf_hw(dv);
          L.sA[rr * kApitch + col] = h[e];
        }
        store_t4(dtl, B, col, r0 + 8 * qd + 4 * hr, h);
      }
      sd += __shfl_xor(sd, 32);
      if (hr == 0) atomicAdd(T.acc + kAccB + (m - 2) * kW + col, sd);
    }
  }
#pragma unroll
  for (int j = 0; j < TPW; ++j) acc[j] = f32x16{};
}

__global__ __launch_bounds__(kRowThreads) void k_train_rows(RowArgs T, MlpDev M, const uint16_t *wt2,
                                                    const uint16_t *wt3, const uint16_t *wt4) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[kLdsA];
  __shared__ __attribute__((aligned(16))) uint16_t sW[kLdsW];
  __shared__ __attribute__((aligned(16))) uint16_t sS[3 * kLdsS];  // s_1, s_2, s_3
  __shared__ float s_row[kRowWaves][kRB];
  __shared__ float s_dy[kRB];
  const RowLds L{sA, sS, s_row, s_dy};
  const int B = T.B, r0 = blockIdx.x * kRB;

  // layer 1's weights (256 x 32) and x0 of the block's rows
  {
    ChunkLoader<0, kW, kIn, kRowThreads> l1;
    l1.load(nullptr, 0, 0, M.wb[1], kIn, 0, 0);
    // x0: 32 rows x 32 features, thread t takes row t / (N / 32), features
    // E (t % (N / 32)) .. + E - 1 for N threads, E = 1024 / N
    constexpr int E = kRB * kIn / kRowThreads;
    const int row = threadIdx.x / (kIn / E), k0 = (threadIdx.x % (kIn / E)) * E;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = k0 + e;
      const float x = k < M.in_dim
                          ? (T.q[(size_t)(r0 + row) * M.in_dim + k] - M.lo[k]) * M.scale[k] - 1.f
                          : 0.f;
      const uint16_t h = f2bf(x);
      sA[row * kApitch + k] = h;
      T.at0[(size_t)k * B + r0 + row] = h;
    }
    l1.store(sW);
  }
  // Chunks n = 1..12 stream through sW, two in flight at a time in
  // registers: odd chunks in `ldo`, even ones in `lde`.
  ChunkLoader<0, kW, kKC, kRowThreads> ldo, lde;
  __syncthreads();
  ldo.load(nullptr, 0, 0, stream_matrix(M, wt2, wt3, wt4, 1), kW, 0, 0);
  lde.load(nullptr, 0, 0, stream_matrix(M, wt2, wt3, wt4, 2), kW, 0, kKC);
  f32x16 acc[kRowTiles];
#pragma unroll
  for (int j = 0; j < kRowTiles; ++j) acc[j] = f32x16{};
  chunk_mfma<kIn, kIn + 8, kRowTiles>(sA, 0, sW, acc);
  __syncthreads();  // every wave is done reading sA and sW
  row_epilogue<kRowTiles>(0, T, M, L, acc);
  ldo.store(sW);
  __syncthreads();
  ldo.load(nullptr, 0, 0, stream_matrix(M, wt2, wt3, wt4, 3), kW, 0, 0);
  for (int pr = 1; pr <= 6; ++pr) {
    chunk_mfma<kKC, kWpitch, kRowTiles>(sA, 0, sW, acc);  // chunk 2 pr - 1
    __syncthreads();
    lde.store(sW);
    __syncthreads();
    if (pr < 6) lde.load(nullptr, 0, 0, stream_matrix(M, wt2, wt3, wt4, 2 * pr + 2), kW, 0, kKC);
    chunk_mfma<kKC, kWpitch, kRowTiles>(sA, kKC, sW, acc);  // chunk 2 pr
    __syncthreads();
    row_epilogue<kRowTiles>(2 * pr, T, M, L, acc);
    if (pr < 6) {
      ldo.store(sW);  // chunk 2 pr + 1
      __syncthreads();
      if (pr < 5) ldo.load(nullptr, 0, 0, stream_matrix(M, wt2, wt3, wt4, 2 * pr + 3), kW, 0, 0);
    }
  }
}

// dW_l = delta_l^T a_{l-1}: a block takes one 32 WA x 32 WB tile of dW_l
// (WA * WB = 4 waves of one 32x32 MFMA tile) over a share of the batch and
// writes its partial sum (no atomics).
struct WgradArgs {
  int B;
  int first[kHid + 1];       // first block of layer l (l = 1..4); first[0] = total
  int split[kHid + 1];
  const uint16_t *dt[kHid + 1], *at[kHid + 1];  // delta_l^T [256][B], a_{l-1}^T [K][B]
  float *wparts[kHid + 1];
};

template <int WA, int WB>
__device__ __forceinline__ void wgrad_block(const uint16_t *A, const uint16_t *Bm, int K, int B,
                                            int S, int blk, float *wparts, uint16_t *lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hr = lane >> 5, c = lane & 31;
  constexpr int BM = 32 * WA, BN = 32 * WB;
  const int tn = K / BN;
  const int tile = blk / S, sp = blk - tile * S;
  const int m0 = (tile / tn) * BM, n0 = (tile % tn) * BN;
  const int nch = B / kKC, cb = sp * nch / S, ce = (sp + 1) * nch / S;
  const int wa = wave / WB, wb = wave % WB;
  using L = ChunkLoader<BM, BN, kKC>;
  L ld;
  f32x16 acc = f32x16{};
  if (cb < ce) {  // block-uniform
    ld.load(A, B, m0, Bm, B, n0, cb * kKC);
    ld.store(lds);
  }
  __syncthreads();
  for (int kc = cb; kc < ce; ++kc) {
    const bool more = kc + 1 < ce;
    if (more) ld.load(A, B, m0, Bm, B, n0, (kc + 1) * kKC);
#pragma unroll
    for (int ks = 0; ks < kKC; ks += 16)
      acc = mfma(ld8(lds + (32 * wa + c) * L::kPitch + ks + 8 * hr),
                 ld8(lds + (BM + 32 * wb + c) * L::kPitch + ks + 8 * hr), acc);
    if (more) {  // every wave has read the chunk before it is overwritten
      __syncthreads();
      ld.store(lds);
      __syncthreads();
    }
  }
  float *out = wparts + (size_t)sp * kW * K;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int o = m0 + 32 * wa + 8 * (i >> 2) + 4 * hr + (i & 3);
    out[(size_t)o * K + n0 + 32 * wb + c] = acc[i];
  }
}

__global__ __launch_bounds__(256) void k_train_wgrad(WgradArgs T) {
  __shared__ __attribute__((aligned(16))) uint16_t lds[(128 + 32) * (kKC + 8)];
  const int b = blockIdx.x;
  // layers in order 4, 3, 2, 1 of the block index
  if (b < T.first[3]) {
    wgrad_block<2, 2>(T.dt[4], T.at[4], kW, T.B, T.split[4], b - T.first[4], T.wparts[4], lds);
  } else if (b < T.first[2]) {
    wgrad_block<2, 2>(T.dt[3], T.at[3], kW, T.B, T.split[3], b - T.first[3], T.wparts[3], lds);
  } else if (b < T.first[1]) {
    wgrad_block<2, 2>(T.dt[2], T.at[2], kW, T.B, T.split[2], b - T.first[2], T.wparts[2], lds);
  } else {
    wgrad_block<4, 1>(T.dt[1], T.at[1], kIn, T.B, T.split[1], b - T.first[1], T.wparts[1], lds);
  }
}

// The gradient of each parameter (dW partials summed over the splits in a
// fixed order; biases, w5 and b5 from the accumulators) into g and, with
// lr > 0, one Adam step (FP32 master, bias-corrected moments) with the bf16
// copies of the updated weights. Thread 0 also writes the loss.
struct UpdateArgs {
  int B;
  const float *wparts[kHid];  // dW_l partials, l = 1..4
  int split[kHid];
  float *acc;  // each entry is read by one thread, which then clears it
};

__global__ void k_train_update(UpdateArgs T, float *__restrict__ p, float *__restrict__ g,
                               float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                               float b2, float eps, float bc1, float bc2, uint16_t *wb1,
                               uint16_t *wb2, uint16_t *wb3, uint16_t *wb4, uint16_t *wt2,
                               uint16_t *wt3, uint16_t *wt4, uint16_t *w5b, float *loss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    if (loss) *loss = T.acc[kAccLoss] / (float)T.B;
    T.acc[kAccLoss] = 0.f;
  }
  if (i >= (int64_t)MlpLayout::kTotal) return;
  int l = 0, o = 0, k = 0;  // W_l [o][k], or l = 0
  float gi;
  if (i < (int64_t)MlpLayout::kB1) {
    l = 1;
  } else if (i >= (int64_t)MlpLayout::kW2 && i < (int64_t)MlpLayout::kB2) {
    l = 2;
  } else if (i >= (int64_t)MlpLayout::kW3 && i < (int64_t)MlpLayout::kB3) {
    l = 3;
  } else if (i >= (int64_t)MlpLayout::kW4 && i < (int64_t)MlpLayout::kB4) {
    l = 4;
  }
  if (l) {
    const int K = MlpLayout::K(l);
    const int64_t e = i - (int64_t)MlpLayout::w(l);
    o = (int)(e / K);
    k = (int)(e - (int64_t)o * K);
    const float *wp = (l == 1 ? T.wparts[0] : l == 2 ? T.wparts[1] : l == 3 ? T.wparts[2]
                                                                           : T.wparts[3]) + e;
    const int S = l == 1 ? T.split[0] : l == 2 ? T.split[1] : l == 3 ? T.split[2] : T.split[3];
    const size_t stride = (size_t)kW * K;
    // four independent sums keep four loads in flight; fixed order
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int sp = 0;
    for (; sp + 4 <= S; sp += 4) {
      s0 += wp[(size_t)sp * stride];
      s1 += wp[(size_t)(sp + 1) * stride];
      s2 += wp[(size_t)(sp + 2) * stride];
      s3 += wp[(size_t)(sp + 3) * stride];
    }
    for (; sp < S; ++sp) s0 += wp[(size_t)sp * stride];
    gi = (s0 + s1) + (s2 + s3);
  } else {
    int a;  // the accumulator of this bias, w5 entry or b5
    if (i == (int64_t)MlpLayout::kB5) {
      a = kAccB5;
    } else if (i >= (int64_t)MlpLayout::kW5) {
      k = (int)(i - (int64_t)MlpLayout::kW5);
      a = kAccW5 + k;
    } else {
      const int lb = i < (int64_t)MlpLayout::kW2 ? 1 : i < (int64_t)MlpLayout::kW3 ? 2
                     : i < (int64_t)MlpLayout::kW4 ? 3 : 4;
      a = kAccB + (lb - 1) * kW + (int)(i - (int64_t)MlpLayout::b(lb));
    }
    gi = T.acc[a];
    T.acc[a] = 0.f;  // zero for the next step's atomics
  }
  g[i] = gi;
  if (!(lr > 0.f)) return;
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  p[i] = pi;
  const uint16_t h = f2bf(pi);
  if (l == 1) {
    wb1[(size_t)o * kIn + k] = h;
  } else if (l) {
    uint16_t *wb = l == 2 ? wb2 : l == 3 ? wb3 : wb4;
    uint16_t *wt = l == 2 ? wt2 : l == 3 ? wt3 : wt4;
    wb[(size_t)o * kW + k] = h;
    wt[(size_t)k * kW + o] = h;
  } else if (i >= (int64_t)MlpLayout::kW5 && i < (int64_t)MlpLayout::kB5) {
    w5b[k] = h;
  }
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

// training workspace for batch B (multiple of 128), carved from m->ws. The
// accumulators and the loss come first, at the same offset for every batch:
// they are zeroed when the workspace is allocated and cleared again by each
// step's update kernel after it reads them.
int train_split(int l, int B) { return std::min(l == 1 ? 32 : 16, B / kKC); }

size_t carve(void *base, int B, TrainDev *T, float **loss) {
  char *p = (char *)base;
  auto take = [&](size_t bytes) {
    char *r = p;
    p += align256(bytes);
    return (void *)r;
  };
  const size_t act = (size_t)B * kW;
  TrainDev t;
  memset(&t, 0, sizeof(t));
  t.B = B;
  t.acc = (float *)take((kAccLen + 1) * 4);
  float *lp = t.acc + kAccLen;
  t.at[0] = (uint16_t *)take((size_t)B * kIn * 2);
  for (int l = 1; l < kHid; ++l) t.at[l] = (uint16_t *)take(act * 2);
  for (int l = 1; l <= kHid; ++l) t.dt[l] = (uint16_t *)take(act * 2);
  for (int l = 1; l <= kHid; ++l) {
    t.split[l] = train_split(l, B);
    t.wparts[l] = (float *)take((size_t)t.split[l] * kW * MlpLayout::K(l) * 4);
  }
  if (T) *T = t;
  if (loss) *loss = lp;
  return (size_t)(p - (char *)base);
}

// forward + backward for one batch, then the update: the gradient into m->g
// and, with lr > 0, an Adam step; the loss into *loss_dev (workspace)
int train_launches(csd_mlp *m, const float *q, const float *target, int B, float lr,
                   hipStream_t st, float **loss_dev) {
  if (B > m->cap) {
    if (m->ws) hipFree(m->ws);
    m->ws = nullptr;
    m->cap = 0;
    const size_t bytes = carve(nullptr, B, nullptr, nullptr);
    hipError_t e = hipMalloc(&m->ws, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc (MLP training workspace)");
    e = hipMemset(m->ws, 0, (kAccLen + 1) * 4);  // accumulators start at zero
    if (e != hipSuccess) return hip_fail(e, "hipMemset (MLP training workspace)");
    m->cap = B;
    m->ws_bytes = bytes;
  }
  TrainDev T;
  float *loss = nullptr;
  carve(m->ws, B, &T, &loss);
  *loss_dev = loss;
  const MlpDev M = m->dev();
  RowArgs R;
  R.B = B;
  R.q = q;
  R.target = target;
  R.at0 = T.at[0];
  R.at1 = T.at[1];
  R.at2 = T.at[2];
  R.at3 = T.at[3];
  R.dt1 = T.dt[1];
  R.dt2 = T.dt[2];
  R.dt3 = T.dt[3];
  R.dt4 = T.dt[4];
  R.acc = T.acc;
  hipLaunchKernelGGL(k_train_rows, dim3(B / kRB), dim3(kRowThreads), 0, st, R, M, m->wt[2], m->wt[3],
                     m->wt[4]);
  WgradArgs W;
  memset(&W, 0, sizeof(W));
  W.B = B;
  int nb = 0;
  for (int l = kHid; l >= 1; --l) {  // blocks of layers 4, 3, 2, 1 in that order
    W.first[l] = nb;
    W.split[l] = T.split[l];
    W.dt[l] = T.dt[l];
    W.at[l] = T.at[l - 1];
    W.wparts[l] = T.wparts[l];
    const int tiles = l == 1 ? (kW / 128) * (kIn / 32) : (kW / 64) * (kW / 64);
    nb += tiles * T.split[l];
  }
  W.first[0] = nb;
  hipLaunchKernelGGL(k_train_wgrad, dim3(nb), dim3(256), 0, st, W);
  UpdateArgs U;
  U.B = B;
  for (int l = 1; l <= kHid; ++l) {
    U.wparts[l - 1] = T.wparts[l];
    U.split[l - 1] = T.split[l];
  }
  U.acc = T.acc;
  float bc1 = 1.f, bc2 = 1.f;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  if (lr > 0.f) {
    m->step += 1;
    bc1 = (float)(1.0 - pow((double)b1, (double)m->step));
    bc2 = (float)(1.0 - pow((double)b2, (double)m->step));
  }
  const int64_t P = MlpLayout::kTotal;
  hipLaunchKernelGGL(k_train_update, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, U, m->p,
                     m->g, m->m, m->v, lr, b1, b2, eps, bc1, bc2, m->wb[1], m->wb[2], m->wb[3],
                     m->wb[4], m->wt[2], m->wt[3], m->wt[4], m->w5b, loss);
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
  float *ld = nullptr;
  rc = train_launches(m, q, target, batch, 0.f, st, &ld);
  if (rc) return rc;
  if (grads) CSD_HIP(hipMemcpyAsync(grads, m->g, MlpLayout::kTotal * 4, hipMemcpyDeviceToDevice, st));
  if (loss) CSD_HIP(hipMemcpyAsync(loss, ld, 4, hipMemcpyDeviceToDevice, st));
  return CSD_OK;
}

int csd_mlp_train_step(csd_mlp *m, const float *q, const float *target, int32_t batch, float lr,
                       float *loss, void *stream) {
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if (!q || !target) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  if (!(lr > 0.f)) return fail(CSD_ERR_INVALID_ARGUMENT, "learning rate must be positive");
  hipStream_t st = (hipStream_t)stream;
  float *ld = nullptr;
  rc = train_launches(m, q, target, batch, lr, st, &ld);
  if (rc) return rc;
  if (loss) CSD_HIP(hipMemcpyAsync(loss, ld, 4, hipMemcpyDeviceToDevice, st));
  return CSD_OK;
}

}  // extern "C"

int csd::mlp_predict_staged(const csd_mlp *m, const float *stage, int num_blocks, int dim,
                            float *out, hipStream_t st, const uint32_t *live_slots,
                            int blocks_per_slot) {
  if (!m) return fail(CSD_ERR_INVALID_ARGUMENT, "null surrogate");
  if (dim != m->in_dim) return fail(CSD_ERR_INVALID_ARGUMENT, "surrogate input dimension differs from the plant's");
  const int64_t n = (int64_t)num_blocks * 256;
  if (n == 0) return CSD_OK;
  const int64_t blocks = (n + 4 * kRowsWave - 1) / (4 * kRowsWave);
  hipLaunchKernelGGL((k_mlp_infer<StagedInput>), dim3((unsigned)blocks), dim3(kBlock), 0, st,
                     m->dev(), StagedInput{stage, dim, live_slots, (int64_t)blocks_per_slot * 256},
                     n, out);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

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
