// csd_eizo.hip -- batched EI-ZO edge inflation (the round driver).
//
// Reference algorithm: planning/cuda_edge_inflation_zero_order.cu:94-339
// (executeEizo) with the kernels of planning/cuda_set_builder_utils.cu and
// geometry/cuda_hit_and_run_sampling.cu. Differences, all deliberate:
//  * many regions (edges) advance round by round together; one launch per
//    stage per round serves every active region;
//  * chain i of N starts at fraction (i + 1/2) / N of the segment instead of
//    at line_start;
//  * a lane runs one hit-and-run chain; the samples of a block stay in that
//    block's staging segment, where the collision check compacts the
//    colliding ones in place;
//  * colliding samples are taken in sample-index order (block ballot +
//    prefix sums) instead of atomicAdd order, so results are deterministic;
//  * randomness is Philox4x32-10 keyed by the seed, with counters
//    (sample, step, region, round*16 + k): every round draws fresh samples;
//  * all bisection steps of a particle run in one lane with FK + collision
//    check inline;
//  * face placement and the redundancy test run in FP64, one workgroup per
//    region, all faces of a round in one launch; b is rounded up when stored
//    in FP32 so the segment stays inside the stored face.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

#include "csd_eizo_device.h"
#include "csd_mlp_internal.h"

namespace csd {

// Round stage 1a: hit-and-run chains, one sample per lane, written
// unchecked into the block's staging segment as [dim][256] (coalesced).
// Kept apart from the collision check so that the chains run at the high
// occupancy their small register footprint allows.
template <int DPL, bool BROW = false, int BLK = kSampleBlock, int DIMC = 0>
__global__ __launch_bounds__(BLK) void k_eizo_sample(EizoDev D, int dim, int cap, int nblk,
                                                     int mixing, uint32_t k0, uint32_t k1,
                                                     uint32_t round, uint32_t stream_off) {
  constexpr int DP = SamplerShape<DPL>::DP;
  constexpr bool LDS = SamplerShape<DPL>::LDS;
  // a block of BLK threads fills BLK / 256 consecutive staging segments and
  // shares one LDS copy of the polytope among them
  constexpr int kSegs = BLK / kSampleBlock;
  extern __shared__ float csd_smem[];
  const int slot = blockIdx.y;
  if (D.ctl) {  // device-resident round: slots past the active count idle
    if (slot >= (int)D.ctl[1]) return;
    round = D.ctl[2];
  }
  const int r = D.active[slot];
  const uint32_t N = D.nsamp[r];
  if (blockIdx.x * BLK >= N) return;  // whole block past this region's samples
  const int F = D.nf[r];
  const float *A = D.A + (size_t)r * cap * dim, *b = D.b + (size_t)r * cap;
  if constexpr (LDS) {
    stage_polytope<DP, BROW>(csd_smem, csd_smem + (size_t)F * DP, A, b, F, dim);
    __syncthreads();
    A = csd_smem;
    b = csd_smem + (size_t)F * DP;
  }
  const int sblk = (int)blockIdx.x * kSegs + (int)threadIdx.x / kSampleBlock;
  const int lane = (int)threadIdx.x % kSampleBlock;
  // a segment past the round's staging or this region's samples: nothing
  // to draw (the same chains run whatever the block size)
  if (sblk >= nblk || (uint32_t)sblk * kSampleBlock >= N) return;
  const uint32_t i = (uint32_t)sblk * kSampleBlock + lane;
  // chain i starts at fraction (i + 1/2) / N of the segment, which lies in
  // the region: spread starts cover elongated regions better than starts at
  // line_start only
  const float f = ((float)i + 0.5f) / (float)N;
  float x[DP];
#pragma unroll
  for (int d = 0; d < DP; ++d) {
    const float a = d < dim ? D.ls[(size_t)r * dim + d] : 0.f;
    const float e = d < dim ? D.le[(size_t)r * dim + d] : 0.f;
    x[d] = fmaf(f, e - a, a);
  }
  hit_and_run<DP, LDS, BROW, DIMC>(A, b, F, dim, x, i, (uint32_t)r + stream_off, round, mixing,
                                   k0, k1);
  float *seg = D.stage + ((size_t)slot * nblk + sblk) * kSampleBlock * dim;
#pragma unroll
  for (int d = 0; d < DP; ++d)
    if (d < dim) seg[(size_t)d * kSampleBlock + lane] = x[d];
}

// Sampler tile for polytopes of up to `faces` faces: staged in LDS with
// rows padded to 8 or 16 floats (b in the padding slot when dim < DP) by
// 256-thread blocks up to 64 KB, or by 1024-thread blocks -- 16 waves
// sharing one copy -- up to the 160 KB of a gfx950 workgroup; larger
// polytopes are read from global memory. 1024-thread blocks are taken from
// `wide_from` faces on (csd_eizo_tune_sampler measures it) and whenever
// only they fit.
struct SamplerLaunch {
  int dpl = 32;  // SamplerShape: 2 DP + LDS
  bool brow = false;
  int blk = kSampleBlock;
  size_t lds = 0;
};

SamplerLaunch pick_sampler(int dim, int faces, int wide_from) {
  SamplerLaunch L;
  for (int dp : {8, 16}) {
    if (dim > dp) continue;
    const size_t bytes = dp == 8 ? polytope_lds_bytes<8>(faces) : polytope_lds_bytes<16>(faces);
    const bool narrow = bytes <= kLdsLimit, wide = bytes <= kLdsLimitWide;
    if (!narrow && !wide) continue;
    L.dpl = 2 * dp + 1;
    L.brow = dim < dp;
    L.blk = narrow && faces < wide_from ? kSampleBlock : kSampleBlockWide;
    L.lds = bytes;
    return L;
  }
  return L;  // global memory, 256-thread blocks
}

// k_eizo_sample for `slots` regions and `nblk` staging segments each
void launch_sampler(const SamplerLaunch &L, int nblk, int slots, hipStream_t s,
                    const EizoDev &E, int dim, int cap, int mixing, uint32_t k0, uint32_t k1,
                    uint32_t round, uint32_t off) {
  const int segs = L.blk / kSampleBlock;
  const dim3 g((nblk + segs - 1) / segs, slots), bk(L.blk);
#define CSD_SAMPLE(DPL, BR, BLK)                                                              \
  hipLaunchKernelGGL((k_eizo_sample<DPL, BR, BLK>), g, bk, L.lds, s, E, dim, cap, nblk, mixing, \
                     k0, k1, round, off)
  // 14 dimensions (two 7-DOF arms): rows of 16 with b in slot 15, and the
  // dot products skip the zero slot 14
#define CSD_SAMPLE14(BLK)                                                                     \
  hipLaunchKernelGGL((k_eizo_sample<33, true, BLK, 14>), g, bk, L.lds, s, E, dim, cap, nblk,   \
                     mixing, k0, k1, round, off)
  if (L.dpl == 32) CSD_SAMPLE(32, false, kSampleBlock);
  else if (L.blk == kSampleBlock) {
    if (L.dpl == 17) {
      if (L.brow) CSD_SAMPLE(17, true, kSampleBlock);
      else CSD_SAMPLE(17, false, kSampleBlock);
    } else {
      if (L.brow && dim == 14) CSD_SAMPLE14(kSampleBlock);
      else if (L.brow) CSD_SAMPLE(33, true, kSampleBlock);
      else CSD_SAMPLE(33, false, kSampleBlock);
    }
  } else {
    if (L.dpl == 17) {
      if (L.brow) CSD_SAMPLE(17, true, kSampleBlockWide);
      else CSD_SAMPLE(17, false, kSampleBlockWide);
    } else {
      if (L.brow && dim == 14) CSD_SAMPLE14(kSampleBlockWide);
      else if (L.brow) CSD_SAMPLE(33, true, kSampleBlockWide);
      else CSD_SAMPLE(33, false, kSampleBlockWide);
    }
  }
#undef CSD_SAMPLE14
#undef CSD_SAMPLE
}

// Allows the 1024-thread sampler tiles their dynamic LDS above 64 KB.
int allow_wide_sampler_lds() {
  static const int rc = [] {
    const void *fns[] = {
        reinterpret_cast<const void *>(&k_eizo_sample<17, true, kSampleBlockWide>),
        reinterpret_cast<const void *>(&k_eizo_sample<17, false, kSampleBlockWide>),
        reinterpret_cast<const void *>(&k_eizo_sample<33, true, kSampleBlockWide>),
        reinterpret_cast<const void *>(&k_eizo_sample<33, false, kSampleBlockWide>),
        reinterpret_cast<const void *>(&k_eizo_sample<33, true, kSampleBlockWide, 14>)};
    for (const void *f : fns)
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kLdsLimitWide) != hipSuccess)
        return (int)CSD_ERR_HIP;
    return (int)CSD_OK;
  }();
  return rc;
}

// Round stage 1b: collision check of the block's samples and compaction of
// the colliding ones, in sample order, to the front of the same segment as
// [count][dim] (block ballot + prefix sums; all samples are read before the
// first write). With `pred` (a pilot pruned by the surrogate): only samples
// whose prediction is below `pred_below` get the exact check, the others
// count as free, and blk_first counts the samples checked instead of the
// collisions among the first M.
template <int VAR, int DP>
__global__ __launch_bounds__(kSampleBlock) void k_eizo_check(PlantImage P, VoxelArgs V,
                                                             RobotGeomList R, EizoDev D, int dim,
                                                             int nblk, uint32_t M,
                                                             const float *__restrict__ pred,
                                                             float pred_below) {
  const int slot = blockIdx.y;
  if (D.ctl) {
    if (slot >= (int)D.ctl[1]) return;
    M = D.ctl[3];
  }
  const int r = D.active[slot];
  const uint32_t N = D.nsamp[r];
  if (blockIdx.x * kSampleBlock >= N) {
    if (threadIdx.x == 0) {
      D.blk_cnt[(size_t)slot * nblk + blockIdx.x] = 0;
      D.blk_first[(size_t)slot * nblk + blockIdx.x] = 0;
    }
    return;
  }
  extern __shared__ __attribute__((aligned(16))) float csd_smem[];
  const float *tab = stage_tables<VAR>(P, csd_smem);
  const uint32_t i = blockIdx.x * kSampleBlock + threadIdx.x;
  const bool valid = i < N;
  float *seg = D.stage + ((size_t)slot * nblk + blockIdx.x) * kSampleBlock * dim;
  float x[DP];
#pragma unroll
  for (int d = 0; d < DP; ++d) x[d] = d < dim ? seg[(size_t)d * kSampleBlock + threadIdx.x] : 0.f;
  const bool check =
      valid && (!pred || pred[((size_t)slot * nblk + blockIdx.x) * kSampleBlock + threadIdx.x] <
                             pred_below);
  // a wave with nothing to check skips the check (and its forward kinematics)
  const bool col = __any(check) && collides_variant<VAR, false, 3 | kCoherent | kGrouped>(P, QReg<DP>{x}, V, R, !check, tab) && check;

  __shared__ uint32_t wcnt[kSampleBlock / 64], wfirst[kSampleBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t m = __ballot(col);
  const uint64_t mf = pred ? __ballot(check) : __ballot(col && i < M);
  if (lane == 0) {
    wcnt[wave] = (uint32_t)__popcll(m);
    wfirst[wave] = (uint32_t)__popcll(mf);
  }
  __syncthreads();  // also orders every read of seg before the writes below
  uint32_t off = 0, tot = 0, totf = 0;
#pragma unroll
  for (int w = 0; w < kSampleBlock / 64; ++w) {
    off += w < wave ? wcnt[w] : 0;
    tot += wcnt[w];
    totf += wfirst[w];
  }
  if (col) {
    const uint32_t rank = off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    float *dst = seg + (size_t)rank * dim;
#pragma unroll
    for (int d = 0; d < DP; ++d)
      if (d < dim) dst[d] = x[d];
  }
  if (threadIdx.x == 0) {
    D.blk_cnt[(size_t)slot * nblk + blockIdx.x] = tot;
    D.blk_first[(size_t)slot * nblk + blockIdx.x] = totf;
  }
}

// Stage 2: unadaptive test (cuda_edge_inflation_zero_order.cu:165-199), when
// `test`, and gather of the first budget_r colliding samples, one workgroup
// per region. Without the test the collision count goes to pilot_col.
__global__ __launch_bounds__(256) void k_eizo_compact(EizoDev D, int dim, int nblk, uint32_t M,
                                                      float thresh, int U, int test,
                                                      uint32_t round) {
  const int slot = blockIdx.x;
  if (D.ctl) {
    if (slot >= (int)D.ctl[1]) return;
    M = D.ctl[3];
    thresh = __uint_as_float(D.ctl[4]);
    round = D.ctl[2];
  }
  const int r = D.active[slot];
  const uint32_t N = D.nsamp[r];
  __shared__ uint32_t s_off[1024];
  __shared__ uint32_t s_tot, s_first;
  __shared__ int s_u;
  if (threadIdx.x == 0) {
    uint32_t acc = 0, first = 0;
    for (int k = 0; k < nblk; ++k) {
      if (k < 1024) s_off[k] = acc;
      acc += D.blk_cnt[(size_t)slot * nblk + k];
      first += D.blk_first[(size_t)slot * nblk + k];
    }
    s_tot = acc;
    s_first = first;
    D.samples[r] += N;
    D.collisions[r] += acc;
    const bool certified = test == 1 && (float)first <= thresh;
    if (certified) D.status[r] = 1;
    if (test != 1) D.pilot_col[r] = acc;
    s_u = certified ? 0 : (int)min(acc, min(D.budget[r], (uint32_t)U));
    D.u[r] = s_u;
    if (int32_t *st = round_stats(D, r, round)) {
      if (test == 1) {
        st[kStatSamples] = (int32_t)N;
        st[kStatFirstCol] = (int32_t)first;
        st[kStatCol] = (int32_t)acc;
      } else {  // pilot; a later test phase of the round overwrites these
        st[kStatPilot] = (int32_t)N;
        st[kStatPilotCol] = (int32_t)acc;
        st[kStatPilotChecked] = test == 2 ? (int32_t)first : (int32_t)N;
        st[kStatSamples] = 0;
        st[kStatFirstCol] = -1;
        st[kStatCol] = 0;
      }
      st[kStatUpdated] = s_u;
      st[kStatFacesStart] = D.nf[r];
    }
  }
  __syncthreads();
  const int u = s_u;
  if (u == 0) return;
  for (int k = 0; k < nblk; ++k) {
    const uint32_t off = s_off[k];
    if ((int)off >= u) break;
    const uint32_t cnt = D.blk_cnt[(size_t)slot * nblk + k];
    const float *src = D.stage + (((size_t)slot * nblk + k) * kSampleBlock) * dim;
    for (uint32_t e = threadIdx.x; e < cnt * dim; e += blockDim.x) {
      const uint32_t j = e / dim;
      if ((int)(off + j) < u) D.part[((size_t)r * U + off + j) * dim + (e - j * dim)] = src[e];
    }
  }
}

// Surrogate ranking: from the surrogate's values at L points between each
// particle's projection and the particle (pred [slot][cand][L], positive =
// free), the first point predicted in collision gives an estimate of how far
// along the way the boundary lies (linear interpolation; the particle itself
// collides). Particles are ordered by that estimated distance to the segment
// and the `keep` nearest of the first `cand` stay, in that order (ties by
// index), as the particles to bisect.
constexpr int kRankBlock = 1024;
constexpr int kRankMax = 1024;  // particles ranked per region (the first ones)

__global__ __launch_bounds__(kRankBlock) void k_eizo_rank(EizoDev D, const float *__restrict__ pred,
                                                          int dim, int U, int L, int cand,
                                                          int keep, float *__restrict__ tmp,
                                                          uint32_t round) {
  const int slot = blockIdx.x;
  if (D.ctl) {  // device-resident round: slots past the active count idle
    if (slot >= (int)D.ctl[1]) return;
    round = D.ctl[2];
  }
  const int r = D.active[slot];
  if (D.status[r] != 0) return;
  const int u = D.u[r];
  if (u <= keep) return;
  const int n = min(u, cand);
  __shared__ float key[kRankMax];
  __shared__ int idx[kRankMax];
  const float *q1 = D.ls + (size_t)r * dim, *q2 = D.le + (size_t)r * dim;
  for (int i = threadIdx.x; i < n; i += kRankBlock) {
    const float *p = D.part + ((size_t)r * U + i) * dim;
    float a = 0.f, bb = 0.f;
    for (int d = 0; d < dim; ++d) {
      a = fmaf(p[d] - q1[d], q2[d] - q1[d], a);
      bb = fmaf(q2[d] - q1[d], q2[d] - q1[d], bb);
    }
    const float t = fminf(fmaxf(a / bb, 0.f), 1.f);
    float len2 = 0.f;
    for (int d = 0; d < dim; ++d) {
      const float e = p[d] - ((1.f - t) * q1[d] + t * q2[d]);
      len2 = fmaf(e, e, len2);
    }
    const float *pv = pred + ((size_t)slot * cand + i) * L;
    float prev = 1.f, frac = 1.f;  // the projection is taken as free
    for (int l = 0; l < L; ++l) {
      const float v = pv[l];
      if (v < 0.f) {
        const float f0 = (float)l / (float)(L + 1), f1 = (float)(l + 1) / (float)(L + 1);
        frac = f0 + (f1 - f0) * (prev / (prev - v));
        break;
      }
      prev = v;
    }
    key[i] = frac * sqrtf(len2);
  }
  // bitonic sort of (key, index) over the next power of two >= n, ascending
  // key, lower index first on ties; padding sorts last
  int P2 = 1;
  while (P2 < n) P2 <<= 1;
  for (int i = threadIdx.x; i < P2; i += kRankBlock) {
    if (i >= n) key[i] = INFINITY;
    idx[i] = i < n ? i : 0x7fffffff;
  }
  __syncthreads();
  for (int k2 = 2; k2 <= P2; k2 <<= 1)
    for (int j = k2 >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P2; i += kRankBlock) {
        const int l = i ^ j;
        if (l > i) {
          const float ki = key[i], kl = key[l];
          const int ii = idx[i], il = idx[l];
          const bool after = ki > kl || (ki == kl && ii > il);
          if (after == ((i & k2) == 0)) {
            key[i] = kl;
            key[l] = ki;
            idx[i] = il;
            idx[l] = ii;
          }
        }
      }
      __syncthreads();
    }
  const int k = min(keep, n);
  // reorder: selected particles to the front (through tmp [slot][kRankMax][dim])
  float *t = tmp + (size_t)slot * kRankMax * dim;
  for (int e = threadIdx.x; e < k * dim; e += kRankBlock) {
    const int j = e / dim, d = e - j * dim;
    t[e] = D.part[((size_t)r * U + idx[j]) * dim + d];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < k * dim; e += kRankBlock)
    D.part[(size_t)r * U * dim + e] = t[e];
  if (threadIdx.x == 0) {
    D.u[r] = k;
    if (int32_t *st = round_stats(D, r, round)) st[kStatUpdated] = k;
  }
}

// Stage 3: projection + all bisection steps per lane; FP64 distance from
// the last colliding point to the (pre-bisection) projection. DPS = 8 DP + S:
// S waves share each 64 particles, each testing part of the pairs, when a
// launch has few waves (eizo_bisect_split).
template <int VAR, int DPS>
__global__ __launch_bounds__(kSampleBlock) void k_eizo_bisect(PlantImage P, VoxelArgs V,
                                                              RobotGeomList R, EizoDev D,
                                                              int dim, int U, int steps) {
  constexpr int DP = DPS / 8, S = DPS % 8;
  constexpr int kPer = kSampleBlock / S;
  if (D.ctl && blockIdx.y >= D.ctl[1]) return;
  const int r = D.active[blockIdx.y];
  if (D.status[r] != 0) return;
  const int u = D.u[r];
  if ((int)(blockIdx.x * kPer) >= u) return;  // whole block past this region's particles
  extern __shared__ __attribute__((aligned(16))) float csd_smem[];
  __shared__ uint8_t verdict[kSampleBlock];  // the parts' verdicts (S > 1)
  const float *tab = stage_tables<VAR>(P, csd_smem);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = wave % S, grp = wave / S;
  const int i = blockIdx.x * kPer + grp * 64 + lane;
  const bool valid = i < u;
  const int ii = valid ? i : u - 1;
  float p[DP], q1[DP], q2[DP], sc[DP];
#pragma unroll
  for (int d = 0; d < DP; ++d) {
    p[d] = d < dim ? D.part[((size_t)r * U + ii) * dim + d] : 0.f;
    q1[d] = d < dim ? D.ls[(size_t)r * dim + d] : 0.f;
    q2[d] = d < dim ? D.le[(size_t)r * dim + d] : 0.f;
  }
  // the projection goes to D.proj before the steps and is read back for the
  // distance (this lane's own stores; part 0 of the particle's waves owns it)
  const bool owner = valid && sub == 0;
  float *proj = D.proj + ((size_t)r * U + i) * dim;
  bisect_particle<VAR, DP, S>(P, V, R, p, q1, q2, steps, !valid, sc, owner ? proj : nullptr,
                              dim, tab, sub, verdict);
  if (!owner) return;
  double dd = 0.0;
#pragma unroll
  for (int d = 0; d < DP; ++d) {
    if (d < dim) {
      D.scol[((size_t)r * U + i) * dim + d] = sc[d];
      const double e = (double)sc[d] - (double)proj[d];
      dd += e * e;
    }
  }
  D.dist[(size_t)r * U + i] = sqrt(dd);
}

// Waves per 64 particles of k_eizo_bisect: 4 or 2 when the launch would
// otherwise have fewer than 2,048 or 4,096 waves (two or four per SIMD), so
// that each particle's pair tests are split over several waves and the
// steps' latency overlaps; else 1 (the thresholds of csd_bisect's split).
inline int eizo_bisect_split(int regions, uint32_t particles) {
  const int64_t waves = (int64_t)regions * ((particles + 63) / 64);
  return waves < 2048 ? 4 : waves < 4096 ? 2 : 1;
}

// k_eizo_bisect over `regions` slots of up to `particles` particles each
#define CSD_EIZO_BISECT(regions, particles, stream, D_)                                        \
  do {                                                                                         \
    const int split_ = eizo_bisect_split((regions), (particles));                              \
    const int per_ = kSampleBlock / split_;                                                    \
    const dim3 g_((unsigned)(((particles) + per_ - 1) / per_), (unsigned)(regions));           \
    const int st_ = (int)o.bisection_steps;                                                    \
    if (dp == 8 && split_ == 4)                                                                \
      CSD_DISPATCH_VARIANT_X(var, k_eizo_bisect, 68, g_, dim3(kSampleBlock), tab_bytes,        \
                             stream, P, V, RL, D_, dim, U, st_);                               \
    else if (dp == 8 && split_ == 2)                                                           \
      CSD_DISPATCH_VARIANT_X(var, k_eizo_bisect, 66, g_, dim3(kSampleBlock), tab_bytes,        \
                             stream, P, V, RL, D_, dim, U, st_);                               \
    else if (dp == 8)                                                                          \
      CSD_DISPATCH_VARIANT_X(var, k_eizo_bisect, 65, g_, dim3(kSampleBlock), tab_bytes,        \
                             stream, P, V, RL, D_, dim, U, st_);                               \
    else if (split_ == 4)                                                                      \
      CSD_DISPATCH_VARIANT_X(var, k_eizo_bisect, 132, g_, dim3(kSampleBlock), tab_bytes,       \
                             stream, P, V, RL, D_, dim, U, st_);                               \
    else if (split_ == 2)                                                                      \
      CSD_DISPATCH_VARIANT_X(var, k_eizo_bisect, 130, g_, dim3(kSampleBlock), tab_bytes,       \
                             stream, P, V, RL, D_, dim, U, st_);                               \
    else                                                                                       \
      CSD_DISPATCH_VARIANT_X(var, k_eizo_bisect, 129, g_, dim3(kSampleBlock), tab_bytes,       \
                             stream, P, V, RL, D_, dim, U, st_);                               \
  } while (0)

// Device-resident round start (one workgroup): the regions still running
// (status 0), in region order, become this round's slots, with this round's
// sample count -- max(M_k, budget), or `pilot` when the round starts with a
// pilot -- and round count; ctl takes the round number, M_k, the test
// threshold and the slot count for the round's other stages. Past the last
// round the slot count is 0 and every stage idles.
constexpr int kBeginBlock = 1024;

__global__ __launch_bounds__(kBeginBlock) void k_eizo_round_begin(EizoDev D, int Rn,
                                                                  uint32_t rounds,
                                                                  uint32_t pilot) {
  __shared__ uint32_t s_w[kBeginBlock / 64];
  __shared__ uint32_t s_base;
  const uint32_t it = D.ctl[0];
  const bool live = it < rounds;
  const uint32_t M = live ? D.mtab[it] : 0u;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int r0 = 0; r0 < Rn; r0 += kBeginBlock) {
    const int r = r0 + (int)threadIdx.x;
    const bool act = live && r < Rn && D.status[r] == 0;
    const uint64_t m = __ballot(act);
    if (lane == 0) s_w[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = s_base;
    for (int w = 0; w < wave; ++w) off += s_w[w];
    if (act) {
      D.active[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = r;
      D.nsamp[r] = pilot ? pilot : max(M, D.budget[r]);
      D.iters[r] += 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int w = 0; w < kBeginBlock / 64; ++w) t += s_w[w];
      s_base += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    D.ctl[1] = s_base;
    D.ctl[2] = it;
    D.ctl[3] = M;
    D.ctl[4] = live ? __float_as_uint(D.ttab[it]) : 0u;
    if (live) D.ctl[0] = it + 1;
  }
}

// Device-resident pilot decision (one workgroup), after the pilot phase:
// the round's slots whose pilot collision count is at most `limit` keep the
// unadaptive test; they become the slots of T (T.active, in slot order) with
// the sample count max(M_k, budget), and T.ctl takes the slot count and the
// round's number, M_k and threshold. The others take their particles from
// the pilot. The host loop's rule, with the same float comparison.
__global__ __launch_bounds__(kBeginBlock) void k_eizo_pilot_select(EizoDev D, EizoDev T,
                                                                   float limit) {
  __shared__ uint32_t s_w[kBeginBlock / 64];
  __shared__ uint32_t s_base;
  const uint32_t na = D.ctl[1], M = D.ctl[3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (uint32_t s0 = 0; s0 < na; s0 += kBeginBlock) {
    const uint32_t s = s0 + threadIdx.x;
    const int r = s < na ? D.active[s] : 0;
    const bool keep = s < na && (float)D.pilot_col[r] <= limit;
    const uint64_t m = __ballot(keep);
    if (lane == 0) s_w[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = s_base;
    for (int w = 0; w < wave; ++w) off += s_w[w];
    if (keep) {
      T.active[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = r;
      D.nsamp[r] = max(M, D.budget[r]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int w = 0; w < kBeginBlock / 64; ++w) t += s_w[w];
      s_base += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    T.ctl[1] = s_base;
    T.ctl[2] = D.ctl[2];
    T.ctl[3] = M;
    T.ctl[4] = D.ctl[4];
  }
}

}  // namespace csd

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
struct csd_eizo {
  const csd_plant *plant = nullptr;
  csd_eizo_options opt;
  std::vector<int32_t> robot;
  int R = 0, cap = 0, U = 0, Vmax = 0, dim = 0;
  csd::EizoDev D;
  float *vox = nullptr;
  void *arena = nullptr;
  size_t stage_bytes = 0, blk_entries = 0;
  std::vector<void *> owned;
  // guidance (csd_eizo_set_guidance)
  uint32_t pilot = 0;
  float pilot_factor = 0.f;
  const csd_mlp *surrogate = nullptr;
  int keep = 64, line_points = 4;
  float *pred = nullptr;  // [slots][cand][L] surrogate outputs
  size_t pred_bytes = 0;
  // pilot pruned by the surrogate (csd_eizo_set_pruning): only samples
  // predicted below prune_below are checked exactly
  bool prune = false;
  float prune_below = 0.f;
  float *ppred = nullptr;  // [slots][nblk][256] predictions at the pilot samples
  size_t ppred_bytes = 0;
  // device-resident rounds (csd_eizo_set_graph): host status read every
  // graph_sync rounds, 0 for host-driven rounds
  int graph_sync = 0;
  hipStream_t gs = nullptr;
  uint32_t *ctl = nullptr, *mtab = nullptr;
  float *ttab = nullptr;
  // the test phase of a device-resident guided round: its slots and control block
  int32_t *active2 = nullptr;
  uint32_t *ctl2 = nullptr;
  // the last csd_eizo_run: regions, rounds per region, wall time per round
  int last_regions = 0;
  std::vector<int32_t> last_rounds;
  std::vector<float> round_ms;
  // csd_eizo_set_metric
  double *metric_buf = nullptr;
  // sampler tile: 1024-thread blocks from this many faces on (when 256-thread
  // blocks also fit); csd_eizo_tune_sampler / csd_eizo_set_sampler_tile
  int wide_from = std::numeric_limits<int>::max();
};

using namespace csd;

namespace {
template <class T>
int dalloc(csd_eizo *w, T **p, size_t count) {
  void *q = nullptr;
  CSD_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
  w->owned.push_back(q);
  *p = (T *)q;
  return CSD_OK;
}

// Samples of the unadaptive test in round `it`
// (cuda_edge_inflation_zero_order.cu:113-119, float arithmetic).
uint32_t unadaptive_samples(const csd_eizo_options &o, uint32_t it) {
  const float delta_k = (float)(6 * o.delta / (M_PI * M_PI * (it + 1) * (it + 1)));
  return (uint32_t)ceilf(2 * logf(1 / delta_k) / (o.epsilon * o.tau * o.tau));
}

// Grows the staging segments and block counters to na regions x nblk blocks.
int grow_stage(csd_eizo *w, int na, int nblk) {
  EizoDev &D = w->D;
  const size_t stage_need =
      std::max((size_t)na * nblk * kSampleBlock, (size_t)na * 256) * w->dim * sizeof(float);
  if (stage_need > w->stage_bytes) {
    if (D.stage) hipFree(D.stage);
    D.stage = nullptr;
    w->stage_bytes = 0;
    CSD_HIP(hipMalloc(&D.stage, stage_need));
    w->stage_bytes = stage_need;
  }
  const size_t blk_need = (size_t)na * nblk;
  if (blk_need > w->blk_entries) {
    if (D.blk_cnt) hipFree(D.blk_cnt);
    if (D.blk_first) hipFree(D.blk_first);
    D.blk_cnt = D.blk_first = nullptr;
    w->blk_entries = 0;
    CSD_HIP(hipMalloc(&D.blk_cnt, blk_need * 4));
    CSD_HIP(hipMalloc(&D.blk_first, blk_need * 4));
    w->blk_entries = blk_need;
  }
  return CSD_OK;
}
}  // namespace

extern "C" {

int csd_eizo_create(const csd_plant *plant, const int32_t *robot_geoms, int32_t num_robot_geoms,
                    const csd_eizo_options *options, int32_t max_regions, int32_t face_capacity,
                    int32_t max_particles, int32_t max_voxels, csd_eizo **out) {
  if (!plant || !options || !out) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const int dim = plant->img.dof;
  if (dim < 1 || dim > kDimMax) return fail(CSD_ERR_CAPACITY, "EI-ZO supports 1..16 dims");
  if (plant->img.num_bodies > 16) return fail(CSD_ERR_CAPACITY, "more than 16 moving joints");
  if (max_regions < 1 || face_capacity < 1 || max_particles < 1 || max_voxels < 0)
    return fail(CSD_ERR_INVALID_ARGUMENT, "non-positive capacity");
  if (num_robot_geoms < 0 || num_robot_geoms > kMaxRobotGeoms)
    return fail(CSD_ERR_CAPACITY, "robot geometry list longer than 128");
  if (options->tau <= 0.f || options->tau >= 1.f || options->epsilon <= 0.f ||
      options->delta <= 0.f)
    return fail(CSD_ERR_INVALID_ARGUMENT, "tau must be in (0,1); epsilon, delta positive");
  csd_eizo *w = new csd_eizo();
  w->plant = plant;
  w->opt = *options;
  w->robot.assign(robot_geoms, robot_geoms + num_robot_geoms);
  w->R = max_regions;
  w->cap = face_capacity;
  w->U = max_particles;
  w->Vmax = max_voxels;
  w->dim = dim;
  if (allow_wide_sampler_lds() != CSD_OK) {
    delete w;
    return fail(CSD_ERR_HIP, "hipFuncSetAttribute (sampler LDS)");
  }
  const size_t R = max_regions, C = face_capacity, U = max_particles;
  int rc = CSD_OK;
  EizoDev &D = w->D;
  memset(&D, 0, sizeof(D));
  if (!rc) rc = dalloc(w, &D.A, R * C * dim);
  if (!rc) rc = dalloc(w, &D.b, R * C);
  if (!rc) rc = dalloc(w, &D.nf, R);
  if (!rc) rc = dalloc(w, &D.ls, R * dim);
  if (!rc) rc = dalloc(w, &D.le, R * dim);
  if (!rc) rc = dalloc(w, &D.budget, R);
  if (!rc) rc = dalloc(w, &D.nsamp, R);
  if (!rc) rc = dalloc(w, &D.pilot_col, R);
  if (!rc) rc = dalloc(w, &D.active, R);
  if (!rc) rc = dalloc(w, &D.part, R * U * dim);
  if (!rc) rc = dalloc(w, &D.scol, R * U * dim);
  if (!rc) rc = dalloc(w, &D.proj, R * U * dim);
  if (!rc) rc = dalloc(w, &D.dist, R * U);
  if (!rc) rc = dalloc(w, &D.u, R);
  if (!rc) rc = dalloc(w, &D.status, R);
  if (!rc) rc = dalloc(w, &D.faces_added, R);
  if (!rc) rc = dalloc(w, &D.iters, R);
  if (!rc) rc = dalloc(w, &D.samples, R);
  if (!rc) rc = dalloc(w, &D.collisions, R);
  if (!rc) rc = dalloc(w, &w->vox, std::max(max_voxels, 1) * 3);
  if (!rc) rc = dalloc(w, &w->ctl, 8);
  if (!rc) rc = dalloc(w, &w->ctl2, 8);
  if (!rc) rc = dalloc(w, &w->active2, R);
  if (!rc) rc = dalloc(w, &w->mtab, std::max<uint32_t>(options->max_iterations, 1));
  if (!rc) rc = dalloc(w, &w->ttab, std::max<uint32_t>(options->max_iterations, 1));
  D.T = (int)std::max<uint32_t>(options->max_iterations, 1);
  if (!rc) rc = dalloc(w, &D.stats, R * (size_t)D.T * kStats);
  if (rc) {
    csd_eizo_destroy(w);
    return rc;
  }
  *out = w;
  return CSD_OK;
}

int csd_eizo_destroy(csd_eizo *w) {
  if (!w) return CSD_OK;
  hipDeviceSynchronize();
  for (void *p : w->owned) hipFree(p);
  if (w->D.stage) hipFree(w->D.stage);
  if (w->D.blk_cnt) hipFree(w->D.blk_cnt);
  if (w->D.blk_first) hipFree(w->D.blk_first);
  if (w->pred) hipFree(w->pred);
  if (w->ppred) hipFree(w->ppred);
  if (w->gs) hipStreamDestroy(w->gs);
  mlp_attach(w->surrogate, -1);
  delete w;
  return CSD_OK;
}

int csd_eizo_set_guidance(csd_eizo *w, uint32_t pilot_samples, float pilot_factor,
                          const csd_mlp *surrogate, int32_t keep_particles, int32_t line_points) {
  if (!w) return fail(CSD_ERR_INVALID_ARGUMENT, "null workspace");
  if (pilot_samples > 262144) return fail(CSD_ERR_CAPACITY, "at most 262144 pilot samples");
  if (pilot_samples && !(pilot_factor > 0.f))
    return fail(CSD_ERR_INVALID_ARGUMENT, "pilot_factor must be positive");
  if (surrogate && (keep_particles < 1 || keep_particles > kRankMax || line_points < 1 ||
                    line_points > 64))
    return fail(CSD_ERR_INVALID_ARGUMENT, "keep_particles must be 1..1024 and line_points 1..64");
  w->pilot = pilot_samples;
  w->pilot_factor = pilot_factor;
  if (surrogate != w->surrogate) {
    mlp_attach(surrogate, +1);
    mlp_attach(w->surrogate, -1);
  }
  w->surrogate = surrogate;
  if (surrogate) {
    w->keep = keep_particles;
    w->line_points = line_points;
  }
  return CSD_OK;
}

int csd_eizo_set_pruning(csd_eizo *w, int32_t enable, float predicted_below) {
  if (!w) return fail(CSD_ERR_INVALID_ARGUMENT, "null workspace");
  if (enable && !std::isfinite(predicted_below))
    return fail(CSD_ERR_INVALID_ARGUMENT, "the pruning threshold must be finite");
  w->prune = enable != 0;
  w->prune_below = predicted_below;
  return CSD_OK;
}

int csd_eizo_set_sampler_tile(csd_eizo *w, int32_t wide_from) {
  if (!w) return fail(CSD_ERR_INVALID_ARGUMENT, "null workspace");
  if (wide_from < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative face count");
  w->wide_from = wide_from;
  return CSD_OK;
}

int csd_eizo_tune_sampler(csd_eizo *w, int32_t *wide_from) {
  if (!w) return fail(CSD_ERR_INVALID_ARGUMENT, "null workspace");
  const int dim = w->dim, cap = w->cap;
  EizoDev &D = w->D;
  // synthetic regions that fill the GPU: Rt slots of N samples, short chains
  // (the time per step and face is what the tiles differ in)
  const int Rt = std::min(w->R, 64);
  const uint32_t N = 16384;
  const int nblk = (int)(N / kSampleBlock), steps = 8;
  int rc = grow_stage(w, Rt, nblk);
  if (rc) return rc;
  // face-count classes where both tiles fit: up to the largest polytope a
  // 256-thread block stages
  const int dp = dim <= 8 ? 8 : 16;
  const int narrow_max = (int)(kLdsLimit / ((dp + 1) * sizeof(float)));
  std::vector<int> classes;
  for (int f : {2 * dim + 32, 256, 512, narrow_max})
    if (f <= std::min(cap, narrow_max) && (classes.empty() || f > classes.back()))
      classes.push_back(f);
  // polytope: the box |x_i| <= 1 and unit normals n with b = 1 + u, u in
  // [0, 1): the unit ball is inside; segment along x_0 through the origin
  const int fmax = classes.empty() ? 0 : classes.back();
  std::vector<float> A((size_t)Rt * cap * dim, 0.f), b((size_t)Rt * cap, 0.f), ls(Rt * dim, 0.f),
      le(Rt * dim, 0.f);
  uint64_t seed = 0x5eed;
  auto uni = [&seed] {
    seed = seed * 6364136223846793005ull + 1442695040888963407ull;
    return (float)((seed >> 40) & 0xffffff) / 16777216.f;
  };
  for (int r = 0; r < Rt; ++r) {
    float *Ar = &A[(size_t)r * cap * dim], *br = &b[(size_t)r * cap];
    for (int f = 0; f < fmax; ++f) {
      float *a = Ar + (size_t)f * dim;
      if (f < 2 * dim) {
        a[f / 2] = (f & 1) ? 1.f : -1.f;
        br[f] = 1.f;
        continue;
      }
      float nn = 0.f;
      for (int d = 0; d < dim; ++d) {
        a[d] = uni() - 0.5f;
        nn += a[d] * a[d];
      }
      nn = nn > 0.f ? 1.f / sqrtf(nn) : 0.f;
      for (int d = 0; d < dim; ++d) a[d] *= nn;
      br[f] = 1.f + uni();
    }
    ls[r * dim] = -0.1f;
    le[r * dim] = 0.1f;
  }
  std::vector<int32_t> act(Rt);
  std::vector<uint32_t> ns(w->R, N);
  for (int r = 0; r < Rt; ++r) act[r] = r;
  CSD_HIP(hipMemcpy(D.A, A.data(), A.size() * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.b, b.data(), b.size() * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.ls, ls.data(), ls.size() * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.le, le.data(), le.size() * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.active, act.data(), Rt * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.nsamp, ns.data(), (size_t)w->R * 4, hipMemcpyHostToDevice));
  EizoDev E = D;
  E.ctl = nullptr;
  hipEvent_t e0, e1;
  CSD_HIP(hipEventCreate(&e0));
  CSD_HIP(hipEventCreate(&e1));
  int from = std::numeric_limits<int>::max();
  bool wide_wins_above = true;  // 1024 wins at every larger class tested so far
  for (int k = (int)classes.size() - 1; k >= 0 && rc == CSD_OK; --k) {
    const int F = classes[k];
    std::vector<int32_t> nf(Rt, F);
    if (hipMemcpy(D.nf, nf.data(), Rt * 4, hipMemcpyHostToDevice) != hipSuccess) {
      rc = hip_fail(hipGetLastError(), "hipMemcpy (sampler tuning)");
      break;
    }
    float best[2] = {0.f, 0.f};
    for (int t = 0; t < 2 && rc == CSD_OK; ++t) {
      const SamplerLaunch L = pick_sampler(dim, F, t == 0 ? std::numeric_limits<int>::max() : 0);
      float ms_min = 1e30f;
      for (int rep = 0; rep < 4; ++rep) {  // the first launch warms up
        hipEventRecord(e0, 0);
        launch_sampler(L, nblk, Rt, 0, E, dim, cap, steps, 0x1234u, 0x5678u, 0u, 0u);
        hipEventRecord(e1, 0);
        hipError_t he = hipEventSynchronize(e1);
        if (he == hipSuccess) he = hipGetLastError();
        if (he != hipSuccess) {
          rc = hip_fail(he, "sampler tuning launch");
          break;
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0) ms_min = std::min(ms_min, ms);
      }
      best[t] = ms_min;
    }
    const bool wide = best[1] < 0.97f * best[0];
    wide_wins_above = wide_wins_above && wide;
    if (wide_wins_above) from = F;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  w->wide_from = from;
  if (wide_from) *wide_from = from;
  return CSD_OK;
}

int csd_eizo_set_metric(csd_eizo *w, const float *metric) {
  if (!w) return fail(CSD_ERR_INVALID_ARGUMENT, "null workspace");
  const int n = w->dim;
  if (!metric) {
    w->D.metric = nullptr;  // Euclidean: the reference's placement
    return CSD_OK;
  }
  // symmetric positive definite: a Cholesky factorisation in FP64 succeeds
  std::vector<double> M((size_t)n * n), L((size_t)n * n, 0.0);
  for (int k = 0; k < n * n; ++k) M[k] = metric[k];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      const double a = M[i * n + j], b = M[j * n + i];
      if (!(fabs(a - b) <= 1e-6 * std::max(1.0, std::max(fabs(a), fabs(b)))))
        return fail(CSD_ERR_INVALID_ARGUMENT, "metric is not symmetric");
    }
  for (int j = 0; j < n; ++j) {
    double d = M[j * n + j];
    for (int k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0.0)) return fail(CSD_ERR_INVALID_ARGUMENT, "metric is not positive definite");
    L[j * n + j] = sqrt(d);
    for (int i = j + 1; i < n; ++i) {
      double v = M[i * n + j];
      for (int k = 0; k < j; ++k) v -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = v / L[j * n + j];
    }
  }
  if (!w->metric_buf) {
    const int rc = dalloc(w, &w->metric_buf, (size_t)kDimMax * kDimMax);
    if (rc) return rc;
  }
  CSD_HIP(hipMemcpy(w->metric_buf, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
  w->D.metric = w->metric_buf;
  return CSD_OK;
}

int csd_eizo_set_graph(csd_eizo *w, int32_t rounds_per_sync) {
  if (!w) return fail(CSD_ERR_INVALID_ARGUMENT, "null workspace");
  if (rounds_per_sync < 0 || rounds_per_sync > 64)
    return fail(CSD_ERR_INVALID_ARGUMENT, "rounds_per_sync must be 0..64");
  if (rounds_per_sync && !w->gs) CSD_HIP(hipStreamCreateWithFlags(&w->gs, hipStreamNonBlocking));
  w->graph_sync = rounds_per_sync;
  return CSD_OK;
}

int csd_eizo_run(csd_eizo *w, const float *domain_A, const float *domain_b, int32_t domain_faces,
                 const float *line_start, const float *line_end, int32_t num_regions,
                 const uint32_t *particle_budgets, const float *voxels, int32_t num_voxels,
                 float voxel_radius, float *out_A, float *out_b,
                 csd_eizo_region_info *infos) {
  if (!w) return fail(CSD_ERR_INVALID_ARGUMENT, "null workspace");
  if (num_regions < 0 || num_regions > w->R)
    return fail(CSD_ERR_CAPACITY, "more regions than the workspace holds");
  if (num_regions == 0) return CSD_OK;
  if (domain_faces < 0 || domain_faces >= w->cap)
    return fail(CSD_ERR_CAPACITY, "domain does not leave room for new faces");
  if (num_voxels < 0 || num_voxels > w->Vmax)
    return fail(CSD_ERR_CAPACITY, "more voxels than the workspace holds");
  if (!line_start || !line_end || !out_A || !out_b || !infos || (domain_faces && (!domain_A || !domain_b)) ||
      (num_voxels && !voxels))
    return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  const int dim = w->dim, cap = w->cap, U = w->U, Rn = num_regions;
  EizoDev &D = w->D;
  const csd_eizo_options &o = w->opt;

  // ---- region state ----
  std::vector<float> A((size_t)Rn * cap * dim, 0.f), b((size_t)Rn * cap, 0.f);
  for (int r = 0; r < Rn; ++r) {
    memcpy(&A[(size_t)r * cap * dim], domain_A, sizeof(float) * domain_faces * dim);
    memcpy(&b[(size_t)r * cap], domain_b, sizeof(float) * domain_faces);
  }
  std::vector<int32_t> nf(Rn, domain_faces), zeros(Rn, 0), active(Rn);
  std::vector<int64_t> zeros64(Rn, 0);
  std::vector<uint32_t> budget(Rn);
  uint32_t max_budget = 0;
  for (int r = 0; r < Rn; ++r) {
    budget[r] = particle_budgets ? particle_budgets[r] : o.num_particles;
    if (budget[r] == 0 || (int)budget[r] > U) return fail(CSD_ERR_CAPACITY, "particle budget 0 or above the workspace capacity");
    max_budget = std::max(max_budget, budget[r]);
    active[r] = r;
  }
  CSD_HIP(hipMemcpy(D.A, A.data(), A.size() * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.b, b.data(), b.size() * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.nf, nf.data(), Rn * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.ls, line_start, (size_t)Rn * dim * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.le, line_end, (size_t)Rn * dim * 4, hipMemcpyHostToDevice));
  CSD_HIP(hipMemcpy(D.budget, budget.data(), Rn * 4, hipMemcpyHostToDevice));
  for (int32_t *p : {D.u, D.status, D.faces_added, D.iters})
    CSD_HIP(hipMemcpy(p, zeros.data(), Rn * 4, hipMemcpyHostToDevice));
  for (int64_t *p : {D.samples, D.collisions})
    CSD_HIP(hipMemcpy(p, zeros64.data(), Rn * 8, hipMemcpyHostToDevice));
  if (num_voxels)
    CSD_HIP(hipMemcpy(w->vox, voxels, (size_t)num_voxels * 12, hipMemcpyHostToDevice));
  CSD_HIP(hipMemset(D.stats, 0, (size_t)Rn * D.T * kStats * sizeof(int32_t)));
  w->last_regions = 0;
  w->round_ms.assign((size_t)D.T, 0.f);
  using clock = std::chrono::steady_clock;

  RobotGeomList RL;
  VoxelArgs V{w->vox, num_voxels, voxel_radius, 0, 0};
  int rc = CSD_OK;
  bool all_slots = fill_robot_list(w->plant, w->robot.data(), (int)w->robot.size(), RL, V, &rc);
  if (rc) return rc;
  const PlantImage &P = w->plant->img;
  const int var = check_variant(P, num_voxels > 0 && !all_slots);
  const size_t tab_bytes = check_lds_bytes(P, var);
  const uint32_t k0 = (uint32_t)o.seed, k1 = (uint32_t)(o.seed >> 32);
  std::vector<int32_t> status(Rn), nfh(nf);  // nfh: face counts on the host

  std::vector<int32_t> rounds(Rn, 0), pcol(Rn, 0);
  std::vector<uint32_t> nsamp(Rn, 0);
  const int dp = dim <= 8 ? 8 : 16;

  // one sampling phase over the regions in `slots`: nsamp[r] samples each
  // (stream offset: independent random numbers for the pilot), collision
  // check, compaction (+ the unadaptive test when `test`)
  // prune (pilot only): the surrogate predicts every sample and only those
  // predicted below prune_below get the exact check
  auto phase = [&](const std::vector<int32_t> &slots, uint32_t it, uint32_t M, bool test,
                   uint32_t stream_off, bool prune) -> int {
    const int na = (int)slots.size();
    uint32_t nmax = 0;
    for (int r : slots) nmax = std::max(nmax, nsamp[r]);
    const int nblk = (int)((nmax + kSampleBlock - 1) / kSampleBlock);
    if (nblk > 1024) return fail(CSD_ERR_CAPACITY, "more than 262144 samples in one round");
    const int grc = grow_stage(w, na, nblk);
    if (grc) return grc;
    CSD_HIP(hipMemcpy(D.active, slots.data(), na * 4, hipMemcpyHostToDevice));
    CSD_HIP(hipMemcpy(D.nsamp, nsamp.data(), Rn * 4, hipMemcpyHostToDevice));
    const float thresh = (1 - o.tau) * o.epsilon * M;
    // the largest polytope of the phase decides whether they are staged in LDS
    int nfmax = 0;
    for (int r : slots) nfmax = std::max(nfmax, nfh[r]);
    const dim3 sg(nblk, na), sb(kSampleBlock);
    launch_sampler(pick_sampler(dim, nfmax, w->wide_from), nblk, na, 0, D, dim, cap,
                   (int)o.mixing_steps, k0, k1, it, stream_off);
    const float *pr = nullptr;
    if (prune) {
      const size_t need = (size_t)na * nblk * kSampleBlock * sizeof(float);
      if (need > w->ppred_bytes) {
        if (w->ppred) hipFree(w->ppred);
        w->ppred = nullptr;
        w->ppred_bytes = 0;
        CSD_HIP(hipMalloc(&w->ppred, need));
        w->ppred_bytes = need;
      }
      const int prc = mlp_predict_staged(w->surrogate, D.stage, na * nblk, dim, w->ppred, 0);
      if (prc) return prc;
      pr = w->ppred;
    }
    if (dp == 8)
      CSD_DISPATCH_VARIANT_X(var, k_eizo_check, 8, sg, sb, tab_bytes, 0, P, V, RL, D, dim, nblk,
                             M, pr, w->prune_below);
    else
      CSD_DISPATCH_VARIANT_X(var, k_eizo_check, 16, sg, sb, tab_bytes, 0, P, V, RL, D, dim,
                             nblk, M, pr, w->prune_below);
    hipLaunchKernelGGL(k_eizo_compact, dim3(na), dim3(256), 0, 0, D, dim, nblk, M, thresh, U,
                       test ? 1 : prune ? 2 : 0, it);
    CSD_HIP(hipGetLastError());
    return CSD_OK;
  };

  if (w->graph_sync) {
    // Device-resident rounds: one round is captured as a graph and replayed;
    // the host reads the region status every graph_sync rounds only. A round
    // is the host loop's, stage for stage: start (slots, sample counts), the
    // pilot (sampling, surrogate predictions when pruned, check, count) and
    // its decision, the test phase (sampling, check, unadaptive test and
    // gather), the surrogate ranking, bisection and faces. Every stage
    // launches with every region as a slot and idles past the slot count its
    // control block holds.
    const uint32_t T = o.max_iterations;
    std::vector<uint32_t> mt(std::max<uint32_t>(T, 1), 0u);
    std::vector<float> tt(mt.size(), 0.f);
    for (uint32_t it = 0; it < T; ++it) {
      mt[it] = unadaptive_samples(o, it);
      tt[it] = (1 - o.tau) * o.epsilon * mt[it];
    }
    const uint32_t zero8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    CSD_HIP(hipMemcpy(w->ctl, zero8, sizeof(zero8), hipMemcpyHostToDevice));
    CSD_HIP(hipMemcpy(w->ctl2, zero8, sizeof(zero8), hipMemcpyHostToDevice));
    CSD_HIP(hipMemcpy(w->mtab, mt.data(), mt.size() * 4, hipMemcpyHostToDevice));
    CSD_HIP(hipMemcpy(w->ttab, tt.data(), tt.size() * 4, hipMemcpyHostToDevice));
    const hipStream_t s = w->gs;
    const bool ranked = w->surrogate != nullptr;
    const bool pruned = w->pilot && w->prune && w->surrogate;
    const uint32_t bmax = ranked ? (uint32_t)w->keep : max_budget;
    const int pblk = (int)((w->pilot + kSampleBlock - 1) / kSampleBlock);
    const float limit = w->pilot_factor * o.epsilon * (float)w->pilot;
    const int L = w->line_points;
    // the sampler stages the largest polytope the workspace holds
    const SamplerLaunch sl = pick_sampler(dim, cap, w->wide_from);
    // buffers for every region as a slot, sized outside the capture
    if (ranked) {
      const size_t need = (size_t)Rn * kRankMax * L * sizeof(float);
      if (need > w->pred_bytes) {
        if (w->pred) hipFree(w->pred);
        w->pred = nullptr;
        w->pred_bytes = 0;
        CSD_HIP(hipMalloc(&w->pred, need));
        w->pred_bytes = need;
      }
    }
    if (pruned) {
      const size_t need = (size_t)Rn * pblk * kSampleBlock * sizeof(float);
      if (need > w->ppred_bytes) {
        if (w->ppred) hipFree(w->ppred);
        w->ppred = nullptr;
        w->ppred_bytes = 0;
        CSD_HIP(hipMalloc(&w->ppred, need));
        w->ppred_bytes = need;
      }
    }
    auto sample = [&](const EizoDev &E, int nb, uint32_t off) {
      launch_sampler(sl, nb, Rn, s, E, dim, cap, (int)o.mixing_steps, k0, k1, 0u, off);
    };
    auto check = [&](const EizoDev &E, int nb, const float *pr, float below) {
      const dim3 sg(nb, Rn), sb(kSampleBlock);
      if (dp == 8)
        CSD_DISPATCH_VARIANT_X(var, k_eizo_check, 8, sg, sb, tab_bytes, s, P, V, RL, E, dim, nb,
                               0u, pr, below);
      else
        CSD_DISPATCH_VARIANT_X(var, k_eizo_check, 16, sg, sb, tab_bytes, s, P, V, RL, E, dim,
                               nb, 0u, pr, below);
    };
    bool done = false;
    uint32_t it = 0;
    while (it < T && !done) {
      // rounds [it, end) share one graph, sized for the largest M_k among them
      const uint32_t end = std::min<uint32_t>(T, std::max<uint32_t>(16u, it * 4u));
      const int nblk = (int)((std::max(mt[end - 1], max_budget) + kSampleBlock - 1) / kSampleBlock);
      if (nblk > 1024) return fail(CSD_ERR_CAPACITY, "more than 262144 samples in one round");
      rc = grow_stage(w, Rn, std::max(nblk, pblk));
      if (rc) return rc;
      if (ranked) {  // the ranking reorders through the staging buffer
        const size_t need = (size_t)Rn * kRankMax * dim * sizeof(float);
        if (need > w->stage_bytes) {
          if (D.stage) hipFree(D.stage);
          D.stage = nullptr;
          w->stage_bytes = 0;
          CSD_HIP(hipMalloc(&D.stage, need));
          w->stage_bytes = need;
        }
      }
      EizoDev G = D;
      G.ctl = w->ctl;
      G.mtab = w->mtab;
      G.ttab = w->ttab;
      EizoDev G2 = G;  // the test phase after a pilot: its own slots and control block
      G2.active = w->active2;
      G2.ctl = w->ctl2;
      const EizoDev &GT = w->pilot ? G2 : G;
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      CSD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int crc = CSD_OK;
      hipLaunchKernelGGL(k_eizo_round_begin, dim3(1), dim3(kBeginBlock), 0, s, G, Rn, T,
                         w->pilot);
      if (w->pilot) {
        sample(G, pblk, 0x40000000u);
        const float *pr = nullptr;
        if (pruned) {
          crc = mlp_predict_staged(w->surrogate, G.stage, Rn * pblk, dim, w->ppred, s, w->ctl + 1,
                                   pblk);
          pr = w->ppred;
        }
        check(G, pblk, pr, w->prune_below);
        hipLaunchKernelGGL(k_eizo_compact, dim3(Rn), dim3(256), 0, s, G, dim, pblk, 0u, 0.f, U,
                           pruned ? 2 : 0, 0u);
        hipLaunchKernelGGL(k_eizo_pilot_select, dim3(1), dim3(kBeginBlock), 0, s, G, G2, limit);
      }
      sample(GT, nblk, 0u);
      check(GT, nblk, nullptr, 0.f);
      hipLaunchKernelGGL(k_eizo_compact, dim3(Rn), dim3(256), 0, s, GT, dim, nblk, 0u, 0.f, U, 1,
                         0u);
      if (ranked && crc == CSD_OK) {
        SegmentPoints sp{G.part, G.ls, G.le, G.active, G.u, U, dim, L, kRankMax, w->ctl + 1};
        crc = mlp_predict_segment_points(w->surrogate, sp, Rn, w->pred, s);
        hipLaunchKernelGGL(k_eizo_rank, dim3(Rn), dim3(kRankBlock), 0, s, G, w->pred, dim, U, L,
                           kRankMax, w->keep, G.stage, 0u);
      }
      CSD_EIZO_BISECT(Rn, bmax, s, G);
      hipLaunchKernelGGL(k_eizo_faces, dim3(Rn), dim3(kFaceBlock), 0, s, G, dim, U, cap,
                         o.configuration_margin, (int)o.max_hyperplanes_per_iteration, 1, 0u);
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (crc != CSD_OK) {
        if (graph) hipGraphDestroy(graph);
        return crc;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (graph) hipGraphDestroy(graph);
      while (e == hipSuccess && it < end && !done) {
        const uint32_t n = std::min<uint32_t>((uint32_t)w->graph_sync, end - it);
        const auto t0 = clock::now();
        for (uint32_t k = 0; k < n && e == hipSuccess; ++k) e = hipGraphLaunch(exec, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(status.data(), D.status, Rn * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        // the replays between two status reads share their wall time
        const float ms = std::chrono::duration<float, std::milli>(clock::now() - t0).count();
        for (uint32_t k = it; k < it + n && k < (uint32_t)D.T; ++k) w->round_ms[k] = ms / n;
        it += n;
        done = std::none_of(status.begin(), status.end(), [](int32_t v) { return v == 0; });
      }
      if (exec) hipGraphExecDestroy(exec);
      if (e != hipSuccess) return hip_fail(e, "EI-ZO round graph");
    }
    CSD_HIP(hipMemcpy(rounds.data(), D.iters, Rn * 4, hipMemcpyDeviceToHost));
    active.clear();  // the host-driven loop below has nothing left to do
  }

  for (uint32_t it = 0; it < o.max_iterations && !active.empty(); ++it) {
    const auto t_round = clock::now();
    const int na = (int)active.size();
    for (int r : active) rounds[r] += 1;
    const uint32_t M = unadaptive_samples(o, it);
    // pilot: regions whose pilot collision fraction is far above epsilon skip
    // this round's test and take their particles from the pilot
    std::vector<int32_t> full = active;
    if (w->pilot) {
      for (int r : active) nsamp[r] = w->pilot;
      rc = phase(active, it, M, false, 0x40000000u, w->prune && w->surrogate);
      if (rc) return rc;
      CSD_HIP(hipMemcpy(pcol.data(), D.pilot_col, Rn * 4, hipMemcpyDeviceToHost));
      full.clear();
      const float limit = w->pilot_factor * o.epsilon * (float)w->pilot;
      for (int r : active)
        if ((float)pcol[r] <= limit) full.push_back(r);
    }
    if (!full.empty()) {
      for (int r : full) nsamp[r] = std::max(M, budget[r]);
      rc = phase(full, it, M, true, 0u, false);
      if (rc) return rc;
    }
    if (w->pilot || full.size() != active.size())
      CSD_HIP(hipMemcpy(D.active, active.data(), na * 4, hipMemcpyHostToDevice));
    if (w->surrogate) {
      const int L = w->line_points, cand = kRankMax;
      const size_t need = (size_t)na * cand * L * sizeof(float);
      if (need > w->pred_bytes) {
        if (w->pred) hipFree(w->pred);
        w->pred = nullptr;
        w->pred_bytes = 0;
        CSD_HIP(hipMalloc(&w->pred, need));
        w->pred_bytes = need;
      }
      // the ranking reorders through the staging buffer (dead at this point)
      const size_t rank_need = (size_t)na * kRankMax * dim * sizeof(float);
      if (rank_need > w->stage_bytes) {
        if (D.stage) hipFree(D.stage);
        D.stage = nullptr;
        w->stage_bytes = 0;
        CSD_HIP(hipMalloc(&D.stage, rank_need));
        w->stage_bytes = rank_need;
      }
      SegmentPoints sp{D.part, D.ls, D.le, D.active, D.u, U, dim, L, cand};
      rc = mlp_predict_segment_points(w->surrogate, sp, na, w->pred, 0);
      if (rc) return rc;
      hipLaunchKernelGGL(k_eizo_rank, dim3(na), dim3(kRankBlock), 0, 0, D, w->pred, dim, U, L,
                         cand, w->keep, D.stage, it);
    }
    const uint32_t bmax = w->surrogate ? (uint32_t)w->keep : max_budget;
    CSD_EIZO_BISECT(na, bmax, 0, D);
    hipLaunchKernelGGL(k_eizo_faces, dim3(na), dim3(kFaceBlock), 0, 0, D, dim, U, cap,
                       o.configuration_margin, (int)o.max_hyperplanes_per_iteration, 1, it);
    CSD_HIP(hipGetLastError());
    CSD_HIP(hipMemcpy(status.data(), D.status, Rn * 4, hipMemcpyDeviceToHost));
    CSD_HIP(hipMemcpy(nfh.data(), D.nf, Rn * 4, hipMemcpyDeviceToHost));
    std::vector<int32_t> next;
    for (int r : active)
      if (status[r] == 0) next.push_back(r);  // 1 certified, 2 face capacity full
    active.swap(next);
    w->round_ms[it] = std::chrono::duration<float, std::milli>(clock::now() - t_round).count();
  }

  CSD_HIP(hipMemcpy(out_A, D.A, (size_t)Rn * cap * dim * 4, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(out_b, D.b, (size_t)Rn * cap * 4, hipMemcpyDeviceToHost));
  std::vector<int32_t> fa(Rn);
  std::vector<int64_t> smp(Rn), colls(Rn);
  CSD_HIP(hipMemcpy(nf.data(), D.nf, Rn * 4, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(fa.data(), D.faces_added, Rn * 4, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(smp.data(), D.samples, Rn * 8, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(colls.data(), D.collisions, Rn * 8, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(status.data(), D.status, Rn * 4, hipMemcpyDeviceToHost));
  w->last_regions = Rn;
  w->last_rounds = rounds;
  for (int r = 0; r < Rn; ++r) {
    infos[r].iterations = rounds[r];
    infos[r].converged = status[r] == 1;
    infos[r].capacity_reached = status[r] == 2;
    infos[r].pad = 0;
    infos[r].num_faces = nf[r];
    infos[r].faces_added = fa[r];
    infos[r].samples_drawn = smp[r];
    infos[r].collisions_found = colls[r];
  }
  return CSD_OK;
}

int csd_eizo_get_rounds(const csd_eizo *w, int32_t region, csd_eizo_round_info *out,
                        int32_t capacity, int32_t *count) {
  if (!w || !count) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  if (region < 0 || region >= w->last_regions)
    return fail(CSD_ERR_INVALID_ARGUMENT, "region outside the last csd_eizo_run");
  const int n = std::min(w->last_rounds[region], w->D.T);
  *count = n;
  if (n == 0 || !out) return CSD_OK;
  if (capacity < n) return fail(CSD_ERR_CAPACITY, "fewer entries than rounds");
  std::vector<int32_t> st((size_t)n * kStats);
  CSD_HIP(hipMemcpy(st.data(), w->D.stats + (size_t)region * w->D.T * kStats,
                    st.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int k = 0; k < n; ++k) {
    const int32_t *e = &st[(size_t)k * kStats];
    csd_eizo_round_info &o = out[k];
    o.iteration_index = k;
    o.num_samples_drawn = e[kStatSamples];
    o.collisions_threshold = (1 - w->opt.tau) * w->opt.epsilon * unadaptive_samples(w->opt, k);
    o.num_collisions_unadaptive = e[kStatFirstCol];
    o.num_collisions_found = e[kStatCol];
    o.num_collisions_updated = e[kStatUpdated];
    o.num_faces_placed = e[kStatFaces];
    o.num_faces_at_start = e[kStatFacesStart];
    o.pilot_samples = e[kStatPilot];
    o.pilot_collisions = e[kStatPilotCol];
    o.pilot_checked = e[kStatPilotChecked];
    o.iteration_time_ms = k < (int)w->round_ms.size() ? w->round_ms[k] : 0.f;
  }
  return CSD_OK;
}

int csd_eizo_last_particles(const csd_eizo *w, int32_t region, float *particles, float *scol,
                            float *proj, int32_t capacity, int32_t *count) {
  if (!w || !count) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  if (region < 0 || region >= w->last_regions)
    return fail(CSD_ERR_INVALID_ARGUMENT, "region outside the last csd_eizo_run");
  int32_t u = 0;
  CSD_HIP(hipMemcpy(&u, w->D.u + region, 4, hipMemcpyDeviceToHost));
  *count = u;
  if (u == 0 || (!particles && !scol && !proj)) return CSD_OK;  // count only
  if (capacity < u) return fail(CSD_ERR_CAPACITY, "fewer rows than particles");
  const size_t off = (size_t)region * w->U * w->dim, bytes = (size_t)u * w->dim * sizeof(float);
  if (particles) CSD_HIP(hipMemcpy(particles, w->D.part + off, bytes, hipMemcpyDeviceToHost));
  if (scol) CSD_HIP(hipMemcpy(scol, w->D.scol + off, bytes, hipMemcpyDeviceToHost));
  if (proj) CSD_HIP(hipMemcpy(proj, w->D.proj + off, bytes, hipMemcpyDeviceToHost));
  return CSD_OK;
}

}  // extern "C"
