// csd_screen.hip -- screened collision checking for fast-path plants.
//
// In cluttered scenes most configurations collide, and most of those hit one
// of a few pairs (in a 7-DOF arm among 16 boxes, 40 of 784 pair tests decide
// 96% of uniform configurations). The fused checker still tests every pair of
// a wave whenever one of its 64 lanes is free. The screened checker runs:
//   1. k_screen: one lane per configuration: FK and the screened pair tests
//      only; writes 0 (collides) or 2 (undecided) to the output;
//   2. k_finish: each workgroup takes a range of configurations, compacts the
//      undecided ones in LDS and gives every batch of 64 to all of its waves,
//      each wave testing its share of the plant's full pair list; writes
//      0 or 1.
// The screen is a subset of the full check's own pair tests with the same
// contact distances and arithmetic, so the result equals the unscreened
// checker's bit for bit. The screen is chosen on the host from the hit bits
// of a sample of the workload (greedy set cover, csd_tune_check).
#include <string.h>

#include <algorithm>
#include <vector>

#include "csd_collide.h"
#include "csd_screen.h"

namespace csd {

constexpr int kScreenBlock = 256;
constexpr int kFinishBlock = 512;    // 8 waves, 2 per SIMD (<= 256 VGPRs)
constexpr int kFinishWaves = kFinishBlock / 64;
constexpr int kMaxRange = 16384;     // configurations per survivor workgroup

template <int VAR>
struct FastShape {
  static constexpr int NB = (VAR == kFast8x4 || VAR == kFast8x8) ? 8 : 16;
  static constexpr int KM = (VAR == kFast8x4 || VAR == kFast16x4) ? 4 : 8;
};

__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// byte offsets in the screen program (see ScreenImage)
template <int NB, int KM>
struct ScreenLayout {
  static constexpr int NS = NB * KM;
  static constexpr int kActive = 0;
  static constexpr int kHot = 8 * NB;
  static constexpr int kSelfR2 = kHot + 8 * NB * NB;
  static constexpr int kObst = kSelfR2 + 4 * NB * NB * KM * KM;
  static constexpr int kRec = 64 + 4 * NS + 16;  // one obstacle record
  static constexpr int kRecR2 = 64, kRecMask = 64 + 4 * NS, kRecType = kRecMask + 8;
};

// The screened pair tests of one lane's configuration, reading the screen
// program from LDS. Tests run in groups of the KM = 4 slots of one body (a
// self row: sphere ka of body ba against the 4 spheres of body bb; an
// obstacle against the 4 spheres of body b), branch-free inside a group
// with one 16-byte read of contact distances (-1 = not screened, never a
// hit). Group masks are made wave-uniform (readfirstlane), so a group that
// is not screened costs a scalar branch and no vector work.
template <int NB, int KM, class Q>
__device__ __forceinline__ bool screen_fast(const PlantImage &P, const char *prog, int num_obst,
                                            const Q &qa, bool col) {
  static_assert(KM == 4, "screen groups are the 4 sphere slots of a body");
  using L = ScreenLayout<NB, KM>;
  constexpr int NS = NB * KM;
  float cx[NS], cy[NS], cz[NS], rad[NS];
  fast_spheres<NB, KM, false>(P, qa, cx, cy, cz, rad);
  __syncthreads();  // the program is in LDS
  col = col || P.always_colliding;
  const uint64_t *act = (const uint64_t *)(prog + L::kActive);
  const uint64_t *hot = (const uint64_t *)(prog + L::kHot);
  const float4 *sr2 = (const float4 *)(prog + L::kSelfR2);
#pragma unroll
  for (int ba = 0; ba < NB; ++ba) {
    const uint64_t a = uniform_u64(act[ba]);
    if (a == 0) continue;
#pragma unroll
    for (int bb = ba + 1; bb < NB; ++bb) {
      if (!((a >> bb) & 1)) continue;
      const uint64_t hm = uniform_u64(hot[ba * NB + bb]);
#pragma unroll
      for (int ka = 0; ka < KM; ++ka) {
        if (!((hm >> ka) & 1)) continue;
        const float4 r = sr2[(ba * NB + bb) * KM + ka];
        const int sa = ba * KM + ka, sb = bb * KM;
        col |= hit_sphere_sphere(cx[sa], cy[sa], cz[sa], cx[sb], cy[sb], cz[sb], r.x);
        col |= hit_sphere_sphere(cx[sa], cy[sa], cz[sa], cx[sb + 1], cy[sb + 1], cz[sb + 1], r.y);
        col |= hit_sphere_sphere(cx[sa], cy[sa], cz[sa], cx[sb + 2], cy[sb + 2], cz[sb + 2], r.z);
        col |= hit_sphere_sphere(cx[sa], cy[sa], cz[sa], cx[sb + 3], cy[sb + 3], cz[sb + 3], r.w);
      }
    }
  }
  for (int h = 0; h < num_obst; ++h) {
    if (__all(col)) break;
    const char *rec = prog + L::kObst + h * L::kRec;
    const uint64_t m = uniform_u64(*(const uint64_t *)(rec + L::kRecMask));
    const int t = __builtin_amdgcn_readfirstlane(*(const int32_t *)(rec + L::kRecType));
    const float4 *r2 = (const float4 *)(rec + L::kRecR2);
    float ov[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) ov[k] = ((const float *)rec)[k];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (!((m >> b) & 1)) continue;
      const float4 r4 = r2[b];
      const float rr[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        const int s = b * KM + k;
        if (t == CSD_BOX)
          col |= hit_sphere_obox(cx[s], cy[s], cz[s], rr[k], ov);
        else if (t == CSD_SPHERE)
          col |= hit_sphere_sphere(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], rr[k]);
        else
          col |= seg_point_d2(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], ov[3], ov[4], ov[5],
                              ov[6]) < rr[k];
      }
    }
  }
  return col;
}

template <int VAR>
__global__ __launch_bounds__(kScreenBlock) void k_screen(PlantImage P, ScreenImage S,
                                                         const float *__restrict__ q, int64_t n,
                                                         int64_t sn, int64_t sj,
                                                         uint8_t *__restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) uint4 sprog[];
  for (int w = threadIdx.x; w < S.prog_words; w += kScreenBlock) sprog[w] = S.prog[w];
  const int64_t gi = (int64_t)blockIdx.x * kScreenBlock + threadIdx.x;
  const bool valid = gi < n;
  const int64_t i = valid ? gi : n - 1;
  constexpr int NB = FastShape<VAR>::NB;
  const bool col = screen_fast<NB, FastShape<VAR>::KM>(
      P, (const char *)sprog, S.num_obst, preload_q<NB>(QMem{q, i * sn, sj}, P.dof), !valid);
  if (valid) out[gi] = col ? 0 : 2;
}

// Full check of the configurations k_screen left undecided (out == 2) in
// [blockIdx.x * range, + range): compacted in LDS, then taken in batches of
// 64 by the 8 waves of the workgroup -- 1 batch with 8 waves (parts of the
// pair tests) each, 2 with 4, 4 with 2 or 8 with 1, by the number of
// undecided configurations; the parts are OR-ed through LDS.
template <int VAR>
__global__ __launch_bounds__(kFinishBlock) void k_finish(PlantImage P,
                                                         const float *__restrict__ q, int64_t n,
                                                         int64_t sn, int64_t sj, int32_t range,
                                                         uint8_t *__restrict__ out) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  extern __shared__ __attribute__((aligned(16))) float csd_smem[];
  __shared__ int16_t list[kMaxRange];
  __shared__ uint8_t part[kFinishBlock];
  __shared__ int s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  const float *tab = stage_tables<VAR>(P, csd_smem);
  __syncthreads();  // s_cnt is zero (stage_tables has no barrier for 32 slots)
  const int64_t base = (int64_t)blockIdx.x * range;
  for (int e0 = 0; e0 < range; e0 += kFinishBlock) {
    const int e = e0 + threadIdx.x;
    const bool und = e < range && base + e < n && out[base + e] == 2;
    const uint64_t m = __ballot(und);
    int off = 0;
    if (lane == 0 && m) off = atomicAdd(&s_cnt, __popcll(m));
    off = __shfl(off, 0);
    if (und) list[off + __popcll(m & ((1ull << lane) - 1ull))] = (int16_t)e;
  }
  __syncthreads();
  const int total = s_cnt;
  RobotGeomList R;
  R.n = 0;
  const VoxelArgs V{nullptr, 0, 0.f, 0, 0};
  // nb batches of 64 at a time, each shared by 8 / nb waves (parts)
  const int nb = total <= 64 ? 1 : total <= 128 ? 2 : total <= 256 ? 4 : 8;
  const int parts = kFinishWaves / nb;
  const int my_batch = wave / parts, my_part = wave % parts;
  for (int g = 0; g < total; g += 64 * nb) {
    const int k = g + my_batch * 64 + lane;
    const bool ok = k < total;
    const int64_t j = base + list[ok ? k : total - 1];
    const bool c = collides_at<VAR>(P, QMem{q, j * sn, sj}, V, R, !ok, tab, my_part, parts);
    part[threadIdx.x] = c ? 1 : 0;
    __syncthreads();
    if (my_part == 0 && ok) {
      bool cc = false;
      for (int w = 0; w < parts; ++w) cc |= part[(wave + w) * 64 + lane] != 0;
      out[j] = cc ? 0 : 1;
    }
    __syncthreads();
  }
}

// Hit bits of every pair test of a fast-path plant for a sample of
// configurations: bit p of configuration i is word p / 32 at bits[(p / 32) *
// n + i]. Positions: self pair (ba, bb, ka, kb), ba < bb, at
// ((ba * NB + bb) * KM + ka) * KM + kb; obstacle pair (o, s) at
// NB * NB * KM * KM + o * NS + s. Tuning only: runtime loops (the sphere
// centres are indexed at run time and may live in scratch).
template <int VAR>
__global__ __launch_bounds__(kScreenBlock) void k_pair_hits(PlantImage P,
                                                            const float *__restrict__ q,
                                                            int64_t n, int64_t sn, int64_t sj,
                                                            uint32_t *__restrict__ bits) {
  constexpr int NB = FastShape<VAR>::NB, KM = FastShape<VAR>::KM, NS = NB * KM;
  constexpr bool kTab = NS <= 64;
  const int64_t i = (int64_t)blockIdx.x * kScreenBlock + threadIdx.x;
  if (i >= n) return;
  float cx[NS], cy[NS], cz[NS], rad[NS];
  fast_spheres<NB, KM, true>(P, QMem{q, i * sn, sj}, cx, cy, cz, rad);
  auto mark = [&](int p) { bits[(size_t)(p >> 5) * n + i] |= 1u << (p & 31); };
  for (int ba = 0; ba < NB; ++ba) {
    const uint64_t act = P.self_active[ba];
    for (int bb = ba + 1; bb < NB; ++bb) {
      if (!((act >> bb) & 1)) continue;
      const uint64_t pb = P.self_bits[ba * NB + bb];
      for (int ka = 0; ka < KM; ++ka)
        for (int kb = 0; kb < KM; ++kb) {
          if (!((pb >> (ka * KM + kb)) & 1)) continue;
          const int sa = ba * KM + ka, sb = bb * KM + kb;
          float r2;
          if (kTab) {
            r2 = P.self_r2[((size_t)ba * NB + bb) * (KM * KM) + ka * KM + kb];
          } else {
            const float rs = rad[sa] + rad[sb];
            r2 = rs * rs;
          }
          if (hit_sphere_sphere(cx[sa], cy[sa], cz[sa], cx[sb], cy[sb], cz[sb], r2))
            mark(((ba * NB + bb) * KM + ka) * KM + kb);
        }
    }
  }
  for (int o = 0; o < P.num_obstacles; ++o) {
    const int t = P.obstacle_type[o];
    float ov[16];
    for (int k = 0; k < 16; ++k) ov[k] = P.obstacles[o].v[k];
    const uint64_t m0 = P.obstacle_mask[2 * o], m1 = P.obstacle_mask[2 * o + 1];
    for (int s = 0; s < NS; ++s) {
      const bool on = s < 64 ? ((m0 >> s) & 1) : ((m1 >> (s - 64)) & 1);
      if (!on) continue;
      bool hit;
      if (t == CSD_BOX) {
        const float r2 = kTab ? P.obstacle_r2[(size_t)o * NS + s] : rad[s] * rad[s];
        hit = hit_sphere_obox(cx[s], cy[s], cz[s], r2, ov);
      } else if (t == CSD_SPHERE) {
        const float rs = rad[s] + ov[3];
        const float r2 = kTab ? P.obstacle_r2[(size_t)o * NS + s] : rs * rs;
        hit = hit_sphere_sphere(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], r2);
      } else {
        const float rs = rad[s] + ov[7];
        const float r2 = kTab ? P.obstacle_r2[(size_t)o * NS + s] : rs * rs;
        hit = seg_point_d2(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], ov[3], ov[4], ov[5],
                           ov[6]) < r2;
      }
      if (hit) mark(NB * NB * KM * KM + o * NS + s);
    }
  }
}

namespace {
int sfail(int code, const char *msg) {
  set_error(msg);
  return code;
}

template <int VAR>
void launch_finish(const csd_plant *pl, const float *q, int64_t n, int64_t sn, int64_t sj,
                   int32_t range, uint8_t *out, hipStream_t st) {
  range = std::min(std::max(range, 64), kMaxRange);  // the LDS list holds kMaxRange entries
  const size_t lds = check_lds_bytes(pl->img, VAR);
  const dim3 g((unsigned)((n + range - 1) / range));
  hipLaunchKernelGGL((k_screen<VAR>), dim3((unsigned)((n + kScreenBlock - 1) / kScreenBlock)),
                     dim3(kScreenBlock), (size_t)pl->screen.prog_words * 16, st, pl->img,
                     pl->screen, q, n, sn, sj, out);
  hipLaunchKernelGGL((k_finish<VAR>), g, dim3(kFinishBlock), lds, st, pl->img, q, n, sn, sj,
                     range, out);
}
}  // namespace

// the 8-slot-per-body shapes hold 64 or 128 sphere centres per lane and
// would spill in 512-thread survivor workgroups
bool screen_supported(int var) { return var == kFast8x4 || var == kFast16x4; }

void launch_screened(int var, const csd_plant *pl, const float *q, int64_t n, int64_t sn,
                     int64_t sj, int32_t range, uint8_t *out, hipStream_t st) {
  if (var == kFast8x4)
    launch_finish<kFast8x4>(pl, q, n, sn, sj, range, out, st);
  else
    launch_finish<kFast16x4>(pl, q, n, sn, sj, range, out, st);
}

// Greedy set cover over the sampled hit bits (below), stopping at the
// prefix that minimises (screen tests) + kFinishWeight * (undecided
// fraction) * (all tests). Returns the expected undecided fraction.
namespace {
constexpr double kFinishWeight = 2.0;  // a survivor's full check costs ~2x its share
constexpr int kScreenMaxGroups = 48;    // groups of 4 pair tests
}  // namespace

int build_screen(int var, csd_plant *pl, const float *q, int64_t n, int64_t sn, int64_t sj,
                 hipStream_t st, double *undecided) {
  if (!screen_supported(var)) return sfail(CSD_ERR_INVALID_ARGUMENT, "no screen for this plant shape");
  const PlantImage &P = pl->img;
  const int NB = (var == kFast8x4 || var == kFast8x8) ? 8 : 16;
  const int KM = (var == kFast8x4 || var == kFast16x4) ? 4 : 8;
  const int NS = NB * KM;
  const int self_pos = NB * NB * KM * KM;
  const int npos = self_pos + P.num_obstacles * NS;
  const int nwords = (npos + 31) / 32;
  const int64_t m = std::min<int64_t>(n, 4096);  // sample: every (n/m)-th configuration
  const int64_t stride = n / m;
  uint32_t *dbits = nullptr;
  CSD_HIP(hipMalloc((void **)&dbits, (size_t)nwords * m * 4));
  hipError_t e = hipMemsetAsync(dbits, 0, (size_t)nwords * m * 4, st);
  if (e == hipSuccess) {
#define CSD_HITS(V_)                                                                           \
  hipLaunchKernelGGL((k_pair_hits<V_>), dim3((unsigned)((m + kScreenBlock - 1) / kScreenBlock)), \
                     dim3(kScreenBlock), 0, st, P, q, m, sn * stride, sj, dbits)
    if (var == kFast8x4)
      CSD_HITS(kFast8x4);
    else
      CSD_HITS(kFast16x4);
#undef CSD_HITS
    e = hipGetLastError();
  }
  std::vector<uint32_t> bits((size_t)nwords * m);
  if (e == hipSuccess) e = hipMemcpyAsync(bits.data(), dbits, bits.size() * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(dbits);
  if (e != hipSuccess) return hip_fail(e, "screen statistics");

  // per pair position: the set of sample configurations it hits
  const int64_t cw = (m + 63) / 64;
  std::vector<int> cand;
  std::vector<std::vector<uint64_t>> sets;
  for (int p = 0; p < npos; ++p) {
    const uint32_t *row = &bits[(size_t)(p >> 5) * m];
    std::vector<uint64_t> s(cw, 0);
    bool any = false;
    for (int64_t i = 0; i < m; ++i)
      if ((row[i] >> (p & 31)) & 1) {
        s[i >> 6] |= 1ull << (i & 63);
        any = true;
      }
    if (any) {
      cand.push_back(p);
      sets.push_back(std::move(s));
    }
  }
  std::vector<uint64_t> open(cw, ~0ull);
  if (m % 64) open[cw - 1] = (1ull << (m % 64)) - 1;
  // Tests run in groups of 4 (position / 4: a self row or an obstacle-body
  // pair): a pair in a group already screened costs nothing more, so those
  // are taken first; otherwise the pair that decides the most still
  // undecided sample configurations opens a new group.
  const double full_tests = (double)P.num_pairs;
  int64_t left = m;
  std::vector<int> order;
  std::vector<int64_t> left_after;
  std::vector<int> groups_after;
  std::vector<char> used(cand.size(), 0);
  std::vector<char> in_group((size_t)npos / 4 + 1, 0);
  int groups = 0;
  while (groups <= kScreenMaxGroups) {
    int best = -1, best_free = -1;
    int64_t gain = 0, gain_free = 0;
    for (size_t c = 0; c < cand.size(); ++c) {
      if (used[c]) continue;
      int64_t g = 0;
      for (int64_t w = 0; w < cw; ++w) g += __builtin_popcountll(sets[c][w] & open[w]);
      if (in_group[cand[c] / 4]) {
        if (g > gain_free) {
          gain_free = g;
          best_free = (int)c;
        }
      } else if (g > gain) {
        gain = g;
        best = (int)c;
      }
    }
    if (best_free >= 0) {
      best = best_free;
      gain = gain_free;
    } else if (best < 0 || groups == kScreenMaxGroups) {
      break;
    } else {
      in_group[cand[best] / 4] = 1;
      ++groups;
    }
    used[best] = 1;
    for (int64_t w = 0; w < cw; ++w) open[w] &= ~sets[best][w];
    left -= gain;
    order.push_back(cand[best]);
    left_after.push_back(left);
    groups_after.push_back(groups);
  }
  int k_best = 0;
  double c_best = kFinishWeight * full_tests;
  for (size_t k = 0; k < order.size(); ++k) {
    const double c =
        4.0 * groups_after[k] + kFinishWeight * full_tests * (double)left_after[k] / m;
    if (c < c_best) {
      c_best = c;
      k_best = (int)k + 1;
    }
  }
  *undecided = k_best ? (double)left_after[k_best - 1] / m : 1.0;

  // screen program: self masks and table, then obstacles in first-chosen
  // order (the plant's records and contact distances, downloaded)
  std::vector<uint64_t> self_active(NB, 0), self_hot((size_t)NB * NB, 0);
  std::vector<int32_t> obst;
  std::vector<uint64_t> omask;
  for (int k = 0; k < k_best; ++k) {
    const int p = order[k];
    if (p < self_pos) {
      const int blk = p / (KM * KM), ab = p % (KM * KM);
      const int ba = blk / NB, bb = blk % NB;
      self_active[ba] |= 1ull << bb;
      self_hot[(size_t)ba * NB + bb] |= 1ull << ab;
    } else {
      const int o = (p - self_pos) / NS, s = (p - self_pos) % NS;
      size_t h = std::find(obst.begin(), obst.end(), o) - obst.begin();
      if (h == obst.size()) {
        if ((int)obst.size() == kScreenMaxObst) continue;
        obst.push_back(o);
        omask.push_back(0);
      }
      omask[h] |= 1ull << s;  // NS <= 64
    }
  }
  const size_t self_tab = (size_t)NB * NB * KM * KM;
  const size_t rec = 64 + 4 * (size_t)NS + 16;
  const size_t o_obst = 8 * (size_t)NB + 8 * (size_t)NB * NB + 4 * self_tab;
  const size_t bytes = (o_obst + obst.size() * rec + 15) / 16 * 16;
  std::vector<char> img(bytes, 0);
  // groups of 4 pair tests: self row (ba, bb, ka) and obstacle body (o, b);
  // a group's unscreened tests get contact distance -1 (never a hit)
  std::vector<float> r2tab(self_tab), r2prog(self_tab, -1.f), orow(NS);
  std::vector<uint64_t> rows((size_t)NB * NB, 0);
  std::vector<int32_t> otype(std::max(P.num_obstacles, 1));
  CSD_HIP(hipStreamSynchronize(st));
  CSD_HIP(hipMemcpy(r2tab.data(), P.self_r2, 4 * self_tab, hipMemcpyDeviceToHost));
  if (P.num_obstacles > 0)
    CSD_HIP(hipMemcpy(otype.data(), P.obstacle_type, 4 * (size_t)P.num_obstacles,
                      hipMemcpyDeviceToHost));
  for (int blk = 0; blk < NB * NB; ++blk)
    for (int ab = 0; ab < KM * KM; ++ab)
      if ((self_hot[blk] >> ab) & 1) {
        rows[blk] |= 1ull << (ab / KM);
        r2prog[(size_t)blk * KM * KM + ab] = r2tab[(size_t)blk * KM * KM + ab];
      }
  memcpy(&img[0], self_active.data(), 8 * NB);
  memcpy(&img[8 * NB], rows.data(), 8 * (size_t)NB * NB);
  memcpy(&img[8 * NB + 8 * (size_t)NB * NB], r2prog.data(), 4 * self_tab);
  for (size_t h = 0; h < obst.size(); ++h) {
    char *r = &img[o_obst + h * rec];
    CSD_HIP(hipMemcpy(r, P.obstacles + obst[h], 64, hipMemcpyDeviceToHost));
    CSD_HIP(hipMemcpy(orow.data(), P.obstacle_r2 + (size_t)obst[h] * NS, 4 * (size_t)NS,
                      hipMemcpyDeviceToHost));
    uint64_t bodies = 0;
    for (int sl = 0; sl < NS; ++sl) {
      const bool on = (omask[h] >> sl) & 1;
      if (on) bodies |= 1ull << (sl / KM);
      const float v = on ? orow[sl] : -1.f;
      memcpy(r + 64 + 4 * sl, &v, 4);
    }
    memcpy(r + 64 + 4 * NS, &bodies, 8);
    memcpy(r + 64 + 4 * NS + 8, &otype[obst[h]], 4);
  }
  constexpr size_t kArenaBytes = 64 * 1024;  // >= the largest program (16 x 4 slots)
  if (bytes > kArenaBytes) return sfail(CSD_ERR_CAPACITY, "screen program too large");
  if (!pl->screen_arena) CSD_HIP(hipMalloc(&pl->screen_arena, kArenaBytes));
  CSD_HIP(hipMemcpy(pl->screen_arena, img.data(), bytes, hipMemcpyHostToDevice));
  ScreenImage S;
  S.num_obst = (int32_t)obst.size();
  S.num_tests = k_best;
  S.prog_words = (int32_t)(bytes / 16);
  S.pad = 0;
  S.prog = (const uint4 *)pl->screen_arena;
  pl->screen = S;
  return CSD_OK;
}

}  // namespace csd
