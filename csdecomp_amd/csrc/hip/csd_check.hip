// csd_check.hip -- fused forward kinematics + collision checking for gfx950.
//
// One lane owns one configuration end to end: FK of the rigid bodies, world
// positions of the moving spheres in VGPRs, then every collision pair of the
// plant, with a wave-level early exit once all 64 lanes have found a
// collision. Nothing is materialised in HBM except the 1-byte flag per
// configuration (the reference writes 64 B/link of transforms and 1 B per
// (config, pair): plant/cuda_collision_checker.cu:16-85).
#include <string.h>

#include <algorithm>
#include <vector>

#include "csd_collide.h"
#include "csd_screen.h"

namespace csd {

constexpr int kCheckBlock = 256;

// S waves share each group of 64 configurations: wave w of a block takes
// part w % S of the pair tests of group w / S (all of them when S = 1), and
// the parts are OR-ed through LDS. S > 1 fills the chip when the batch is
// small (65,536 configurations are 1,024 waves, one per SIMD).
template <int VAR, int S>
__global__ __launch_bounds__(kCheckBlock) void k_check(PlantImage P, const float *__restrict__ q,
                                                       int64_t n, int64_t sn, int64_t sj,
                                                       VoxelArgs V, RobotGeomList R,
                                                       uint8_t *__restrict__ out) {
  constexpr int kPerBlock = kCheckBlock / S;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = wave % S, grp = wave / S;
  const int64_t gi = (int64_t)blockIdx.x * kPerBlock + grp * 64 + lane;
  const bool valid = gi < n;
  const int64_t i = valid ? gi : n - 1;  // idle tail lanes re-read the last config
  const QMem qa{q, i * sn, sj};
  extern __shared__ __attribute__((aligned(16))) float csd_smem[];
  const float *tab = stage_tables<VAR>(P, csd_smem);
  // idle lanes start "done" so they never hold the wave back
  const bool col = collides_at<VAR>(P, qa, V, R, !valid, tab, sub, S);
  if constexpr (S == 1) {
    if (valid) out[gi] = col ? 0 : 1;
  } else {
    __shared__ uint8_t part[kCheckBlock];
    part[threadIdx.x] = col ? 1 : 0;
    __syncthreads();
    if (sub == 0 && valid) {
      bool c = false;
#pragma unroll
      for (int s = 0; s < S; ++s) c |= part[(grp * S + s) * 64 + lane] != 0;
      out[gi] = c ? 0 : 1;
    }
  }
}

// Two passes over a tile of kCheckBlock / S1 configurations per workgroup,
// for batches where most configurations collide with the robot itself:
//  1. FK + the moving x moving pairs (S1 waves per 64 configurations, parts
//     OR-ed through LDS); colliding configurations are written at once;
//  2. the survivors, compacted in LDS, 64 at a time: every wave of the
//     workgroup takes the same 64 and tests its share of the obstacles and
//     voxels (o = wave mod 4), verdicts OR-ed through LDS.
// Pass 2 repeats FK (a few percent of a configuration's work) rather than
// keeping sphere centres in LDS.
template <int VAR, int S1>
__global__ __launch_bounds__(kCheckBlock) void k_check2(PlantImage P, const float *__restrict__ q,
                                                        int64_t n, int64_t sn, int64_t sj,
                                                        VoxelArgs V, RobotGeomList R,
                                                        uint8_t *__restrict__ out) {
  constexpr int kTile = kCheckBlock / S1;
  constexpr int kWaves = kCheckBlock / 64;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = wave % S1, grp = wave / S1;
  const int64_t base = (int64_t)blockIdx.x * kTile;
  extern __shared__ __attribute__((aligned(16))) float csd_smem[];
  __shared__ uint8_t part[kCheckBlock];
  __shared__ int16_t surv[kTile];
  __shared__ int s_cnt[kWaves];
  const float *tab = stage_tables<VAR>(P, csd_smem);
  // pass 1
  const int loc = grp * 64 + lane;
  const int64_t gi = base + loc;
  const bool valid = gi < n;
  const int64_t i = valid ? gi : n - 1;
  bool col = collides_at<VAR, false, 1>(P, QMem{q, i * sn, sj}, V, R, !valid, tab, sub, S1);
  if constexpr (S1 > 1) {
    part[threadIdx.x] = col ? 1 : 0;
    __syncthreads();
    col = false;
#pragma unroll
    for (int s = 0; s < S1; ++s) col |= part[(grp * S1 + s) * 64 + lane] != 0;
    __syncthreads();
  }
  if (sub == 0 && valid && col) out[gi] = 0;
  const bool alive = sub == 0 && valid && !col;
  const uint64_t m = __ballot(alive);
  if (lane == 0) s_cnt[wave] = __popcll(m);
  __syncthreads();
  int off = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    off += w < wave ? s_cnt[w] : 0;
    total += s_cnt[w];
  }
  if (alive) surv[off + __popcll(m & ((1ull << lane) - 1ull))] = (int16_t)loc;
  __syncthreads();
  // pass 2 (total is uniform over the workgroup)
  for (int g = 0; g < total; g += 64) {
    const int k = g + lane;
    const bool ok = k < total;
    const int64_t j = base + surv[ok ? k : total - 1];
    const bool c =
        collides_at<VAR, false, 2>(P, QMem{q, j * sn, sj}, V, R, !ok, tab, wave, kWaves);
    part[threadIdx.x] = c ? 1 : 0;
    __syncthreads();
    if (wave == 0 && ok) {
      bool cc = false;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) cc |= part[w * 64 + lane] != 0;
      out[j] = cc ? 0 : 1;
    }
    __syncthreads();
  }
}

// per (config, voxel): 1 iff no robot geometry touches the voxel
template <int NB>
__global__ __launch_bounds__(kCheckBlock) void k_voxel_flags(PlantImage P,
                                                             const float *__restrict__ q,
                                                             int64_t n, int64_t sn, int64_t sj,
                                                             VoxelArgs V, RobotGeomList R,
                                                             uint8_t *__restrict__ out) {
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n) return;
  CSD_BODY_FRAMES(B, NB);
  fk_bodies<NB>(P, QMem{q, gi * sn, sj}, B);
  for (int v = 0; v < V.num; ++v) {
    float Y[12];
    aff_identity(Y);
    Y[3] = V.vox[3 * v];
    Y[7] = V.vox[3 * v + 1];
    Y[11] = V.vox[3 * v + 2];
    const float vd[4] = {V.radius, 0.f, 0.f, 0.f};
    bool hit = false;
    for (int r = 0; r < R.n; ++r) {
      const DevGeom &g = P.geoms[R.ids[r]];
      float X[12];
      geom_world<NB>(g, B, X);
      hit |= hit_posed(0, Y, vd, geom_rank(g.type), X, g.dims);
    }
    out[gi * V.num + v] = hit ? 0 : 1;
  }
}

// standalone FK: transforms [N][L][16] column-major (the reference layout,
// plant/cuda_forward_kinematics.cu:59-77)
template <int NB>
__global__ __launch_bounds__(kCheckBlock) void k_fk(PlantImage P, const float *__restrict__ q,
                                                    int64_t n, int64_t sn, int64_t sj,
                                                    float *__restrict__ tf) {
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n) return;
  CSD_BODY_FRAMES(B, NB);
  fk_bodies<NB>(P, QMem{q, gi * sn, sj}, B);
  for (int l = 0; l < P.num_links; ++l) {
    const DevLink &lk = P.links[l];
    float X[12];
    if (lk.body < 0) {
#pragma unroll
      for (int k = 0; k < 12; ++k) X[k] = lk.O[k];
    } else {
      float Y[12];
      B.get(lk.body, Y);
      aff_mul(Y, lk.O, X);
    }
    float4 *o4 = reinterpret_cast<float4 *>(tf + (gi * P.num_links + l) * 16);
    o4[0] = make_float4(X[0], X[4], X[8], 0.f);
    o4[1] = make_float4(X[1], X[5], X[9], 0.f);
    o4[2] = make_float4(X[2], X[6], X[10], 0.f);
    o4[3] = make_float4(X[3], X[7], X[11], 1.f);
  }
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------
namespace {
int fail(int code, const char *msg) {
  set_error(msg);
  return code;
}

int check_args(const csd_plant *pl, const float *q, int64_t n, int64_t sn, int64_t sj,
               const void *out) {
  if (!pl) return fail(CSD_ERR_INVALID_ARGUMENT, "null plant");
  if (n < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative configuration count");
  if (n > 0 && (!q || !out)) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  if (n > 0 && (sn < 0 || sj < 0)) return fail(CSD_ERR_INVALID_ARGUMENT, "negative stride");
  if (pl->img.num_bodies > 16) return fail(CSD_ERR_CAPACITY, "more than 16 moving joints");
  return CSD_OK;
}

dim3 grid_for(int64_t n) { return dim3((unsigned)((n + kCheckBlock - 1) / kCheckBlock)); }

int size_class(int64_t n) { return n <= 0 ? 0 : 64 - __builtin_clzll((unsigned long long)n); }

// waves per configuration group: the tuned value for this batch-size class,
// else enough to put about four waves on each of the 1,024 SIMDs
int check_split(const csd_plant *pl, int64_t n, int var) {
  if (var >= kGeneric8) return 1;
  const int t = pl->check_split[size_class(n)];
  if (t) return t;
  const int64_t waves = (n + 63) / 64;
  return waves >= 4096 ? 1 : waves >= 2048 ? 2 : 4;
}

// split: 1, 2 or 4 -- one pass, that many waves per 64 configurations;
// 11, 12 or 14 -- two passes (k_check2), pass 1 with 1, 2 or 4 waves per 64
void launch_k_check(int var, int split, const csd_plant *pl, const float *q, int64_t n,
                    int64_t sn, int64_t sj, const VoxelArgs &V, const RobotGeomList &R,
                    uint8_t *out, hipStream_t st) {
  const size_t lds = check_lds_bytes(pl->img, var);
  const dim3 b(kCheckBlock);
  if (split > 10 && var < kGeneric8) {
    const int s1 = split - 10;
    const int64_t tile = kCheckBlock / s1;
    const dim3 g((unsigned)((n + tile - 1) / tile));
#define CSD_CHECK2(V_, S_)                                                                   \
  hipLaunchKernelGGL((k_check2<V_, S_>), g, b, lds, st, pl->img, q, n, sn, sj, V, R, out)
#define CSD_CHECK2_S(V_)                 \
  do {                                   \
    if (s1 == 4) CSD_CHECK2(V_, 4);      \
    else if (s1 == 2) CSD_CHECK2(V_, 2); \
    else CSD_CHECK2(V_, 1);              \
  } while (0)
    switch (var) {
      case kFast8x4: CSD_CHECK2_S(kFast8x4); break;
      case kFast8x8: CSD_CHECK2_S(kFast8x8); break;
      case kFast16x4: CSD_CHECK2_S(kFast16x4); break;
      default: CSD_CHECK2_S(kFast16x8); break;
    }
#undef CSD_CHECK2_S
#undef CSD_CHECK2
    return;
  }
  if (split == 4)
    CSD_DISPATCH_VARIANT_X(var, k_check, 4, dim3((unsigned)((n + 63) / 64)), b, lds, st, pl->img,
                           q, n, sn, sj, V, R, out);
  else if (split == 2)
    CSD_DISPATCH_VARIANT_X(var, k_check, 2, dim3((unsigned)((n + 127) / 128)), b, lds, st,
                           pl->img, q, n, sn, sj, V, R, out);
  else
    CSD_DISPATCH_VARIANT_X(var, k_check, 1, grid_for(n), b, lds, st, pl->img, q, n, sn, sj, V,
                           R, out);
}
}  // namespace

bool fill_robot_list(const csd_plant *pl, const int32_t *robot, int32_t nr, RobotGeomList &R,
                     VoxelArgs &V, int *status) {
  memset(&R, 0, sizeof(R));
  V.mask0 = V.mask1 = 0;
  *status = CSD_OK;
  if (nr < 0 || nr > kMaxRobotGeoms) {
    *status = fail(CSD_ERR_CAPACITY, "robot geometry list longer than 128");
    return false;
  }
  if (nr > 0 && !robot) {
    *status = fail(CSD_ERR_INVALID_ARGUMENT, "null robot geometry list");
    return false;
  }
  bool all_slots = true;
  R.n = nr;
  for (int k = 0; k < nr; ++k) {
    int g = robot[k];
    if (g < 0 || g >= (int)pl->geom_type.size()) {
      *status = fail(CSD_ERR_INVALID_ARGUMENT, "robot geometry index out of range");
      return false;
    }
    R.ids[k] = g;
    int s = pl->geom_slot[g];
    if (s < 0)
      all_slots = false;
    else if (s < 64)
      V.mask0 |= (uint64_t)1 << s;
    else
      V.mask1 |= (uint64_t)1 << (s - 64);
  }
  return all_slots;
}

int launch_check(const csd_plant *pl, const float *q, int64_t n, int64_t sn, int64_t sj,
                 const float *vox, int32_t nvox, float vr, const int32_t *robot, int32_t nr,
                 uint8_t *out, hipStream_t st) {
  int rc = check_args(pl, q, n, sn, sj, out);
  if (rc) return rc;
  if (n == 0) return CSD_OK;
  if (nvox < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative voxel count");
  if (nvox > 0 && !vox) return fail(CSD_ERR_INVALID_ARGUMENT, "null voxel array");
  RobotGeomList R;
  VoxelArgs V{vox, nvox, vr, 0, 0};
  bool all_slots = fill_robot_list(pl, robot, nvox > 0 ? nr : 0, R, V, &rc);
  if (rc) return rc;
  const int var = check_variant(pl->img, nvox > 0 && !all_slots);
  const int32_t range = pl->check_range[size_class(n)];
  if (nvox == 0 && range > 0 && screen_supported(var))
    launch_screened(var, pl, q, n, sn, sj, range, out, st);
  else
    launch_k_check(var, check_split(pl, n, var), pl, q, n, sn, sj, V, R, out, st);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

}  // namespace csd

using namespace csd;

extern "C" {

int csd_check_collision_free(const csd_plant *plant, const float *q, int64_t n,
                             int64_t stride_n, int64_t stride_j, uint8_t *out, void *stream) {
  return launch_check(plant, q, n, stride_n, stride_j, nullptr, 0, 0.f, nullptr, 0, out,
                      (hipStream_t)stream);
}

int csd_check_collision_free_voxels(const csd_plant *plant, const float *q, int64_t n,
                                    int64_t stride_n, int64_t stride_j, const float *voxels,
                                    int32_t num_voxels, float voxel_radius,
                                    const int32_t *robot_geoms, int32_t num_robot_geoms,
                                    uint8_t *out, void *stream) {
  return launch_check(plant, q, n, stride_n, stride_j, voxels, num_voxels, voxel_radius,
                      robot_geoms, num_robot_geoms, out, (hipStream_t)stream);
}

int csd_tune_check(const csd_plant *plant, const float *q, int64_t n, int64_t stride_n,
                   int64_t stride_j, uint8_t *out, void *stream, int32_t *split) {
  int rc = check_args(plant, q, n, stride_n, stride_j, out);
  if (rc) return rc;
  if (n == 0) return CSD_OK;
  hipStream_t st = (hipStream_t)stream;
  RobotGeomList R;
  memset(&R, 0, sizeof(R));
  const VoxelArgs V{nullptr, 0, 0.f, 0, 0};
  const int var = check_variant(plant->img, false);
  csd_plant *pl = const_cast<csd_plant *>(plant);
  const int cls = size_class(n);
  int best = 1;
  const int forced = split ? *split : 0;
  // survivor-workgroup ranges worth timing for a screen that leaves a
  // fraction u undecided: about 32 to 512 undecided per workgroup (1 batch
  // of 64 with 8 waves sharing its pair tests up to 8 batches of 1 wave)
  auto ranges_for = [&](double u) {
    std::vector<int32_t> r;
    for (double t : {32.0, 64.0, 128.0, 256.0, 512.0}) {
      int64_t v = (int64_t)(t / std::max(u, 1e-4));
      v = std::min<int64_t>(std::max<int64_t>((v + 63) / 64 * 64, 512), 16384);
      if (std::find(r.begin(), r.end(), (int32_t)v) == r.end()) r.push_back((int32_t)v);
    }
    return r;
  };
  if (forced != 0) {  // record a given shape instead of timing
    if (!(forced == 1 || forced == 2 || forced == 4 || forced == 11 || forced == 12 ||
          forced == 14 || forced == 20))
      return fail(CSD_ERR_INVALID_ARGUMENT, "split must be 1, 2, 4, 11, 12, 14 or 20");
    best = var < kGeneric8 ? forced : 1;
    if (best == 20 && !screen_supported(var)) best = 1;
    pl->check_range[cls] = 0;
    if (best == 20) {
      double u = 1.0;
      rc = build_screen(var, pl, q, n, stride_n, stride_j, st, &u);
      if (rc) return rc;
      const std::vector<int32_t> rs = ranges_for(u);
      pl->check_range[cls] = rs[std::min<size_t>(1, rs.size() - 1)];
      launch_screened(var, pl, q, n, stride_n, stride_j, pl->check_range[cls], out, st);
    } else {
      pl->check_split[cls] = (int8_t)best;
      launch_k_check(var, best, plant, q, n, stride_n, stride_j, V, R, out, st);
    }
  } else if (var < kGeneric8) {
    hipEvent_t e0, e1;
    CSD_HIP(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
      hipEventDestroy(e0);
      return hip_fail(e, "hipEventCreate");
    }
    float best_ms = INFINITY;
    // one timing of `launch` (a warm-up, then 5 launches between events)
    auto time5 = [&](auto launch) {
      launch();
      e = hipEventRecord(e0, st);
      for (int k = 0; k < 5 && e == hipSuccess; ++k) launch();
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = INFINITY;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      return ms;
    };
    for (int s : {1, 2, 4, 11, 12, 14}) {
      const float ms =
          time5([&] { launch_k_check(var, s, plant, q, n, stride_n, stride_j, V, R, out, st); });
      if (e != hipSuccess) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = s;
      }
    }
    int32_t best_range = 0;
    if (e == hipSuccess && screen_supported(var)) {
      double u = 1.0;
      rc = build_screen(var, pl, q, n, stride_n, stride_j, st, &u);
      if (rc) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return rc;
      }
      for (int32_t r : ranges_for(u)) {
        const float ms = time5(
            [&] { launch_screened(var, pl, q, n, stride_n, stride_j, r, out, st); });
        if (e != hipSuccess) break;
        if (ms < best_ms) {
          best_ms = ms;
          best_range = r;
        }
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "check autotuning");
    pl->check_range[cls] = best_range;
    if (best_range > 0) {
      best = 20;
      launch_screened(var, pl, q, n, stride_n, stride_j, best_range, out, st);
    } else {
      pl->check_split[cls] = (int8_t)best;
      launch_k_check(var, best, plant, q, n, stride_n, stride_j, V, R, out, st);
    }
  }
  else {
    rc = launch_check(plant, q, n, stride_n, stride_j, nullptr, 0, 0.f, nullptr, 0, out, st);
    if (rc) return rc;
  }
  CSD_HIP(hipGetLastError());
  if (split) *split = best;
  return CSD_OK;
}

int csd_voxel_flags(const csd_plant *plant, const float *q, int64_t n, int64_t stride_n,
                    int64_t stride_j, const float *voxels, int32_t num_voxels,
                    float voxel_radius, const int32_t *robot_geoms, int32_t num_robot_geoms,
                    uint8_t *out, void *stream) {
  int rc = check_args(plant, q, n, stride_n, stride_j, out);
  if (rc) return rc;
  if (num_voxels < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative voxel count");
  if (n == 0 || num_voxels == 0) return CSD_OK;
  if (!voxels) return fail(CSD_ERR_INVALID_ARGUMENT, "null voxel array");
  RobotGeomList R;
  VoxelArgs V{voxels, num_voxels, voxel_radius, 0, 0};
  fill_robot_list(plant, robot_geoms, num_robot_geoms, R, V, &rc);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (plant->img.num_bodies <= 8)
    hipLaunchKernelGGL((k_voxel_flags<8>), grid_for(n), dim3(kCheckBlock), 0, st, plant->img, q,
                       n, stride_n, stride_j, V, R, out);
  else
    hipLaunchKernelGGL((k_voxel_flags<16>), grid_for(n), dim3(kCheckBlock), 0, st, plant->img,
                       q, n, stride_n, stride_j, V, R, out);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

int csd_forward_kinematics(const csd_plant *plant, const float *q, int64_t n, int64_t stride_n,
                           int64_t stride_j, float *transforms, void *stream) {
  int rc = check_args(plant, q, n, stride_n, stride_j, transforms);
  if (rc) return rc;
  if (n == 0) return CSD_OK;
  hipStream_t st = (hipStream_t)stream;
  if (plant->img.num_bodies <= 8)
    hipLaunchKernelGGL((k_fk<8>), grid_for(n), dim3(kCheckBlock), 0, st, plant->img, q, n,
                       stride_n, stride_j, transforms);
  else
    hipLaunchKernelGGL((k_fk<16>), grid_for(n), dim3(kCheckBlock), 0, st, plant->img, q, n,
                       stride_n, stride_j, transforms);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

}  // extern "C"
