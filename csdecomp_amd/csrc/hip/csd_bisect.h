// csd_bisect.h -- the standalone projection + bisection kernel (csd_bisect,
// csd_edit_regions). Its instantiations are split over two translation units
// by padded dimension (csd_bisect8.hip, csd_bisect16.hip) so that they
// compile in parallel.
#pragma once
#include "csd_eizo_device.h"

namespace csd {

struct BisectLaunch {
  PlantImage P;
  VoxelArgs V;
  RobotGeomList R;
  const float *part, *ls, *le;
  const uint32_t *seg;
  int n, dim, steps;
  float *scol, *proj;
  dim3 grid, block;
  size_t lds;
  hipStream_t st;
};

// launch k_bisect_standalone for dim <= 8 / dim <= 16 with `split` (1, 2 or
// 4) waves per 64 particles, for plant variant `var`
void launch_bisect_dp8(const BisectLaunch &L, int var, int split);
void launch_bisect_dp16(const BisectLaunch &L, int var, int split);

#ifdef CSD_BISECT_DEFINE_KERNEL
// DPS = 8 * DP + S: DP padded dimensions, S waves per 64 particles
template <int VAR, int DPS>
__global__ __launch_bounds__(kSampleBlock) void k_bisect_standalone(
    PlantImage P, VoxelArgs V, RobotGeomList R, const float *__restrict__ part,
    const float *__restrict__ ls, const float *__restrict__ le,
    const uint32_t *__restrict__ seg, int n, int dim, int steps, float *__restrict__ scol,
    float *__restrict__ proj) {
  constexpr int DP = DPS / 8, S = DPS % 8;
  constexpr int kPer = kSampleBlock / S;
  extern __shared__ __attribute__((aligned(16))) float csd_smem[];
  __shared__ uint8_t verdict[kSampleBlock];
  const float *tab = stage_tables<VAR>(P, csd_smem);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = wave % S, grp = wave / S;
  const int i = blockIdx.x * kPer + grp * 64 + lane;
  const bool valid = i < n;
  const int ii = valid ? i : n - 1;
  const uint32_t s_idx = seg[ii];
  float p[DP], q1[DP], q2[DP], sc[DP];
#pragma unroll
  for (int d = 0; d < DP; ++d) {
    p[d] = d < dim ? part[(size_t)ii * dim + d] : 0.f;
    q1[d] = d < dim ? ls[(size_t)s_idx * dim + d] : 0.f;
    q2[d] = d < dim ? le[(size_t)s_idx * dim + d] : 0.f;
  }
  const bool owner = valid && sub == 0;  // writes the particle's results
  bisect_particle<VAR, DP, S>(P, V, R, p, q1, q2, steps, !valid, sc,
                              owner ? proj + (size_t)i * dim : nullptr, dim, tab, sub, verdict);
  if (!owner) return;
#pragma unroll
  for (int d = 0; d < DP; ++d)
    if (d < dim) scol[(size_t)i * dim + d] = sc[d];
}


template <int DP>
void launch_bisect_dp(const BisectLaunch &L, int var, int split) {
#define CSD_BISECT(DPS)                                                                        \
  CSD_DISPATCH_VARIANT_X(var, k_bisect_standalone, DPS, L.grid, L.block, L.lds, L.st, L.P, L.V, \
                         L.R, L.part, L.ls, L.le, L.seg, L.n, L.dim, L.steps, L.scol, L.proj)
  if (split == 4) CSD_BISECT(8 * DP + 4);
  else if (split == 2) CSD_BISECT(8 * DP + 2);
  else CSD_BISECT(8 * DP + 1);
#undef CSD_BISECT
}
#endif

}  // namespace csd
