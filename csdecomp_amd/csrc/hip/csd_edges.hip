// csd_edges.hip -- consumers of the collision kernel for roadmaps:
// edge validation (configurations generated in-lane, never stored), swept
// voxel occupancy, brute-force k nearest neighbours.
#include <math.h>
#include <string.h>

#include <utility>
#include <vector>
#include <algorithm>

#include "csd_collide.h"

namespace csd {

constexpr int kEdgeBlock = 256;
constexpr int kEdgeDim = 16;
constexpr float kMaxEdgePoints = 1.0e8f;  // longer edges are rejected, not checked

// sequential float sum of squares without contraction, the order the CPU
// restatement uses (so step counts agree bit for bit)
__device__ __forceinline__ float norm_diff_rn(const float *a, const float *b, int dim) {
  float s = 0.f;
#pragma unroll
  for (int d = 0; d < kEdgeDim; ++d)
    if (d < dim) {
      const float e = __fsub_rn(b[d], a[d]);
      s = __fadd_rn(s, __fmul_rn(e, e));
    }
  return sqrtf(s);
}

// Point k of an edge under each rule (see k_validate_edges), the float
// expression the CPU restatement evaluates
__device__ __forceinline__ float edge_point_t(int rule, int64_t k, int64_t npts, float tstep) {
  if (rule == 0) return (float)((double)k / (1.0 * (double)npts));
  if (rule == 2) return (float)((double)(k + 1) / (1.0 * (double)npts));
  return k == 0 ? 0.f : (k == 1 ? 1.f : __fmul_rn(tstep, (float)(k - 1)));
}

// Edge validation in two phases per group of 64 edges, grid-stride over
// groups.
//  1. Lane i takes edge i of the group and checks its first point (k = 0).
//     In a roadmap over uniform nodes most edges end here: the first point
//     is a node (rules 0 and 1) and most nodes collide.
//  2. The edges whose first point is free and that have more points are
//     taken one at a time by the whole wave, lane l checking point
//     1 + l + 64 j, until a pass finds a collision or the points run out.
// Every point is checked with the arithmetic a single configuration gets,
// so the verdict (free iff no point collides) does not depend on the split.
// rule 0: checkEdgeCollisionFree, n = ceil(|q2-q1|/step), t = k/n, k < n
//         (plant/collision_checker.cpp:41-53)
// rule 1: VisibilityGraph, n = (uint)(|q2-q1|/step + 2) points: q1, q2 and
//         t = (k+1)/(n-1) for k < n-2 (planning/cuda_visibility_graph.cu:58-172)
// rule 2: RoadmapBuilder, n = max(0, floor(|q2-q1|/step) - 2), t = (k+1)/n,
//         k < n (q2 checked, q1 not; planning/roadmap_builder.cpp:262-280)
template <int VAR>
__global__ __launch_bounds__(kEdgeBlock) void k_validate_edges(
    PlantImage P, const float *__restrict__ nodes, int num_nodes, int dim,
    const int32_t *__restrict__ edges,
    int64_t E, float step, int rule, uint8_t *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const VoxelArgs V{nullptr, 0, 0.f, 0, 0};
  RobotGeomList R;
  R.n = 0;
  extern __shared__ __attribute__((aligned(16))) float csd_smem[];
  const float *tab = stage_tables<VAR>(P, csd_smem);
  const int64_t groups = (E + 63) / 64;
  for (int64_t grp = wave; grp < groups; grp += nwaves) {
    // ---- phase 1: one edge per lane, its first point ----
    const int64_t e = grp * 64 + lane;
    int a = 0, b = 0;
    bool live = e < E;  // an edge whose first point is still to be checked
    if (live) {
      a = edges[2 * e];
      b = edges[2 * e + 1];
      if (a < 0 || a >= num_nodes || b < 0 || b >= num_nodes) {
        out[e] = 0;  // invalid node index: reported as not free
        live = false;
      }
    }
    float q1[kEdgeDim], q2[kEdgeDim];
#pragma unroll
    for (int d = 0; d < kEdgeDim; ++d) {
      q1[d] = live && d < dim ? nodes[(size_t)a * dim + d] : 0.f;
      q2[d] = live && d < dim ? nodes[(size_t)b * dim + d] : 0.f;
    }
    int64_t npts = 0;
    float tstep = 0.f;
    if (live) {
      const float dist = norm_diff_rn(q1, q2, dim);
      if (!(dist / step < kMaxEdgePoints)) {  // non-finite or absurdly long edge
        out[e] = 0;
        live = false;
      } else if (rule == 0) {
        npts = (int64_t)ceilf(dist / step);
      } else if (rule == 2) {
        npts = max((int64_t)0, (int64_t)floorf(dist / step) - 2);
      } else {
        npts = (int64_t)(uint32_t)(dist / step + 2);
        tstep = (float)(1.0 / (double)(npts - 1));
      }
      if (live && npts == 0) {  // no point to check: free
        out[e] = 1;
        live = false;
      }
    }
    float q[kEdgeDim];
    {
      const float t = edge_point_t(rule, 0, live ? npts : 1, tstep);
#pragma unroll
      for (int d = 0; d < kEdgeDim; ++d)
        q[d] = __fadd_rn(__fmul_rn(t, q2[d]), __fmul_rn(__fsub_rn(1.f, t), q1[d]));
    }
    // lanes without a live edge start as collided, so the wave leaves the
    // pair loop once every live lane has collided
    const bool col0 = collides_variant<VAR>(P, QReg<kEdgeDim>{q}, V, R, !live, tab);
    if (live && (col0 || npts == 1)) out[e] = col0 ? 0 : 1;
    // ---- phase 2: the remaining points of each surviving edge ----
    uint64_t rest = __ballot(live && !col0 && npts > 1);
    while (rest) {
      const int src = __builtin_ctzll(rest);
      rest &= rest - 1;
      const int64_t n = __shfl(npts, src);
      const float ts = __shfl(tstep, src);
      float p1[kEdgeDim], p2[kEdgeDim];
#pragma unroll
      for (int d = 0; d < kEdgeDim; ++d) {
        p1[d] = __shfl(q1[d], src);
        p2[d] = __shfl(q2[d], src);
      }
      bool any = false;
      for (int64_t base = 1; base < n && !any; base += 64) {
        // idle tail lanes re-check the last point: in "any collision" mode a
        // lane that starts as collided would end the wave's checks early
        const int64_t k = min(base + lane, n - 1);
        const float t = edge_point_t(rule, k, n, ts);
#pragma unroll
        for (int d = 0; d < kEdgeDim; ++d)
          q[d] = __fadd_rn(__fmul_rn(t, p2[d]), __fmul_rn(__fsub_rn(1.f, t), p1[d]));
        const bool col = collides_variant<VAR, true, 3 | kCoherent>(P, QReg<kEdgeDim>{q}, V, R, false, tab);
        any = __any(col);
      }
      if (lane == src) out[e] = any ? 0 : 1;
    }
  }
}

// Swept occupancy: every robot geometry at every configuration marks the
// voxels it touches (same voxel test as geomToVoxelCollisionFree, GPU rule).
// Voxel centres follow GetCollisionVoxelCenter (roadmap_builder.cpp:709-741).
struct GridArgs {
  float center[3], size[3];
  float res, radius;
  int nx, ny, nz;
};

__device__ __forceinline__ float voxel_coord(float c, float size, int i, float res) {
  return (float)((double)c - (double)size / 2.0 + (double)(float)(i * 2 * res) + (double)res);
}

// Voxel index box [lo, hi] that can hold a voxel sphere touched by geometry
// g posed at X (its bounding box widened by the voxel radius, plus a margin
// of one voxel for rounding).
__device__ __forceinline__ void geom_voxel_range(const DevGeom &g, const float (&X)[12],
                                                 const GridArgs &G, int (&lo)[3], int (&hi)[3]) {
  const float two_res = 2.f * G.res;
  const int rk = geom_rank(g.type);
  const int nn[3] = {G.nx, G.ny, G.nz};
  for (int k = 0; k < 3; ++k) {
    float ext;
    if (rk == 0) ext = g.dims[0];
    else if (rk == 1)
      ext = fabsf(X[k * 4 + 0]) * g.dims[0] + fabsf(X[k * 4 + 1]) * g.dims[1] +
            fabsf(X[k * 4 + 2]) * g.dims[2];
    else ext = g.dims[0] + fabsf(X[k * 4 + 2]) * g.dims[1];
    ext = ext * 1.0001f + G.radius + 1e-6f;
    const float origin = G.center[k] - 0.5f * G.size[k] + G.res;
    const float p = X[k * 4 + 3];
    lo[k] = max(0, (int)floorf((p - ext - origin) / two_res) - 1);
    hi[k] = min(nn[k] - 1, (int)ceilf((p + ext - origin) / two_res) + 1);
  }
}

__device__ __forceinline__ bool voxel_hit(const DevGeom &g, const float (&X)[12], const GridArgs &G,
                                          int ix, int iy, int iz) {
  float Y[12];
  aff_identity(Y);
  Y[3] = voxel_coord(G.center[0], G.size[0], ix, G.res);
  Y[7] = voxel_coord(G.center[1], G.size[1], iy, G.res);
  Y[11] = voxel_coord(G.center[2], G.size[2], iz, G.res);
  const float vd[4] = {G.radius, 0.f, 0.f, 0.f};
  return hit_posed(0, Y, vd, geom_rank(g.type), X, g.dims);
}

template <int NB>
__global__ __launch_bounds__(kEdgeBlock) void k_occupancy(PlantImage P, const float *__restrict__ q,
                                                          int64_t n, int64_t sn, int64_t sj,
                                                          RobotGeomList R, GridArgs G,
                                                          uint8_t *__restrict__ grid) {
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n) return;
  CSD_BODY_FRAMES(B, NB);
  fk_bodies<NB>(P, QMem{q, gi * sn, sj}, B);
  const float two_res = 2.f * G.res;
  const float vd[4] = {G.radius, 0.f, 0.f, 0.f};
  for (int r = 0; r < R.n; ++r) {
    const DevGeom &g = P.geoms[R.ids[r]];
    float X[12];
    geom_world<NB>(g, B, X);
    float ext[3];
    const int rk = geom_rank(g.type);
    for (int k = 0; k < 3; ++k) {
      if (rk == 0) ext[k] = g.dims[0];
      else if (rk == 1)
        ext[k] = fabsf(X[k * 4 + 0]) * g.dims[0] + fabsf(X[k * 4 + 1]) * g.dims[1] +
                 fabsf(X[k * 4 + 2]) * g.dims[2];
      else ext[k] = g.dims[0] + fabsf(X[k * 4 + 2]) * g.dims[1];
      ext[k] = ext[k] * 1.0001f + G.radius + 1e-6f;
    }
    int lo[3], hi[3];
    const int nn[3] = {G.nx, G.ny, G.nz};
    for (int k = 0; k < 3; ++k) {
      const float origin = G.center[k] - 0.5f * G.size[k] + G.res;
      const float p = X[k * 4 + 3];
      lo[k] = max(0, (int)floorf((p - ext[k] - origin) / two_res) - 1);
      hi[k] = min(nn[k] - 1, (int)ceilf((p + ext[k] - origin) / two_res) + 1);
    }
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
      for (int iy = lo[1]; iy <= hi[1]; ++iy)
        for (int ix = lo[0]; ix <= hi[0]; ++ix) {
          float Y[12];
          aff_identity(Y);
          Y[3] = voxel_coord(G.center[0], G.size[0], ix, G.res);
          Y[7] = voxel_coord(G.center[1], G.size[1], iy, G.res);
          Y[11] = voxel_coord(G.center[2], G.size[2], iz, G.res);
          if (hit_posed(0, Y, vd, rk, X, g.dims))
            grid[(size_t)ix + (size_t)iy * G.nx + (size_t)iz * G.nx * G.ny] = 1;
        }
  }
}

// Voxel collision maps (RoadmapBuilder::BuildCollisionMap and
// BuildEdgeCollisionMap, planning/roadmap_builder.cpp:326-679): for every
// item -- a node, or the interior configurations of an edge -- the voxels
// whose sphere some robot geometry touches. One workgroup per item: for each
// configuration and geometry, its threads split the geometry's candidate
// voxels and mark hits in an LDS bitmap of the grid, which removes
// duplicates within the item. Then every marked voxel counts the item
// (pass 0) or stores it at the voxel's list cursor (pass 1).
// Edge configurations follow the roadmap rule (roadmap_builder.cpp:526-541):
// n = max(0, floor(|q2 - q1| / step) - 2), t = (k + 1) / n for k < n,
// q = t q2 + (1 - t) q1.
constexpr int kMapBlock = 256;
constexpr int kMapMaxWords = 16384;  // 64 KB of LDS: grids up to 524,288 voxels

template <int NB>
__global__ __launch_bounds__(kMapBlock) void k_voxel_map(
    PlantImage P, const float *__restrict__ nodes, int num_nodes, int dim,
    const int32_t *__restrict__ edges, float step, RobotGeomList R, GridArgs G, int pass,
    int32_t *__restrict__ counts, const int64_t *__restrict__ offsets,
    int32_t *__restrict__ cursor, int32_t *__restrict__ items) {
  extern __shared__ uint32_t bits[];
  const int item = blockIdx.x;
  const int nvox = G.nx * G.ny * G.nz;
  const int W = (nvox + 31) / 32;
  for (int w = threadIdx.x; w < W; w += kMapBlock) bits[w] = 0;
  __syncthreads();
  float q1[kEdgeDim], q2[kEdgeDim];
  int nc = 1;
  bool ok = true;
  if (edges) {
    const int a = edges[2 * item], b = edges[2 * item + 1];
    ok = a >= 0 && a < num_nodes && b >= 0 && b < num_nodes;
#pragma unroll
    for (int d = 0; d < kEdgeDim; ++d) {
      q1[d] = ok && d < dim ? nodes[(size_t)a * dim + d] : 0.f;
      q2[d] = ok && d < dim ? nodes[(size_t)b * dim + d] : 0.f;
    }
    const float dist = norm_diff_rn(q1, q2, dim);
    const float steps = floorf(dist / step) - 2.f;
    nc = ok && steps > 0.f && steps < kMaxEdgePoints ? (int)steps : 0;
  } else {
#pragma unroll
    for (int d = 0; d < kEdgeDim; ++d) q1[d] = d < dim ? nodes[(size_t)item * dim + d] : 0.f;
  }
  // configurations and geometries in a uniform loop (every thread computes
  // the same FK); the candidate voxels of a geometry are split over the
  // workgroup
  for (int c = 0; c < nc; ++c) {
    float q[kEdgeDim];
    if (edges) {
      const float t = (float)((double)(c + 1) / (1.0 * (double)nc));
#pragma unroll
      for (int d = 0; d < kEdgeDim; ++d)
        q[d] = __fadd_rn(__fmul_rn(t, q2[d]), __fmul_rn(__fsub_rn(1.f, t), q1[d]));
    } else {
#pragma unroll
      for (int d = 0; d < kEdgeDim; ++d) q[d] = q1[d];
    }
    CSD_BODY_FRAMES(B, NB);
    fk_bodies<NB>(P, QReg<kEdgeDim>{q}, B);
    for (int r = 0; r < R.n; ++r) {
      const DevGeom &g = P.geoms[R.ids[r]];
      float X[12];
      geom_world<NB>(g, B, X);
      int lo[3], hi[3];
      geom_voxel_range(g, X, G, lo, hi);
      const int sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
      if (sx <= 0 || sy <= 0 || sz <= 0) continue;
      const int count = sx * sy * sz;
      for (int k = threadIdx.x; k < count; k += kMapBlock) {
        const int ix = lo[0] + k % sx, iy = lo[1] + (k / sx) % sy, iz = lo[2] + k / (sx * sy);
        if (voxel_hit(g, X, G, ix, iy, iz)) {
          const int v = ix + iy * G.nx + iz * G.nx * G.ny;
          atomicOr(&bits[v >> 5], 1u << (v & 31));
        }
      }
    }
  }
  __syncthreads();
  for (int w = threadIdx.x; w < W; w += kMapBlock) {
    uint32_t m = bits[w];
    while (m) {
      const int v = w * 32 + __ffs(m) - 1;
      m &= m - 1;
      if (pass == 0) {
        atomicAdd(&counts[v], 1);
      } else {
        const int slot = atomicAdd(&cursor[v], 1);
        items[offsets[v] + slot] = item;
      }
    }
  }
}

// k nearest neighbours by squared L2 distance (sequential, uncontracted
// float sum); candidates with d2 <= FLT_EPSILON (the node itself and exact
// duplicates) are skipped as in roadmap_builder.cpp:190-197; ties by index.
// K (= k) is a template value, so the running threshold is the k-th entry of
// a register list with static indices. Node coordinates are staged in LDS in
// rows zero-padded to DP floats; the padding adds +0 to each sum.
constexpr int kKnnMax = 16;
// Paged use: query row i is node queries[i] (all nodes in order when
// queries is null) and, with after_d2 / after_idx, only candidates after
// (after_d2[i], after_idx[i]) in (distance, index) order count, so that
// successive calls return successive pages of the sorted neighbour list.
template <int K, int DP>
__global__ __launch_bounds__(256) void k_knn(const float *__restrict__ nodes, int N, int dim,
                                             const int32_t *__restrict__ queries, int Q,
                                             const float *__restrict__ after_d2,
                                             const int32_t *__restrict__ after_idx,
                                             int32_t *__restrict__ out_idx,
                                             float *__restrict__ out_d2) {
  __shared__ __attribute__((aligned(16))) float tile[256 * DP];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool valid = i < Q;
  const int qn = valid ? (queries ? queries[i] : i) : 0;
  const bool paged = after_d2 != nullptr;
  const float ad = paged && valid ? after_d2[i] : -INFINITY;
  const int ai = paged && valid ? after_idx[i] : -1;
  float qv[DP];
#pragma unroll
  for (int d = 0; d < DP; ++d)
    qv[d] = (valid && d < dim && qn >= 0 && qn < N) ? nodes[(size_t)qn * dim + d] : 0.f;
  float bd[K];
  int bi[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    bd[s] = INFINITY;
    bi[s] = 0x7fffffff;
  }
  for (int t0 = 0; t0 < N; t0 += 256) {
    const int cnt = min(256, N - t0);
    __syncthreads();
    for (int e = threadIdx.x; e < 256 * DP; e += 256) {
      const int j = e / DP, d = e - j * DP;
      tile[e] = (j < cnt && d < dim) ? nodes[(size_t)(t0 + j) * dim + d] : 0.f;
    }
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      float s = 0.f;
#pragma unroll
      for (int d = 0; d < DP; ++d) {
        const float e = __fsub_rn(tile[j * DP + d], qv[d]);
        s = __fadd_rn(s, __fmul_rn(e, e));
      }
      const int jj = t0 + j;
      const bool later = s > ad || (s == ad && jj > ai);
      if (s > 1.1920928955078125e-07f && later &&
          (s < bd[K - 1] || (s == bd[K - 1] && jj < bi[K - 1]))) {
        // insertion into the sorted list (static indices)
        float cd = s;
        int ci = jj;
#pragma unroll
        for (int p = 0; p < K; ++p) {
          const bool lt = cd < bd[p] || (cd == bd[p] && ci < bi[p]);
          const float td = bd[p];
          const int ti = bi[p];
          bd[p] = lt ? cd : td;
          bi[p] = lt ? ci : ti;
          cd = lt ? td : cd;
          ci = lt ? ti : ci;
        }
      }
    }
  }
  if (!valid) return;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    out_idx[(size_t)i * K + s] = bi[s] == 0x7fffffff ? -1 : bi[s];
    if (out_d2) out_d2[(size_t)i * K + s] = bd[s];
  }
}

struct KnnArgs {
  const float *nodes;
  int N, dim;
  const int32_t *queries;
  int Q;
  const float *after_d2;
  const int32_t *after_idx;
  int32_t *idx;
  float *d2;
};

template <int K>
void launch_knn(const KnnArgs &a, hipStream_t st) {
  const dim3 g((a.Q + 255) / 256), b(256);
  if (a.dim <= 8)
    hipLaunchKernelGGL((k_knn<K, 8>), g, b, 0, st, a.nodes, a.N, a.dim, a.queries, a.Q,
                       a.after_d2, a.after_idx, a.idx, a.d2);
  else
    hipLaunchKernelGGL((k_knn<K, 16>), g, b, 0, st, a.nodes, a.N, a.dim, a.queries, a.Q,
                       a.after_d2, a.after_idx, a.idx, a.d2);
}

template <int... Ks>
void launch_knn_k(int k, const KnnArgs &a, hipStream_t st, std::integer_sequence<int, Ks...>) {
  ((k == Ks + 1 ? launch_knn<Ks + 1>(a, st) : void()), ...);
}

namespace {
int fail(int code, const char *msg) {
  set_error(msg);
  return code;
}
}  // namespace

}  // namespace csd

using namespace csd;

extern "C" {

int csd_validate_edges(const csd_plant *plant, const float *nodes, int32_t num_nodes,
                       const int32_t *edges, int64_t num_edges, float step, int32_t rule,
                       uint8_t *out, void *stream) {
  if (!plant) return fail(CSD_ERR_INVALID_ARGUMENT, "null plant");
  const int dim = plant->img.dof;
  if (dim > kEdgeDim) return fail(CSD_ERR_CAPACITY, "edge validation supports up to 16 dims");
  if (plant->img.num_bodies > 16) return fail(CSD_ERR_CAPACITY, "more than 16 moving joints");
  if (num_edges < 0 || num_nodes < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative size");
  if (!(step > 0.f)) return fail(CSD_ERR_INVALID_ARGUMENT, "step must be positive");
  if (rule < 0 || rule > 2) return fail(CSD_ERR_INVALID_ARGUMENT, "rule must be 0, 1 or 2");
  if (num_edges == 0) return CSD_OK;
  if (!nodes || !edges || !out) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  const int var = check_variant(plant->img, false);
  // persistent grid over groups of 64 edges: enough waves to fill 256 CUs
  // several times over
  const int64_t waves = std::min<int64_t>((num_edges + 63) / 64, 256 * 32);
  dim3 g((unsigned)((waves * 64 + kEdgeBlock - 1) / kEdgeBlock));
  CSD_DISPATCH_VARIANT(var, k_validate_edges, g, dim3(kEdgeBlock),
                       check_lds_bytes(plant->img, var), (hipStream_t)stream, plant->img, nodes,
                       num_nodes, dim, edges, num_edges, step, rule, out);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

int csd_occupancy(const csd_plant *plant, const float *q, int64_t n, int64_t stride_n,
                  int64_t stride_j, const int32_t *robot_geoms, int32_t num_robot_geoms,
                  const float *map_center, const float *map_size, float resolution,
                  uint8_t *grid, void *stream) {
  if (!plant || !map_center || !map_size) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  if (plant->img.num_bodies > 16) return fail(CSD_ERR_CAPACITY, "more than 16 moving joints");
  if (n < 0 || !(resolution > 0.f)) return fail(CSD_ERR_INVALID_ARGUMENT, "bad size");
  if (n == 0 || num_robot_geoms == 0) return CSD_OK;
  if (!q || !grid) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  RobotGeomList R;
  VoxelArgs V{nullptr, 0, 0.f, 0, 0};
  int rc = CSD_OK;
  fill_robot_list(plant, robot_geoms, num_robot_geoms, R, V, &rc);
  if (rc) return rc;
  GridArgs G;
  for (int k = 0; k < 3; ++k) {
    G.center[k] = map_center[k];
    G.size[k] = map_size[k];
  }
  G.res = resolution;
  G.radius = (float)(sqrt(3.0) * resolution);  // GetOfflineVoxelRadius, roadmap.h:68
  G.nx = (int)ceilf(map_size[0] / (2 * resolution));
  G.ny = (int)ceilf(map_size[1] / (2 * resolution));
  G.nz = (int)ceilf(map_size[2] / (2 * resolution));
  dim3 g((unsigned)((n + kEdgeBlock - 1) / kEdgeBlock));
  if (plant->img.num_bodies <= 8)
    hipLaunchKernelGGL((k_occupancy<8>), g, dim3(kEdgeBlock), 0, (hipStream_t)stream,
                       plant->img, q, n, stride_n, stride_j, R, G, grid);
  else
    hipLaunchKernelGGL((k_occupancy<16>), g, dim3(kEdgeBlock), 0, (hipStream_t)stream,
                       plant->img, q, n, stride_n, stride_j, R, G, grid);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

int csd_voxel_collision_map(const csd_plant *plant, const float *nodes, int32_t num_nodes,
                            const int32_t *edges, int64_t num_edges, float step,
                            const int32_t *robot_geoms, int32_t num_robot_geoms,
                            const float *map_center, const float *map_size, float resolution,
                            int64_t *offsets, int32_t *items, int64_t capacity, int64_t *total,
                            void *stream) {
  if (!plant || !map_center || !map_size || !total || !offsets)
    return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  if (plant->img.num_bodies > 16) return fail(CSD_ERR_CAPACITY, "more than 16 moving joints");
  const int dim = plant->img.dof;
  if (dim > kEdgeDim) return fail(CSD_ERR_CAPACITY, "collision maps support up to 16 dims");
  if (num_nodes < 0 || num_edges < 0 || capacity < 0 || !(resolution > 0.f))
    return fail(CSD_ERR_INVALID_ARGUMENT, "bad size");
  if (edges && !(step > 0.f)) return fail(CSD_ERR_INVALID_ARGUMENT, "step must be positive");
  if (num_nodes > 0 && !nodes) return fail(CSD_ERR_INVALID_ARGUMENT, "null node array");
  if (edges == nullptr && num_edges > 0) return fail(CSD_ERR_INVALID_ARGUMENT, "null edge array");
  RobotGeomList R;
  VoxelArgs V{nullptr, 0, 0.f, 0, 0};
  int rc = CSD_OK;
  fill_robot_list(plant, robot_geoms, num_robot_geoms, R, V, &rc);
  if (rc) return rc;
  GridArgs G;
  for (int k = 0; k < 3; ++k) {
    G.center[k] = map_center[k];
    G.size[k] = map_size[k];
  }
  G.res = resolution;
  G.radius = (float)(sqrt(3.0) * resolution);  // GetOfflineVoxelRadius, roadmap.h:68
  G.nx = (int)ceilf(map_size[0] / (2 * resolution));
  G.ny = (int)ceilf(map_size[1] / (2 * resolution));
  G.nz = (int)ceilf(map_size[2] / (2 * resolution));
  const int64_t nvox = (int64_t)G.nx * G.ny * G.nz;
  if (G.nx < 1 || G.ny < 1 || G.nz < 1 || nvox > (int64_t)kMapMaxWords * 32)
    return fail(CSD_ERR_CAPACITY, "collision maps support grids of 1..524288 voxels");
  const int64_t nitems = edges ? num_edges : num_nodes;
  if (nitems > 0x7fffffff) return fail(CSD_ERR_CAPACITY, "too many items");
  hipStream_t st = (hipStream_t)stream;
  std::vector<int64_t> off((size_t)nvox + 1, 0);
  *total = 0;
  if (nitems == 0 || R.n == 0) {
    CSD_HIP(hipMemcpyAsync(offsets, off.data(), off.size() * 8, hipMemcpyHostToDevice, st));
    CSD_HIP(hipStreamSynchronize(st));
    return CSD_OK;
  }
  int32_t *counts = nullptr;
  CSD_HIP(hipMalloc(&counts, (size_t)nvox * 4));
  std::vector<int32_t> cnt((size_t)nvox, 0);
  const size_t lds = (size_t)((nvox + 31) / 32) * 4;
  const dim3 g((unsigned)nitems), b(kMapBlock);
  hipError_t e = hipMemcpyAsync(counts, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice, st);
  auto launch = [&](int pass) {
    if (plant->img.num_bodies <= 8)
      hipLaunchKernelGGL((k_voxel_map<8>), g, b, lds, st, plant->img, nodes, num_nodes, dim, edges,
                         step, R, G, pass, counts, offsets, counts, items);
    else
      hipLaunchKernelGGL((k_voxel_map<16>), g, b, lds, st, plant->img, nodes, num_nodes, dim,
                         edges, step, R, G, pass, counts, offsets, counts, items);
    return hipGetLastError();
  };
  if (e == hipSuccess) e = launch(0);
  if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), counts, cnt.size() * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) {
    for (int64_t v = 0; v < nvox; ++v) off[v + 1] = off[v] + cnt[v];
    *total = off[nvox];
    e = hipMemcpyAsync(offsets, off.data(), off.size() * 8, hipMemcpyHostToDevice, st);
  }
  bool fits = items != nullptr && capacity >= *total;
  if (e == hipSuccess && fits) {
    std::fill(cnt.begin(), cnt.end(), 0);  // the counts become the per-voxel cursors
    e = hipMemcpyAsync(counts, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch(1);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(counts);
  if (e != hipSuccess) return hip_fail(e, "voxel collision map");
  if (items != nullptr && !fits) return fail(CSD_ERR_CAPACITY, "item buffer smaller than *total");
  return CSD_OK;
}

int csd_knn(const float *nodes, int32_t num_nodes, int32_t dim, int32_t k, int32_t *out_idx,
            float *out_d2, void *stream) {
  if (dim < 1 || dim > kEdgeDim) return fail(CSD_ERR_CAPACITY, "k-NN supports 1..16 dims");
  if (k < 1 || k > kKnnMax) return fail(CSD_ERR_CAPACITY, "k must be in 1..16");
  if (num_nodes < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative size");
  if (num_nodes == 0) return CSD_OK;
  if (!nodes || !out_idx) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  const KnnArgs a{nodes, num_nodes, dim, nullptr, num_nodes, nullptr, nullptr, out_idx, out_d2};
  launch_knn_k(k, a, (hipStream_t)stream, std::make_integer_sequence<int, kKnnMax>{});
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

int csd_knn_page(const float *nodes, int32_t num_nodes, int32_t dim, const int32_t *queries,
                 int32_t num_queries, int32_t k, const float *after_d2,
                 const int32_t *after_idx, int32_t *out_idx, float *out_d2, void *stream) {
  if (dim < 1 || dim > kEdgeDim) return fail(CSD_ERR_CAPACITY, "k-NN supports 1..16 dims");
  if (k < 1 || k > kKnnMax) return fail(CSD_ERR_CAPACITY, "k must be in 1..16");
  if (num_nodes < 0 || num_queries < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative size");
  if (num_queries == 0) return CSD_OK;
  if (!nodes || !queries || !out_idx || !out_d2 || (after_d2 == nullptr) != (after_idx == nullptr))
    return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  const KnnArgs a{nodes, num_nodes, dim, queries, num_queries, after_d2, after_idx, out_idx,
                  out_d2};
  launch_knn_k(k, a, (hipStream_t)stream, std::make_integer_sequence<int, kKnnMax>{});
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

}  // extern "C"
