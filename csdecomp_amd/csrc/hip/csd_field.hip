// csd_field.hip -- exact Euclidean distance field of a voxel occupancy grid
// (the reference's DistanceField, geometry/cuda_distance_field.h:133-172,
// which runs the PBA Voronoi algorithm). Here the squared distance to the
// nearest occupied voxel is computed by three separable exact passes:
//  1. along x: one wave per row; forward and backward wave scans give the
//     nearest occupied voxel of the row on either side;
//  2. along y, then 3. along z: one lane per column (consecutive lanes take
//     consecutive x, so every access of a wave is one coalesced row), each
//     building the lower envelope of the parabolas f(q) + (u - q)^2 with its
//     stack kept in place (prev[q] is stored at site q of a scratch grid).
// All distances are integers in voxel units, so the result is exact.
// Memory: the uint8 occupancy grid and three int32 grids of G^3.
#include <string.h>

#include <algorithm>

#include "csd_internal.h"

namespace csd {
namespace {

constexpr int32_t kFar = 0x3fffffff;  // no site (larger than any 3 * 1023^2)
constexpr int kFieldBlock = 256;

int fail(int code, const char *msg) {
  set_error(msg);
  return code;
}

__global__ void k_fill_u8(uint8_t *__restrict__ p, int64_t n, uint8_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// voxel of a point: round((p - origin) / voxel) per axis (C round, halves
// away from zero), as OccupancyGrid::setOccupied / VoronoiGrid::getDistance
// (cuda_distance_field.h:76-131)
__device__ __forceinline__ bool voxel_of(const float *p, float ox, float oy, float oz, float vs,
                                         int G, int &x, int &y, int &z) {
  x = (int)roundf((p[0] - ox) / vs);
  y = (int)roundf((p[1] - oy) / vs);
  z = (int)roundf((p[2] - oz) / vs);
  return x >= 0 && y >= 0 && z >= 0 && x < G && y < G && z < G;
}

__global__ void k_set_occupancy(uint8_t *__restrict__ occ, const float *__restrict__ pts,
                                int64_t m, float ox, float oy, float oz, float vs, int G) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int x, y, z;
  if (voxel_of(pts + 3 * i, ox, oy, oz, vs, G, x, y, z))
    occ[((int64_t)z * G + y) * G + x] = 1;
}

// pass 1: one wave per row (y, z); lane l owns the L = ceil(G / 64) <= 16
// consecutive voxels [l L, l L + L). Each lane scans its segment, one wave
// scan of the segment summaries (last site so far from the left, first
// site from the right) supplies the carries, and out = min(x - last,
// next - x)^2 or kFar.
__global__ __launch_bounds__(kFieldBlock) void k_edt_rows(const uint8_t *__restrict__ occ,
                                                          int32_t *__restrict__ out, int G) {
  constexpr int kMaxSeg = 16;
  constexpr int kNone = 1 << 20;
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (kFieldBlock / 64) + (threadIdx.x >> 6);
  if (row >= (int64_t)G * G) return;
  const int L = (G + 63) / 64;
  const int x0 = lane * L;
  const uint8_t *o = occ + row * G;
  int32_t *d = out + row * G;
  bool site[kMaxSeg];
  int seg_last = -kNone, seg_first = kNone;
#pragma unroll
  for (int i = 0; i < kMaxSeg; ++i) {
    const int x = x0 + i;
    site[i] = i < L && x < G && o[x] != 0;
    if (site[i]) {
      seg_last = x;
      seg_first = min(seg_first, x);
    }
  }
  // exclusive scans over lanes: last site left of the segment, first site
  // right of it
  int left = seg_last, right = seg_first;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const int vl = __shfl_up(left, s);
    const int vr = __shfl_down(right, s);
    if (lane >= s) left = max(left, vl);
    if (lane + s < 64) right = min(right, vr);
  }
  int carry_l = __shfl_up(left, 1);
  int carry_r = __shfl_down(right, 1);
  if (lane == 0) carry_l = -kNone;
  if (lane == 63) carry_r = kNone;
  int lastv[kMaxSeg];
  int last = carry_l;
#pragma unroll
  for (int i = 0; i < kMaxSeg; ++i) {
    if (site[i]) last = x0 + i;
    lastv[i] = last;
  }
  int next = carry_r;
#pragma unroll
  for (int i = kMaxSeg - 1; i >= 0; --i) {
    const int x = x0 + i;
    if (site[i]) next = x;
    if (i < L && x < G) {
      const int dm = min(x - lastv[i], next - x);
      d[x] = dm < (1 << 19) ? dm * dm : kFar;
    }
  }
}

// passes 2 and 3: column c = (a, b) with a fastest; element u of the column
// at f[a + b * sb + u * su]. Lower envelope of p_q(u) = f[q] + (u - q)^2
// over the sites q (f[q] < kFar); prev[q] links the envelope's stack.
__global__ __launch_bounds__(kFieldBlock) void k_edt_columns(const int32_t *__restrict__ f,
                                                             int32_t *__restrict__ prev,
                                                             int32_t *__restrict__ out, int G,
                                                             int64_t sb, int64_t su) {
  const int64_t c = (int64_t)blockIdx.x * kFieldBlock + threadIdx.x;
  if (c >= (int64_t)G * G) return;
  const int64_t base = (c % G) + (c / G) * sb;
  // the top of the stack, the site below it and their values stay in
  // registers; memory holds the rest of the stack (prev) and f
  int top = -1, below = -1;
  int64_t ftop = 0;
  constexpr int kAhead = 16;  // loads issued together
  for (int q0 = 0; q0 < G; q0 += kAhead) {
    int32_t fa[kAhead];
#pragma unroll
    for (int j = 0; j < kAhead; ++j) fa[j] = q0 + j < G ? f[base + (int64_t)(q0 + j) * su] : kFar;
#pragma unroll
    for (int j = 0; j < kAhead; ++j) {
    const int q = q0 + j;
    const int32_t fq = fa[j];
    if (fq >= kFar) continue;
    // pop top while q's parabola overtakes it no later than it overtook the
    // one below: ((fq + q^2) - (ft + t^2)) (t - p) <= ((ft + t^2) - (fp + p^2)) (q - t)
    while (below >= 0) {
      const int t = top, p = below;
      const int64_t fp = f[base + p * su];
      const int64_t lhs = ((int64_t)fq + (int64_t)q * q - ftop - (int64_t)t * t) * (t - p);
      const int64_t rhs = (ftop + (int64_t)t * t - fp - (int64_t)p * p) * (q - t);
      if (lhs > rhs) break;
      top = p;
      ftop = fp;
      below = prev[base + p * su];
    }
    // a single parabola on the stack stays: two distinct sites never have
    // one parabola below the other everywhere
    prev[base + q * su] = top;
    below = top;
    top = q;
    ftop = fq;
    }
  }
  // walk the envelope from the end: u belongs to the top parabola until the
  // one below it is lower
  int64_t fbelow = below >= 0 ? f[base + below * su] : 0;
  for (int u = G - 1; u >= 0; --u) {
    int32_t v = kFar;
    if (top >= 0) {
      int64_t vt = ftop + (int64_t)(u - top) * (u - top);
      while (below >= 0) {
        const int64_t vb = fbelow + (int64_t)(u - below) * (u - below);
        if (vb > vt) break;
        top = below;
        ftop = fbelow;
        vt = vb;
        below = prev[base + top * su];
        fbelow = below >= 0 ? f[base + below * su] : 0;
      }
      v = vt < kFar ? (int32_t)vt : kFar;
    }
    out[base + u * su] = v;
  }
}

__global__ void k_field_distance(const int32_t *__restrict__ d2, const float *__restrict__ pts,
                                 int64_t m, float ox, float oy, float oz, float vs, int G,
                                 float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int x, y, z;
  float r = 0.f;  // outside the grid
  if (voxel_of(pts + 3 * i, ox, oy, oz, vs, G, x, y, z)) {
    const int32_t v = d2[((int64_t)z * G + y) * G + x];
    r = v >= kFar ? INFINITY : sqrtf((float)v) * vs;
  }
  out[i] = r;
}

}  // namespace
}  // namespace csd

struct csd_distance_field {
  float origin[3];
  float voxel_size;
  int32_t G;
  uint8_t *occ = nullptr;
  int32_t *a = nullptr, *b = nullptr, *c = nullptr;  // pass buffers; result in a
  bool computed = false;
};

using namespace csd;

extern "C" {

int csd_distance_field_create(const float *origin, int32_t grid_size, float voxel_size,
                              csd_distance_field **out) {
  if (!origin || !out) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  if (grid_size < 1 || grid_size > 1024)
    return fail(CSD_ERR_CAPACITY, "grid size must be 1..1024");
  if (!(voxel_size > 0.f)) return fail(CSD_ERR_INVALID_ARGUMENT, "voxel size must be positive");
  auto *f = new csd_distance_field();
  memcpy(f->origin, origin, sizeof(f->origin));
  f->voxel_size = voxel_size;
  f->G = grid_size;
  const size_t n = (size_t)grid_size * grid_size * grid_size;
  hipError_t e = hipMalloc(&f->occ, n);
  if (e == hipSuccess) e = hipMalloc(&f->a, n * 4);
  if (e == hipSuccess) e = hipMalloc(&f->b, n * 4);
  if (e == hipSuccess) e = hipMalloc(&f->c, n * 4);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr,
                       f->occ, (int64_t)n, (uint8_t)0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    (void)hipFree(f->occ);
    (void)hipFree(f->a);
    (void)hipFree(f->b);
    (void)hipFree(f->c);
    delete f;
    return e == hipErrorOutOfMemory ? fail(CSD_ERR_OUT_OF_MEMORY, "distance field allocation")
                                    : hip_fail(e, "distance field allocation");
  }
  *out = f;
  return CSD_OK;
}

int csd_distance_field_destroy(csd_distance_field *f) {
  if (!f) return CSD_OK;
  (void)hipFree(f->occ);
  (void)hipFree(f->a);
  (void)hipFree(f->b);
  (void)hipFree(f->c);
  delete f;
  return CSD_OK;
}

int csd_distance_field_reset_occupancy(csd_distance_field *f, void *stream) {
  if (!f) return fail(CSD_ERR_INVALID_ARGUMENT, "null field");
  const int64_t n = (int64_t)f->G * f->G * f->G;
  hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, f->occ, n, (uint8_t)0);
  CSD_HIP(hipGetLastError());
  f->computed = false;
  return CSD_OK;
}

int csd_distance_field_set_occupancy(csd_distance_field *f, const float *points, int64_t m,
                                     void *stream) {
  if (!f) return fail(CSD_ERR_INVALID_ARGUMENT, "null field");
  if (m < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative point count");
  if (m == 0) return CSD_OK;
  if (!points) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  hipLaunchKernelGGL(k_set_occupancy, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, f->occ, points, m, f->origin[0], f->origin[1],
                     f->origin[2], f->voxel_size, f->G);
  CSD_HIP(hipGetLastError());
  f->computed = false;
  return CSD_OK;
}

int csd_distance_field_compute(csd_distance_field *f, void *stream) {
  if (!f) return fail(CSD_ERR_INVALID_ARGUMENT, "null field");
  hipStream_t st = (hipStream_t)stream;
  const int G = f->G;
  const int64_t rows = (int64_t)G * G;
  hipLaunchKernelGGL(k_edt_rows, dim3((unsigned)((rows + 3) / 4)), dim3(kFieldBlock), 0, st,
                     f->occ, f->c, G);
  const dim3 cg((unsigned)((rows + kFieldBlock - 1) / kFieldBlock));
  // along y: columns (x, z), element stride G
  hipLaunchKernelGGL(k_edt_columns, cg, dim3(kFieldBlock), 0, st, f->c, f->b, f->a, G,
                     (int64_t)G * G, (int64_t)G);
  // along z: columns (x, y), element stride G^2
  hipLaunchKernelGGL(k_edt_columns, cg, dim3(kFieldBlock), 0, st, f->a, f->b, f->c, G,
                     (int64_t)G, (int64_t)G * G);
  CSD_HIP(hipGetLastError());
  std::swap(f->a, f->c);
  f->computed = true;
  return CSD_OK;
}

int csd_distance_field_get_distance(const csd_distance_field *f, const float *points, int64_t m,
                                    float *out, void *stream) {
  if (!f) return fail(CSD_ERR_INVALID_ARGUMENT, "null field");
  if (!f->computed) return fail(CSD_ERR_INVALID_ARGUMENT, "distance field not computed");
  if (m < 0) return fail(CSD_ERR_INVALID_ARGUMENT, "negative point count");
  if (m == 0) return CSD_OK;
  if (!points || !out) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  hipLaunchKernelGGL(k_field_distance, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, f->a, points, m, f->origin[0], f->origin[1],
                     f->origin[2], f->voxel_size, f->G, out);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

int csd_distance_field_squared(const csd_distance_field *f, const int32_t **d2) {
  if (!f || !d2) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  if (!f->computed) return fail(CSD_ERR_INVALID_ARGUMENT, "distance field not computed");
  *d2 = f->a;
  return CSD_OK;
}

}  // extern "C"
