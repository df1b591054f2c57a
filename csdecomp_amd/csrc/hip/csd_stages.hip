// csd_stages.hip -- EI-ZO stages as standalone entry points: hit-and-run
// sampling in many polytopes (UniformSampleInHPolyhedronCuda), projection +
// bisection of given particles, and greedy face placement for one region.
// The batched driver in csd_eizo.hip runs the same device code fused per
// round.
#include <algorithm>
#include <mutex>

#include "csd_bisect.h"

namespace csd {

// standalone sampler: polytope p = blockIdx.y. Without `ends` every chain
// of polytope p starts at starts[p] and draws random stream p; with `ends`
// chain i starts at fraction (i + 1/2) / n of the segment starts[p]-ends[p]
// and draws stream p + stream_off, as the EI-ZO round's sampler does.
template <int DPL, bool BROW = false>
__global__ __launch_bounds__(kSampleBlock) void k_hit_and_run(
    const float *__restrict__ A, const float *__restrict__ b, const int32_t *__restrict__ nf,
    int face_stride, int dim, const float *__restrict__ starts, const float *__restrict__ ends,
    uint32_t stream_off, int64_t n, int mixing, uint32_t k0, uint32_t k1, uint32_t round,
    float *__restrict__ out) {
  constexpr int DP = SamplerShape<DPL>::DP;
  constexpr bool LDS = SamplerShape<DPL>::LDS;
  extern __shared__ float csd_smem[];
  const int p = blockIdx.y;
  const int F = nf[p];
  const float *Ap = A + (size_t)p * face_stride * dim, *bp = b + (size_t)p * face_stride;
  if constexpr (LDS) {
    stage_polytope<DP, BROW>(csd_smem, csd_smem + (size_t)F * DP, Ap, bp, F, dim);
    __syncthreads();
    Ap = csd_smem;
    bp = csd_smem + (size_t)F * DP;
  }
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[DP];
  if (ends) {
    const float f = ((float)i + 0.5f) / (float)n;
#pragma unroll
    for (int d = 0; d < DP; ++d) {
      const float a = d < dim ? starts[(size_t)p * dim + d] : 0.f;
      const float e = d < dim ? ends[(size_t)p * dim + d] : 0.f;
      x[d] = fmaf(f, e - a, a);
    }
  } else {
#pragma unroll
    for (int d = 0; d < DP; ++d) x[d] = d < dim ? starts[(size_t)p * dim + d] : 0.f;
  }
  hit_and_run<DP, LDS, BROW>(Ap, bp, F, dim, x, (uint32_t)i,
                             (uint32_t)p + (ends ? stream_off : 0u), round, mixing, k0, k1);
#pragma unroll
  for (int d = 0; d < DP; ++d)
    if (d < dim) out[((size_t)p * n + i) * dim + d] = x[d];
}

__global__ void k_face_prep(const float *__restrict__ scol, const float *__restrict__ proj,
                            int n, int dim, double *__restrict__ dist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double e = (double)scol[(size_t)i * dim + d] - (double)proj[(size_t)i * dim + d];
    s += e * e;
  }
  dist[i] = sqrt(s);
}

// Region editing (planning/cuda_edit_regions.cu:236-293), one workgroup per
// region, FP64 arithmetic on the stored FP32 faces. The region's collisions
// are [off[r], off[r+1]):
//  1. keep those whose optimised point is strictly inside the region, keyed by
//     |scol - proj|;
//  2. repeatedly take the smallest key (lower index on ties), add its face, and
//     drop every remaining point the new face leaves outside (a.x - b >= 0),
//     so each point is tested against the region as edited so far.
constexpr int kEditBlock = 256;  // 4 waves: four regions share a CU (32 KB of keys each)
constexpr int kEditWaves = kEditBlock / 64;
constexpr int kEditMax = 4096;  // collisions per region (LDS keys)

__global__ __launch_bounds__(kEditBlock) void k_edit_faces(
    const float *__restrict__ scol, const float *__restrict__ proj,
    const int32_t *__restrict__ off, int C, const float *__restrict__ ls,
    const float *__restrict__ le, float *__restrict__ A, float *__restrict__ b,
    int32_t *__restrict__ nf, int cap, int dim, float margin, int32_t *__restrict__ status) {
  __shared__ double s_key[kEditMax];
  __shared__ double s_v[kEditWaves];
  __shared__ int s_i[kEditWaves];
  __shared__ double s_a[kDimMax];
  __shared__ double s_b;
  __shared__ int s_stop, s_nf;
  const int r = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int first = off[r], last = off[r + 1];
  if (first < 0 || last < first || last > C) {  // offsets outside the batch
    if (tid == 0) status[r] = -4;
    return;
  }
  const int total = last - first;
  if (total > kEditMax) {
    if (tid == 0) status[r] = -1;
    return;
  }
  if (total == 0) {
    if (tid == 0) status[r] = 0;
    return;
  }
  const int F0 = nf[r];
  if (F0 < 0 || F0 > cap) {  // a face count the buffers cannot hold
    if (tid == 0) status[r] = -3;
    return;
  }
  float *Ar = A + (size_t)r * cap * dim;
  float *br = b + (size_t)r * cap;
  scol += (size_t)first * dim;
  proj += (size_t)first * dim;
  // 1. keys: distance for points strictly inside the region, +inf otherwise
  for (int k = tid; k < total; k += kEditBlock) {
    float o[kDimMax];
    double dd = 0.0;
#pragma unroll
    for (int d = 0; d < kDimMax; ++d) {
      o[d] = d < dim ? scol[(size_t)k * dim + d] : 0.f;
      const double e = d < dim ? (double)o[d] - (double)proj[(size_t)k * dim + d] : 0.0;
      dd += e * e;
    }
    bool inside = true;
    for (int f = 0; f < F0 && inside; ++f) {
      double s = -(double)br[f];
#pragma unroll
      for (int d = 0; d < kDimMax; ++d)
        if (d < dim) s += (double)Ar[(size_t)f * dim + d] * (double)o[d];
      inside = s < 0.0;
    }
    s_key[k] = inside ? sqrt(dd) : INFINITY;
  }
  if (tid == 0) s_nf = F0;
  __syncthreads();
  // 2. greedy placement
  int added = 0, full = 0;
  for (;;) {
    double v = INFINITY;
    int vi = 0x7fffffff;
    for (int k = tid; k < total; k += kEditBlock) argmin_pair(v, vi, s_key[k], k);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(v, o);
      const int oi = __shfl_xor(vi, o);
      argmin_pair(v, vi, ov, oi);
    }
    if (lane == 0) {
      s_v[wave] = v;
      s_i[wave] = vi;
    }
    __syncthreads();
    if (tid == 0) {
      double bv = s_v[0];
      int bi = s_i[0];
      for (int w = 1; w < kEditWaves; ++w) argmin_pair(bv, bi, s_v[w], s_i[w]);
      s_stop = 0;
      if (!(bv < INFINITY)) {
        s_stop = 1;  // no contained point left
      } else if (s_nf >= cap) {
        s_stop = 2;  // face capacity exhausted
      } else {
        const float *sc = scol + (size_t)bi * dim, *pr = proj + (size_t)bi * dim;
        const float *q1 = ls + (size_t)r * dim, *q2 = le + (size_t)r * dim;
        double a[kDimMax], nrm = 0.0;
        for (int d = 0; d < dim; ++d) {
          a[d] = (double)sc[d] - (double)pr[d];
          nrm += a[d] * a[d];
        }
        nrm = sqrt(nrm);
        const int f = s_nf;
        double bb = -(double)margin, v1 = 0.0, v2 = 0.0;
        for (int d = 0; d < dim; ++d) {
          // a zero direction stays zero (Eigen's normalize leaves it alone)
          const float af = nrm > 0.0 ? (float)(a[d] / nrm) : 0.f;
          Ar[(size_t)f * dim + d] = af;
          s_a[d] = af;
          bb += (double)af * sc[d];
          v1 += (double)af * q1[d];
          v2 += (double)af * q2[d];
        }
        const double relax = fmax(v1 - bb, v2 - bb);
        if (relax > 0.0) bb += relax + 2e-6;  // keep the segment inside the face
        const float bf = __double2float_ru(bb);
        br[f] = bf;
        s_b = bf;
        s_nf = f + 1;
        s_key[bi] = INFINITY;
      }
    }
    __syncthreads();
    if (s_stop) {
      full = s_stop == 2;
      break;
    }
    ++added;
    for (int k = tid; k < total; k += kEditBlock) {
      if (!(s_key[k] < INFINITY)) continue;
      double s = -s_b;
      for (int d = 0; d < dim; ++d) s += s_a[d] * (double)scol[(size_t)k * dim + d];
      if (!(s < 0.0)) s_key[k] = INFINITY;
    }
    __syncthreads();
  }
  if (tid == 0) {
    nf[r] = s_nf;
    status[r] = full ? -2 : added;
  }
}

// fused projection + bisection for a batch of particles (shared by csd_bisect
// and csd_edit_regions); arguments validated by the callers
int launch_bisect(const csd_plant *plant, const float *particles, const float *line_start,
                  const float *line_end, const uint32_t *segment_idx, int num_particles,
                  int steps, const float *voxels, int num_voxels, float voxel_radius,
                  const int32_t *robot_geoms, int num_robot_geoms, float *scol, float *proj,
                  hipStream_t st) {
  const int dim = plant->img.dof;
  RobotGeomList R;
  VoxelArgs V{voxels, num_voxels, voxel_radius, 0, 0};
  int rc = CSD_OK;
  bool all_slots = fill_robot_list(plant, robot_geoms, num_voxels ? num_robot_geoms : 0, R, V, &rc);
  if (rc) return rc;
  const int var = check_variant(plant->img, num_voxels > 0 && !all_slots);
  // waves per 64 particles: small batches spread each particle's pairs over
  // several waves so that the chip holds enough waves to hide latency
  const int64_t waves = ((int64_t)num_particles + 63) / 64;
  const bool generic = var == kGeneric8 || var == kGeneric16;  // no split of its pairs
  const int split = generic || waves >= 4096 ? 1 : waves >= 2048 ? 2 : 4;
  const size_t lds = check_lds_bytes(plant->img, var);
  const dim3 g((unsigned)((num_particles + kSampleBlock / split - 1) / (kSampleBlock / split)));
  const dim3 blk(kSampleBlock);
  const BisectLaunch L{plant->img, V, R, particles, line_start, line_end, segment_idx,
                      num_particles, dim, steps, scol, proj, g, blk, lds, st};
  if (dim <= 8) launch_bisect_dp8(L, var, split);
  else launch_bisect_dp16(L, var, split);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}
// csd_place_faces_noalloc: its scratch (the four counters of the one-region
// EizoDev, then the FP64 distance of every particle) lives in this
// file-scope device array instead of per-call allocations. Each GPU holds
// its own copy for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns on that device.
constexpr int kNoAllocMaxParticles = 65536;
constexpr size_t kNoAllocCounterBytes = 256;
constexpr size_t kNoAllocBytes = kNoAllocCounterBytes + (size_t)kNoAllocMaxParticles * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// counters: u = n, status = 0, faces added = 0
__global__ void k_noalloc_init(int32_t *ctr, int n) {
  if (threadIdx.x == 0) {
    ctr[0] = n;
    ctr[1] = 0;
    ctr[2] = 0;
    ctr[3] = 0;
  }
}

__global__ void k_noalloc_finish(const int32_t *ctr, int32_t *added, int32_t *status) {
  if (threadIdx.x == 0) {
    *added = ctr[2];
    if (status) *status = ctr[1];
  }
}

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int noalloc_scratch(void **out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = 0;
  CSD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(CSD_ERR_CAPACITY, "device index above 63");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress (face placement scratch)");
    addr[dev] = p;
  }
  *out = addr[dev];
  return CSD_OK;
}

}  // namespace csd

using namespace csd;

extern "C" {

namespace {
int launch_hit_and_run(const float *A, const float *b, const int32_t *num_faces,
                       int32_t face_stride, int32_t dim, int32_t num_polytopes,
                       const float *starts, const float *ends, uint32_t stream_offset,
                       int64_t num_samples, int32_t mixing_steps, uint64_t seed, uint32_t round,
                       float *out, void *stream) {
  if (dim < 1 || dim > kDimMax) return fail(CSD_ERR_CAPACITY, "sampler supports 1..16 dims");
  if (num_polytopes < 0 || num_samples < 0 || mixing_steps < 0 || face_stride < 0)
    return fail(CSD_ERR_INVALID_ARGUMENT, "negative size");
  if (num_polytopes == 0 || num_samples == 0) return CSD_OK;
  if (!A || !b || !num_faces || !starts || !out)
    return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  if (num_samples > 0x7fffffffLL) return fail(CSD_ERR_CAPACITY, "too many samples per polytope");
  dim3 g((unsigned)((num_samples + kSampleBlock - 1) / kSampleBlock), (unsigned)num_polytopes);
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  hipStream_t st = (hipStream_t)stream;
  const dim3 blk(kSampleBlock);
  // polytopes of up to face_stride rows go to LDS when they fit, with b in
  // the rows' padding slot when there is one (the EI-ZO sampler's shapes)
#define CSD_HNR(DPL, BROW, LDS)                                                                 \
  hipLaunchKernelGGL((k_hit_and_run<DPL, BROW>), g, blk, LDS, st, A, b, num_faces, face_stride, \
                     dim, starts, ends, stream_offset, num_samples, mixing_steps, k0, k1, round, \
                     out)
  const size_t l8 = polytope_lds_bytes<8>(face_stride), l16 = polytope_lds_bytes<16>(face_stride);
  if (dim <= 7 && l8 <= kLdsLimit) CSD_HNR(17, true, l8);
  else if (dim <= 8 && l8 <= kLdsLimit) CSD_HNR(17, false, l8);
  else if (dim <= 15 && l16 <= kLdsLimit) CSD_HNR(33, true, l16);
  else if (l16 <= kLdsLimit) CSD_HNR(33, false, l16);
  else CSD_HNR(32, false, 0);
#undef CSD_HNR
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}
}  // namespace

int csd_hit_and_run(const float *A, const float *b, const int32_t *num_faces, int32_t face_stride,
                    int32_t dim, int32_t num_polytopes, const float *starts, int64_t num_samples,
                    int32_t mixing_steps, uint64_t seed, uint32_t round, float *out,
                    void *stream) {
  return launch_hit_and_run(A, b, num_faces, face_stride, dim, num_polytopes, starts, nullptr, 0u,
                            num_samples, mixing_steps, seed, round, out, stream);
}

int csd_hit_and_run_segments(const float *A, const float *b, const int32_t *num_faces,
                             int32_t face_stride, int32_t dim, int32_t num_polytopes,
                             const float *segment_start, const float *segment_end,
                             int64_t num_samples, int32_t mixing_steps, uint64_t seed,
                             uint32_t stream_offset, uint32_t round, float *out, void *stream) {
  if (!segment_end) return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  return launch_hit_and_run(A, b, num_faces, face_stride, dim, num_polytopes, segment_start,
                            segment_end, stream_offset, num_samples, mixing_steps, seed, round,
                            out, stream);
}

int csd_bisect(const csd_plant *plant, const float *particles, const float *line_start,
               const float *line_end, const uint32_t *segment_idx, int32_t num_particles,
               int32_t steps, const float *voxels, int32_t num_voxels, float voxel_radius,
               const int32_t *robot_geoms, int32_t num_robot_geoms, float *scol, float *proj,
               void *stream) {
  if (!plant) return fail(CSD_ERR_INVALID_ARGUMENT, "null plant");
  const int dim = plant->img.dof;
  if (dim < 1 || dim > kDimMax) return fail(CSD_ERR_CAPACITY, "bisection supports 1..16 dims");
  if (plant->img.num_bodies > 16) return fail(CSD_ERR_CAPACITY, "more than 16 moving joints");
  if (num_particles < 0 || steps < 0 || num_voxels < 0)
    return fail(CSD_ERR_INVALID_ARGUMENT, "negative size");
  if (num_particles == 0) return CSD_OK;
  if (!particles || !line_start || !line_end || !segment_idx || !scol || !proj ||
      (num_voxels && !voxels))
    return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  return launch_bisect(plant, particles, line_start, line_end, segment_idx, num_particles, steps,
                       voxels, num_voxels, voxel_radius, robot_geoms, num_robot_geoms, scol, proj,
                       (hipStream_t)stream);
}

int csd_edit_regions(const csd_plant *plant, const float *collisions, const uint32_t *segment_idx,
                     const int32_t *region_offsets, int32_t num_collisions,
                     const float *line_start, const float *line_end,
                     int32_t num_regions, float *A, float *b, int32_t *num_faces,
                     int32_t capacity, const float *voxels, int32_t num_voxels,
                     float voxel_radius, const int32_t *robot_geoms, int32_t num_robot_geoms,
                     float margin, int32_t bisection_steps, float *scol, float *proj,
                     int32_t *status, void *stream) {
  if (!plant) return fail(CSD_ERR_INVALID_ARGUMENT, "null plant");
  const int dim = plant->img.dof;
  if (dim < 1 || dim > kDimMax) return fail(CSD_ERR_CAPACITY, "region editing supports 1..16 dims");
  if (plant->img.num_bodies > 16) return fail(CSD_ERR_CAPACITY, "more than 16 moving joints");
  if (num_collisions < 0 || num_regions < 0 || bisection_steps < 0 || num_voxels < 0 ||
      capacity < 1)
    return fail(CSD_ERR_INVALID_ARGUMENT, "bad size");
  if (num_regions == 0) return CSD_OK;
  if (!line_start || !line_end || !A || !b || !num_faces || !status || !region_offsets ||
      (num_collisions && (!collisions || !segment_idx || !scol || !proj)) ||
      (num_voxels && !voxels))
    return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  hipStream_t st = (hipStream_t)stream;
  if (num_collisions) {
    const int rc = launch_bisect(plant, collisions, line_start, line_end, segment_idx,
                                 num_collisions, bisection_steps, voxels, num_voxels,
                                 voxel_radius, robot_geoms, num_robot_geoms, scol, proj, st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_edit_faces, dim3(num_regions), dim3(kEditBlock), 0, st, scol, proj,
                     region_offsets, num_collisions, line_start, line_end, A, b, num_faces,
                     capacity, dim, margin, status);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

int csd_place_faces(float *A, float *b, int32_t *num_faces, int32_t capacity, int32_t dim,
                    const float *scol, const float *proj, int32_t num_particles,
                    const float *line_start, const float *line_end, float margin,
                    int32_t max_faces, int32_t *added, void *stream) {
  if (dim < 1 || dim > kDimMax) return fail(CSD_ERR_CAPACITY, "face placement supports 1..16 dims");
  if (num_particles < 0 || capacity < 1 || max_faces < 0)
    return fail(CSD_ERR_INVALID_ARGUMENT, "bad size");
  if (!A || !b || !num_faces || !line_start || !line_end || !added ||
      (num_particles && (!scol || !proj)))
    return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  hipStream_t st = (hipStream_t)stream;
  // a one-region EizoDev over the caller's buffers plus small temporaries
  int32_t *tmp = nullptr;
  double *dist = nullptr;
  CSD_HIP(hipMalloc(&tmp, 4 * sizeof(int32_t)));
  hipError_t e = hipMalloc(&dist, std::max(num_particles, 1) * sizeof(double));
  if (e != hipSuccess) {
    hipFree(tmp);
    return hip_fail(e, "hipMalloc");
  }
  int32_t init[4] = {num_particles, 0, 0, 0};  // u, status, faces_added, spare
  e = hipMemcpyAsync(tmp, init, sizeof(init), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && num_particles)
    hipLaunchKernelGGL(k_face_prep, dim3((num_particles + 255) / 256), dim3(256), 0, st, scol,
                       proj, num_particles, dim, dist);
  EizoDev D;
  memset(&D, 0, sizeof(D));
  D.A = A;
  D.b = b;
  D.nf = num_faces;
  D.ls = const_cast<float *>(line_start);
  D.le = const_cast<float *>(line_end);
  D.scol = const_cast<float *>(scol);
  D.proj = const_cast<float *>(proj);
  D.dist = dist;
  D.u = tmp;
  D.status = tmp + 1;
  D.faces_added = tmp + 2;
  if (e == hipSuccess)
    hipLaunchKernelGGL(k_eizo_faces, dim3(1), dim3(kFaceBlock), 0, st, D, dim,
                       std::max(num_particles, 1), capacity, margin, max_faces, 0, 0u);
  if (e == hipSuccess) e = hipGetLastError();
  int32_t res[4] = {0, 0, 0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(res, tmp, sizeof(res), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipMemcpy(added, tmp + 2, 4, hipMemcpyDeviceToDevice);
  hipFree(tmp);
  hipFree(dist);
  if (e != hipSuccess) return hip_fail(e, "face placement");
  if (res[1] == 2) return fail(CSD_ERR_CAPACITY, "face capacity exhausted");
  return CSD_OK;
}

int csd_place_faces_noalloc(float *A, float *b, int32_t *num_faces, int32_t capacity,
                            int32_t dim, const float *scol, const float *proj,
                            int32_t num_particles, const float *line_start,
                            const float *line_end, float margin, int32_t max_faces,
                            int32_t *added, int32_t *status, void *stream) {
  if (dim < 1 || dim > kDimMax) return fail(CSD_ERR_CAPACITY, "face placement supports 1..16 dims");
  if (num_particles < 0 || capacity < 1 || max_faces < 0)
    return fail(CSD_ERR_INVALID_ARGUMENT, "bad size");
  if (num_particles > kNoAllocMaxParticles)
    return fail(CSD_ERR_CAPACITY,
                "csd_place_faces_noalloc takes at most 65536 particles (its fixed scratch)");
  if (!A || !b || !num_faces || !line_start || !line_end || !added ||
      (num_particles && (!scol || !proj)))
    return fail(CSD_ERR_INVALID_ARGUMENT, "null device pointer");
  void *base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  int32_t *ctr = (int32_t *)base;
  double *dist = (double *)((unsigned char *)base + kNoAllocCounterBytes);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_noalloc_init, dim3(1), dim3(64), 0, st, ctr, num_particles);
  if (num_particles)
    hipLaunchKernelGGL(k_face_prep, dim3((num_particles + 255) / 256), dim3(256), 0, st, scol,
                       proj, num_particles, dim, dist);
  EizoDev D;
  memset(&D, 0, sizeof(D));
  D.A = A;
  D.b = b;
  D.nf = num_faces;
  D.ls = const_cast<float *>(line_start);
  D.le = const_cast<float *>(line_end);
  D.scol = const_cast<float *>(scol);
  D.proj = const_cast<float *>(proj);
  D.dist = dist;
  D.u = ctr;
  D.status = ctr + 1;
  D.faces_added = ctr + 2;
  hipLaunchKernelGGL(k_eizo_faces, dim3(1), dim3(kFaceBlock), 0, st, D, dim,
                     std::max(num_particles, 1), capacity, margin, max_faces, 0, 0u);
  hipLaunchKernelGGL(k_noalloc_finish, dim3(1), dim3(64), 0, st, ctr, added, status);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

}  // extern "C"
