// csd_stages.hip -- EI-ZO stages as standalone entry points: hit-and-run
// sampling in many polytopes (UniformSampleInHPolyhedronCuda), projection +
// bisection of given particles, and greedy face placement for one region.
// The batched driver in csd_eizo.hip runs the same device code fused per
// round.
#include <algorithm>

#include "csd_eizo_device.h"

namespace csd {

// standalone sampler: polytope p = blockIdx.y
template <int DPL>
__global__ __launch_bounds__(kSampleBlock) void k_hit_and_run(
    const float *__restrict__ A, const float *__restrict__ b, const int32_t *__restrict__ nf,
    int face_stride, int dim, const float *__restrict__ starts, int64_t n, int mixing,
    uint32_t k0, uint32_t k1, uint32_t round, float *__restrict__ out) {
  constexpr int DP = SamplerShape<DPL>::DP;
  constexpr bool LDS = SamplerShape<DPL>::LDS;
  extern __shared__ float csd_smem[];
  const int p = blockIdx.y;
  const int F = nf[p];
  const float *Ap = A + (size_t)p * face_stride * dim, *bp = b + (size_t)p * face_stride;
  if constexpr (LDS) {
    stage_polytope<DP>(csd_smem, csd_smem + (size_t)F * DP, Ap, bp, F, dim);
    __syncthreads();
    Ap = csd_smem;
    bp = csd_smem + (size_t)F * DP;
  }
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[DP];
#pragma unroll
  for (int d = 0; d < DP; ++d) x[d] = d < dim ? starts[(size_t)p * dim + d] : 0.f;
  hit_and_run<DP, LDS>(Ap, bp, F, dim, x, (uint32_t)i, (uint32_t)p, round, mixing, k0, k1);
#pragma unroll
  for (int d = 0; d < DP; ++d)
    if (d < dim) out[((size_t)p * n + i) * dim + d] = x[d];
}

template <int VAR, int DP>
__global__ __launch_bounds__(kSampleBlock) void k_bisect_standalone(
    PlantImage P, VoxelArgs V, RobotGeomList R, const float *__restrict__ part,
    const float *__restrict__ ls, const float *__restrict__ le,
    const uint32_t *__restrict__ seg, int n, int dim, int steps, float *__restrict__ scol,
    float *__restrict__ proj) {
  extern __shared__ __attribute__((aligned(16))) float csd_smem[];
  const float *tab = stage_tables<VAR>(P, csd_smem);
  const int i = blockIdx.x * kSampleBlock + threadIdx.x;
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
  const float t = bisect_particle<VAR, DP>(P, V, R, p, q1, q2, steps, !valid, sc, tab);
  if (!valid) return;
#pragma unroll
  for (int d = 0; d < DP; ++d)
    if (d < dim) {
      scol[(size_t)i * dim + d] = sc[d];
      proj[(size_t)i * dim + d] = (1.f - t) * q1[d] + t * q2[d];
    }
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
}  // namespace csd

using namespace csd;

extern "C" {

int csd_hit_and_run(const float *A, const float *b, const int32_t *num_faces, int32_t face_stride,
                    int32_t dim, int32_t num_polytopes, const float *starts, int64_t num_samples,
                    int32_t mixing_steps, uint64_t seed, uint32_t round, float *out,
                    void *stream) {
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
  // polytopes of up to face_stride rows go to LDS when they fit
  if (dim <= 8 && polytope_lds_bytes<8>(face_stride) <= kLdsLimit)
    hipLaunchKernelGGL((k_hit_and_run<17>), g, dim3(kSampleBlock), polytope_lds_bytes<8>(face_stride),
                       st, A, b, num_faces, face_stride, dim, starts, num_samples, mixing_steps,
                       k0, k1, round, out);
  else if (polytope_lds_bytes<16>(face_stride) <= kLdsLimit)
    hipLaunchKernelGGL((k_hit_and_run<33>), g, dim3(kSampleBlock),
                       polytope_lds_bytes<16>(face_stride), st, A, b, num_faces, face_stride, dim,
                       starts, num_samples, mixing_steps, k0, k1, round, out);
  else
    hipLaunchKernelGGL((k_hit_and_run<32>), g, dim3(kSampleBlock), 0, st, A, b, num_faces,
                       face_stride, dim, starts, num_samples, mixing_steps, k0, k1, round, out);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
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
  RobotGeomList R;
  VoxelArgs V{voxels, num_voxels, voxel_radius, 0, 0};
  int rc = CSD_OK;
  bool all_slots = fill_robot_list(plant, robot_geoms, num_voxels ? num_robot_geoms : 0, R, V, &rc);
  if (rc) return rc;
  const int var = check_variant(plant->img, num_voxels > 0 && !all_slots);
  dim3 g((num_particles + kSampleBlock - 1) / kSampleBlock);
  if (dim <= 8)
    CSD_DISPATCH_VARIANT_X(var, k_bisect_standalone, 8, g, dim3(kSampleBlock),
                           check_lds_bytes(plant->img, var), (hipStream_t)stream, plant->img, V, R, particles, line_start, line_end,
                           segment_idx, num_particles, dim, steps, scol, proj);
  else
    CSD_DISPATCH_VARIANT_X(var, k_bisect_standalone, 16, g, dim3(kSampleBlock),
                           check_lds_bytes(plant->img, var), (hipStream_t)stream, plant->img, V, R, particles, line_start, line_end,
                           segment_idx, num_particles, dim, steps, scol, proj);
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
                       std::max(num_particles, 1), capacity, margin, max_faces, 0);
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

}  // extern "C"
