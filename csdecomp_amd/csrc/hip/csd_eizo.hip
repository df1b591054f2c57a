// csd_eizo.hip -- hit-and-run sampling and batched EI-ZO edge inflation.
//
// Reference algorithm: planning/cuda_edge_inflation_zero_order.cu:94-339
// (executeEizo) with the kernels of planning/cuda_set_builder_utils.cu and
// geometry/cuda_hit_and_run_sampling.cu. Differences, all deliberate:
//  * many regions (edges) advance round by round together; one launch per
//    stage per round serves every active region;
//  * sampling and collision checking are fused: a lane runs its hit-and-run
//    chain and checks the final sample in registers; only colliding samples
//    are written to HBM;
//  * colliding samples are taken in sample-index order (block ballot +
//    prefix sums) instead of atomicAdd order, so results are deterministic;
//  * randomness is Philox4x32-10 keyed by the seed, with counters
//    (sample, step, region, round*16 + k): every round draws fresh samples;
//  * all bisection steps of a particle run in one lane with FK + collision
//    check inline;
//  * face placement and the redundancy test run in FP64, one workgroup per
//    region, all faces of a round in one launch; b is rounded up when stored
//    in FP32 so the segment stays inside the stored face.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "csd_collide.h"

namespace csd {

constexpr int kDimMax = 16;     // configuration variables handled by EI-ZO
constexpr int kSampleBlock = 256;
constexpr int kFaceBlock = 1024;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 24-bit uniform in (0, 1]
__device__ __forceinline__ float u01(uint32_t x) {
  return (float)(x >> 8) * (1.0f / 16777216.0f) + (1.0f / 33554432.0f);
}

// One hit-and-run chain (cuda_hit_and_run_sampling.cu:74-112) of `mixing`
// steps from x, in the polytope (A row-major [F][dim], b). A and b are
// wave-uniform, so every face row is a scalar load. A failed step (empty
// chord) leaves the chain where it is.
__device__ __forceinline__ void hit_and_run(const float *__restrict__ A,
                                            const float *__restrict__ b, int F, int dim,
                                            float (&x)[kDimMax], uint32_t chain,
                                            uint32_t stream, uint32_t round, int mixing,
                                            uint32_t k0, uint32_t k1) {
  const int npairs = (dim + 1) >> 1;
  const int ncalls = (2 * npairs + 1 + 3) >> 2;
  for (int s = 0; s < mixing; ++s) {
    float u[kDimMax + 4];
#pragma unroll
    for (int c = 0; c < (kDimMax + 4) / 4; ++c) {
      if (c < ncalls) {
        const uint4 o = philox4x32_10(make_uint4(chain, (uint32_t)s, stream, round * 16u + c),
                                      k0, k1);
        u[4 * c] = u01(o.x);
        u[4 * c + 1] = u01(o.y);
        u[4 * c + 2] = u01(o.z);
        u[4 * c + 3] = u01(o.w);
      } else {
        u[4 * c] = u[4 * c + 1] = u[4 * c + 2] = u[4 * c + 3] = 0.5f;
      }
    }
    float dir[kDimMax];
#pragma unroll
    for (int k = 0; k < kDimMax / 2; ++k) {
      if (k < npairs) {
        const float r = sqrtf(-2.0f * logf(u[2 * k]));
        float sn, cs;
        sincosf(6.2831853071795864769f * u[2 * k + 1], &sn, &cs);
        dir[2 * k] = r * cs;
        dir[2 * k + 1] = r * sn;
      }
    }
    float tmax = 3.402823466e+38f, tmin = -3.402823466e+38f;
    for (int f = 0; f < F; ++f) {
      const float *a = A + (size_t)f * dim;
      float ax = 0.f, ad = 0.f;
#pragma unroll
      for (int d = 0; d < kDimMax; ++d) {
        if (d < dim) {
          ax = fmaf(a[d], x[d], ax);
          ad = fmaf(a[d], dir[d], ad);
        }
      }
      const float lb = b[f] - ax;
      if (ad < 0.f) tmin = fmaxf(tmin, lb / ad);
      else if (ad > 0.f) tmax = fminf(tmax, lb / ad);
    }
    if (isinf(tmax) || isinf(tmin) || tmax < tmin) continue;
    // theta = u (0.999999 tmax - tmin) + tmin, in double as the reference
    const float theta = (float)((double)pick<kDimMax + 4>(u, 2 * npairs) *
                                    (0.999999 * (double)tmax - (double)tmin) +
                                (double)tmin);
#pragma unroll
    for (int d = 0; d < kDimMax; ++d)
      if (d < dim) x[d] = fmaf(theta, dir[d], x[d]);
  }
}

// standalone sampler: polytope p = blockIdx.y
__global__ __launch_bounds__(kSampleBlock) void k_hit_and_run(
    const float *__restrict__ A, const float *__restrict__ b, const int32_t *__restrict__ nf,
    int face_stride, int dim, const float *__restrict__ starts, int64_t n, int mixing,
    uint32_t k0, uint32_t k1, uint32_t round, float *__restrict__ out) {
  const int p = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[kDimMax];
#pragma unroll
  for (int d = 0; d < kDimMax; ++d) x[d] = d < dim ? starts[(size_t)p * dim + d] : 0.f;
  hit_and_run(A + (size_t)p * face_stride * dim, b + (size_t)p * face_stride, nf[p], dim, x,
              (uint32_t)i, (uint32_t)p, round, mixing, k0, k1);
#pragma unroll
  for (int d = 0; d < kDimMax; ++d)
    if (d < dim) out[((size_t)p * n + i) * dim + d] = x[d];
}

// ---------------------------------------------------------------------------
// EI-ZO device state (one slot per region)
// ---------------------------------------------------------------------------
struct EizoDev {
  float *A, *b;            // [R][cap][dim], [R][cap]
  int32_t *nf;             // [R]
  float *ls, *le;          // [R][dim]
  uint32_t *budget;        // [R]
  int32_t *active;         // [R] region ids of this round
  float *stage;            // [nactive][nblk][256][dim] colliding samples per block
  uint32_t *blk_cnt;       // [nactive][nblk]
  uint32_t *blk_first;     // [nactive][nblk]
  float *part;             // [R][U][dim] particles to bisect
  float *scol, *proj;      // [R][U][dim]
  double *dist;            // [R][U]
  int32_t *u;              // [R] particles this round
  int32_t *status;         // [R] 0 running, 1 certified, 2 face overflow
  int32_t *faces_added;    // [R]
  int32_t *iters;          // [R]
  int64_t *samples;        // [R]
  int64_t *collisions;     // [R]
};

// Fused round stage 1: hit-and-run chain + collision check per lane;
// per-block compaction of colliding samples in sample order.
template <int VAR>
__global__ __launch_bounds__(kSampleBlock) void k_eizo_sample_check(
    PlantImage P, VoxelArgs V, RobotGeomList R, EizoDev D, int dim, int cap, int nblk,
    uint32_t M, int mixing, uint32_t k0, uint32_t k1, uint32_t round) {
  const int slot = blockIdx.y;
  const int r = D.active[slot];
  const uint32_t N = max(M, D.budget[r]);
  const uint32_t i = blockIdx.x * kSampleBlock + threadIdx.x;
  const bool valid = i < N;
  float x[kDimMax];
#pragma unroll
  for (int d = 0; d < kDimMax; ++d) x[d] = d < dim ? D.ls[(size_t)r * dim + d] : 0.f;
  if (blockIdx.x * kSampleBlock < N)
    hit_and_run(D.A + (size_t)r * cap * dim, D.b + (size_t)r * cap, D.nf[r], dim, x, i,
                (uint32_t)r, round, mixing, k0, k1);
  const bool col = collides_variant<VAR>(P, QReg<kDimMax>{x}, V, R, !valid) && valid;

  __shared__ uint32_t wcnt[kSampleBlock / 64], wfirst[kSampleBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t m = __ballot(col);
  const uint64_t mf = __ballot(col && i < M);
  if (lane == 0) {
    wcnt[wave] = (uint32_t)__popcll(m);
    wfirst[wave] = (uint32_t)__popcll(mf);
  }
  __syncthreads();
  uint32_t off = 0, tot = 0, totf = 0;
#pragma unroll
  for (int w = 0; w < kSampleBlock / 64; ++w) {
    off += w < wave ? wcnt[w] : 0;
    tot += wcnt[w];
    totf += wfirst[w];
  }
  if (col) {
    const uint32_t rank = off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    float *dst = D.stage + (((size_t)slot * nblk + blockIdx.x) * kSampleBlock + rank) * dim;
#pragma unroll
    for (int d = 0; d < kDimMax; ++d)
      if (d < dim) dst[d] = x[d];
  }
  if (threadIdx.x == 0) {
    D.blk_cnt[(size_t)slot * nblk + blockIdx.x] = tot;
    D.blk_first[(size_t)slot * nblk + blockIdx.x] = totf;
  }
}

// Stage 2: unadaptive test (cuda_edge_inflation_zero_order.cu:165-199) and
// gather of the first budget_r colliding samples, one workgroup per region.
__global__ __launch_bounds__(256) void k_eizo_compact(EizoDev D, int dim, int nblk, uint32_t M,
                                                      float thresh, int U) {
  const int slot = blockIdx.x;
  const int r = D.active[slot];
  const uint32_t N = max(M, D.budget[r]);
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
    D.iters[r] += 1;
    const bool certified = (float)first <= thresh;
    if (certified) D.status[r] = 1;
    s_u = certified ? 0 : (int)min(acc, min(D.budget[r], (uint32_t)U));
    D.u[r] = s_u;
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

// Stage 3: projection onto the segment + all bisection steps per lane
// (cuda_set_builder_utils.cu:25-67, 239-357).
template <int VAR>
__global__ __launch_bounds__(kSampleBlock) void k_eizo_bisect(PlantImage P, VoxelArgs V,
                                                              RobotGeomList R, EizoDev D,
                                                              int dim, int U, int steps) {
  const int r = D.active[blockIdx.y];
  if (D.status[r] != 0) return;
  const int u = D.u[r];
  const int i = blockIdx.x * kSampleBlock + threadIdx.x;
  if ((int)(blockIdx.x * kSampleBlock) >= u) return;
  const bool valid = i < u;
  const int ii = valid ? i : u - 1;
  float p[kDimMax], q1[kDimMax], q2[kDimMax];
  float a = 0.f, bb = 0.f;
#pragma unroll
  for (int d = 0; d < kDimMax; ++d) {
    p[d] = d < dim ? D.part[((size_t)r * U + ii) * dim + d] : 0.f;
    q1[d] = d < dim ? D.ls[(size_t)r * dim + d] : 0.f;
    q2[d] = d < dim ? D.le[(size_t)r * dim + d] : 0.f;
    a = fmaf(p[d] - q1[d], q2[d] - q1[d], a);
    bb = fmaf(q2[d] - q1[d], q2[d] - q1[d], bb);
  }
  const float t = fminf(fmaxf(a / bb, 0.f), 1.f);
  float lo[kDimMax], hi[kDimMax], sc[kDimMax], mid[kDimMax];
#pragma unroll
  for (int d = 0; d < kDimMax; ++d) {
    lo[d] = (1.f - t) * q1[d] + t * q2[d];  // projection
    hi[d] = p[d];
    sc[d] = p[d];
  }
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int d = 0; d < kDimMax; ++d) mid[d] = 0.5f * (hi[d] + lo[d]);
    const bool col = collides_variant<VAR>(P, QReg<kDimMax>{mid}, V, R, !valid);
#pragma unroll
    for (int d = 0; d < kDimMax; ++d) {
      hi[d] = col ? mid[d] : hi[d];
      sc[d] = col ? mid[d] : sc[d];
      lo[d] = col ? lo[d] : mid[d];
    }
  }
  if (!valid) return;
  double dd = 0.0;
#pragma unroll
  for (int d = 0; d < kDimMax; ++d) {
    if (d < dim) {
      const float pr = (1.f - t) * q1[d] + t * q2[d];
      D.scol[((size_t)r * U + i) * dim + d] = sc[d];
      D.proj[((size_t)r * U + i) * dim + d] = pr;
      const double e = (double)sc[d] - (double)pr;
      dd += e * e;
    }
  }
  D.dist[(size_t)r * U + i] = sqrt(dd);
}

// Stage 4: greedy face placement in FP64 (maybePlaceFace +
// checkParticlesMadeRedundantByLastFace, cuda_set_builder_utils.cu:359-424),
// up to max_faces faces, one workgroup per region.
__device__ __forceinline__ void argmin_pair(double &v, int &i, double ov, int oi) {
  if (ov < v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}

__global__ __launch_bounds__(kFaceBlock) void k_eizo_faces(EizoDev D, int dim, int U, int cap,
                                                           float margin, int max_faces,
                                                           int by_slot) {
  const int r = by_slot ? D.active[blockIdx.x] : blockIdx.x;
  if (D.status[r] != 0) return;
  const int u = D.u[r];
  if (u == 0) return;
  __shared__ double s_v[kFaceBlock / 64];
  __shared__ int s_i[kFaceBlock / 64];
  __shared__ double s_a[kDimMax];
  __shared__ double s_b;
  __shared__ int s_best, s_stop;
  double *dist = D.dist + (size_t)r * U;
  const float *scol = D.scol + (size_t)r * U * dim;
  const float *proj = D.proj + (size_t)r * U * dim;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int added = 0;
  for (; added < max_faces; ++added) {
    double v = INFINITY;
    int vi = 0x7fffffff;
    for (int i = threadIdx.x; i < u; i += kFaceBlock) argmin_pair(v, vi, dist[i], i);
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
    if (threadIdx.x == 0) {
      double bv = s_v[0];
      int bi = s_i[0];
      for (int w = 1; w < kFaceBlock / 64; ++w) argmin_pair(bv, bi, s_v[w], s_i[w]);
      s_stop = 0;
      if (!(bv < INFINITY)) {
        s_stop = 1;  // every particle is redundant
      } else if (D.nf[r] >= cap) {
        s_stop = 1;
        D.status[r] = 2;  // face capacity exhausted
      } else {
        const float *sc = scol + (size_t)bi * dim, *pr = proj + (size_t)bi * dim;
        const float *ls = D.ls + (size_t)r * dim, *le = D.le + (size_t)r * dim;
        double a[kDimMax], nrm = 0.0;
        for (int d = 0; d < dim; ++d) {
          a[d] = (double)sc[d] - (double)pr[d];
          nrm += a[d] * a[d];
        }
        nrm = sqrt(nrm);
        const int f = D.nf[r];
        double bb = -(double)margin, v1 = 0.0, v2 = 0.0;
        for (int d = 0; d < dim; ++d) {
          const float af = (float)(a[d] / nrm);
          D.A[((size_t)r * cap + f) * dim + d] = af;
          a[d] = af;
          s_a[d] = af;
          bb += a[d] * sc[d];
          v1 += a[d] * ls[d];
          v2 += a[d] * le[d];
        }
        const double relax = fmax(v1 - bb, v2 - bb);
        if (relax > 0.0) bb += relax;  // keep the segment inside the face
        D.b[(size_t)r * cap + f] = __double2float_ru(bb);
        D.nf[r] = f + 1;
        s_b = bb;
        dist[bi] = INFINITY;
      }
    }
    __syncthreads();
    if (s_stop) break;
    for (int i = threadIdx.x; i < u; i += kFaceBlock) {
      if (!(dist[i] < INFINITY)) continue;
      double val = -s_b;
      for (int d = 0; d < dim; ++d) val += s_a[d] * (double)scol[(size_t)i * dim + d];
      if (val > 0.0) dist[i] = INFINITY;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) D.faces_added[r] += added;
}

namespace {
int fail(int code, const char *msg) {
  set_error(msg);
  return code;
}
}  // namespace

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
}  // namespace

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
  hipLaunchKernelGGL(k_hit_and_run, g, dim3(kSampleBlock), 0, (hipStream_t)stream, A, b,
                     num_faces, face_stride, dim, starts, num_samples, mixing_steps,
                     (uint32_t)seed, (uint32_t)(seed >> 32), round, out);
  CSD_HIP(hipGetLastError());
  return CSD_OK;
}

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
  if (rc) {
    csd_eizo_destroy(w);
    return rc;
  }
  *out = w;
  return CSD_OK;
}

int csd_eizo_destroy(csd_eizo *w) {
  if (!w) return CSD_OK;
  for (void *p : w->owned) hipFree(p);
  if (w->D.stage) hipFree(w->D.stage);
  if (w->D.blk_cnt) hipFree(w->D.blk_cnt);
  if (w->D.blk_first) hipFree(w->D.blk_first);
  delete w;
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

  RobotGeomList RL;
  VoxelArgs V{w->vox, num_voxels, voxel_radius, 0, 0};
  int rc = CSD_OK;
  bool all_slots = fill_robot_list(w->plant, w->robot.data(), (int)w->robot.size(), RL, V, &rc);
  if (rc) return rc;
  const PlantImage &P = w->plant->img;
  const int var = check_variant(P, num_voxels > 0 && !all_slots);
  const uint32_t k0 = (uint32_t)o.seed, k1 = (uint32_t)(o.seed >> 32);
  std::vector<int32_t> status(Rn);

  for (uint32_t it = 0; it < o.max_iterations && !active.empty(); ++it) {
    const int na = (int)active.size();
    // cuda_edge_inflation_zero_order.cu:113-119 (float arithmetic)
    const float delta_k = (float)(6 * o.delta / (M_PI * M_PI * (it + 1) * (it + 1)));
    const uint32_t M = (uint32_t)ceilf(2 * logf(1 / delta_k) / (o.epsilon * o.tau * o.tau));
    const uint32_t N = std::max(M, max_budget);
    const int nblk = (int)((N + kSampleBlock - 1) / kSampleBlock);
    if (nblk > 1024) return fail(CSD_ERR_CAPACITY, "more than 262144 samples in one round");
    const size_t stage_need = (size_t)na * nblk * kSampleBlock * dim * sizeof(float);
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
    CSD_HIP(hipMemcpy(D.active, active.data(), na * 4, hipMemcpyHostToDevice));
    const float thresh = (1 - o.tau) * o.epsilon * M;
    CSD_DISPATCH_VARIANT(var, k_eizo_sample_check, dim3(nblk, na), dim3(kSampleBlock), 0, 0, P,
                         V, RL, D, dim, cap, nblk, M, (int)o.mixing_steps, k0, k1, it);
    hipLaunchKernelGGL(k_eizo_compact, dim3(na), dim3(256), 0, 0, D, dim, nblk, M, thresh, U);
    const int ublk = (int)((max_budget + kSampleBlock - 1) / kSampleBlock);
    CSD_DISPATCH_VARIANT(var, k_eizo_bisect, dim3(ublk, na), dim3(kSampleBlock), 0, 0, P, V, RL,
                         D, dim, U, (int)o.bisection_steps);
    hipLaunchKernelGGL(k_eizo_faces, dim3(na), dim3(kFaceBlock), 0, 0, D, dim, U, cap,
                       o.configuration_margin, (int)o.max_hyperplanes_per_iteration, 1);
    CSD_HIP(hipGetLastError());
    CSD_HIP(hipMemcpy(status.data(), D.status, Rn * 4, hipMemcpyDeviceToHost));
    std::vector<int32_t> next;
    for (int r : active) {
      if (status[r] == 2) return fail(CSD_ERR_CAPACITY, "face capacity exhausted in a region");
      if (status[r] == 0) next.push_back(r);
    }
    active.swap(next);
  }

  CSD_HIP(hipMemcpy(out_A, D.A, (size_t)Rn * cap * dim * 4, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(out_b, D.b, (size_t)Rn * cap * 4, hipMemcpyDeviceToHost));
  std::vector<int32_t> iters(Rn), fa(Rn);
  std::vector<int64_t> smp(Rn), colls(Rn);
  CSD_HIP(hipMemcpy(nf.data(), D.nf, Rn * 4, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(iters.data(), D.iters, Rn * 4, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(fa.data(), D.faces_added, Rn * 4, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(smp.data(), D.samples, Rn * 8, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(colls.data(), D.collisions, Rn * 8, hipMemcpyDeviceToHost));
  CSD_HIP(hipMemcpy(status.data(), D.status, Rn * 4, hipMemcpyDeviceToHost));
  for (int r = 0; r < Rn; ++r) {
    infos[r].iterations = iters[r];
    infos[r].converged = status[r] == 1;
    infos[r].num_faces = nf[r];
    infos[r].faces_added = fa[r];
    infos[r].samples_drawn = smp[r];
    infos[r].collisions_found = colls[r];
  }
  return CSD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// stage-wise entry points (used by the stage parity tests)
// ---------------------------------------------------------------------------
namespace csd {
template <int VAR>
__global__ __launch_bounds__(kSampleBlock) void k_bisect_standalone(
    PlantImage P, VoxelArgs V, RobotGeomList R, const float *__restrict__ part,
    const float *__restrict__ ls, const float *__restrict__ le,
    const uint32_t *__restrict__ seg, int n, int dim, int steps, float *__restrict__ scol,
    float *__restrict__ proj) {
  const int i = blockIdx.x * kSampleBlock + threadIdx.x;
  const bool valid = i < n;
  const int ii = valid ? i : n - 1;
  const uint32_t s_idx = seg[ii];
  float p[kDimMax], q1[kDimMax], q2[kDimMax];
  float a = 0.f, bb = 0.f;
#pragma unroll
  for (int d = 0; d < kDimMax; ++d) {
    p[d] = d < dim ? part[(size_t)ii * dim + d] : 0.f;
    q1[d] = d < dim ? ls[(size_t)s_idx * dim + d] : 0.f;
    q2[d] = d < dim ? le[(size_t)s_idx * dim + d] : 0.f;
    a = fmaf(p[d] - q1[d], q2[d] - q1[d], a);
    bb = fmaf(q2[d] - q1[d], q2[d] - q1[d], bb);
  }
  const float t = fminf(fmaxf(a / bb, 0.f), 1.f);
  float lo[kDimMax], hi[kDimMax], sc[kDimMax], mid[kDimMax];
#pragma unroll
  for (int d = 0; d < kDimMax; ++d) {
    lo[d] = (1.f - t) * q1[d] + t * q2[d];
    hi[d] = p[d];
    sc[d] = p[d];
  }
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int d = 0; d < kDimMax; ++d) mid[d] = 0.5f * (hi[d] + lo[d]);
    const bool col = collides_variant<VAR>(P, QReg<kDimMax>{mid}, V, R, !valid);
#pragma unroll
    for (int d = 0; d < kDimMax; ++d) {
      hi[d] = col ? mid[d] : hi[d];
      sc[d] = col ? mid[d] : sc[d];
      lo[d] = col ? lo[d] : mid[d];
    }
  }
  if (!valid) return;
#pragma unroll
  for (int d = 0; d < kDimMax; ++d)
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

extern "C" {

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
  CSD_DISPATCH_VARIANT(var, k_bisect_standalone, g, dim3(kSampleBlock), 0, (hipStream_t)stream,
                       plant->img, V, R, particles, line_start, line_end, segment_idx,
                       num_particles, dim, steps, scol, proj);
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
