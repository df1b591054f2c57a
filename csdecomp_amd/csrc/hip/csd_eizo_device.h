// csd_eizo_device.h -- device code shared by the batched EI-ZO driver
// (csd_eizo.hip) and the stage-wise entry points (csd_stages.hip): Philox
// random numbers, hit-and-run chains, projection + bisection, greedy FP64
// face placement and the per-region device state.
#pragma once
#include <math.h>
#include <string.h>

#include "csd_collide.h"

namespace csd {

constexpr int kDimMax = 16;     // configuration variables handled by EI-ZO
constexpr int kSampleBlock = 256;
constexpr int kSampleBlockWide = 1024;  // the sampler's wide tile (16 waves)
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

// Polytopes are read from LDS when they fit: rows zero-padded to DP floats
// (one or more 16-byte LDS reads per row, no per-element branches), then b.
// Zero padding adds exact zeros to the dot products (x and dir are 0 past
// dim), so the arithmetic equals the unpadded loop. Larger polytopes are
// read from global memory row by row.
constexpr size_t kLdsLimit = 65536;       // 256-thread sampler blocks
constexpr size_t kLdsLimitWide = 163840;  // 1024-thread blocks: a gfx950 workgroup's LDS

template <int DP>
__host__ __device__ constexpr size_t polytope_lds_bytes(int faces) {
  return (size_t)faces * (DP + 1) * sizeof(float);
}

// BROW (dim < DP only): b_f is stored in the row's last padding slot, so a
// face is read with the row alone.
template <int DP, bool BROW = false>
__device__ __forceinline__ void stage_polytope(float *sA, float *sb, const float *A,
                                               const float *b, int F, int dim) {
  for (int e = threadIdx.x; e < F * DP; e += blockDim.x) {
    const int f = e / DP, d = e - f * DP;
    sA[e] = d < dim ? A[(size_t)f * dim + d] : (BROW && d == DP - 1) ? b[f] : 0.f;
  }
  if constexpr (!BROW)
    for (int f = threadIdx.x; f < F; f += blockDim.x) sb[f] = b[f];
}

// One hit-and-run chain (cuda_hit_and_run_sampling.cu:74-112) of `mixing`
// steps from x, in the polytope (A row-major [F][stride], b). A and b are
// wave-uniform: rows are LDS broadcasts (LDS = true, stride DP) or scalar
// loads (stride dim). A failed step (empty chord) leaves the chain where it
// is. DIMC > 0: dim == DIMC, and the dot products stop there (the padding
// slots they skip only added exact zeros).
template <int DP, bool LDS, bool BROW = false, int DIMC = 0>
__device__ __forceinline__ void hit_and_run(const float *__restrict__ A,
                                            const float *__restrict__ b, int F, int dim,
                                            float (&x)[DP], uint32_t chain, uint32_t stream,
                                            uint32_t round, int mixing, uint32_t k0,
                                            uint32_t k1) {
  const int stride = LDS ? DP : dim;
  const int npairs = (dim + 1) >> 1;
  const int ncalls = (2 * npairs + 1 + 3) >> 2;
  // uniform number 2*npairs (after the Gaussian pairs) sets the step length;
  // it is picked while the Philox outputs are formed, so no array is
  // indexed by a runtime value
  const int jt = 2 * npairs;
  for (int s = 0; s < mixing; ++s) {
    float u[DP + 4];
    float ut = 0.5f;
#pragma unroll
    for (int c = 0; c < (DP + 4) / 4; ++c) {
      if (c < ncalls) {
        const uint4 o = philox4x32_10(make_uint4(chain, (uint32_t)s, stream, round * 16u + c),
                                      k0, k1);
        u[4 * c] = u01(o.x);
        u[4 * c + 1] = u01(o.y);
        u[4 * c + 2] = u01(o.z);
        u[4 * c + 3] = u01(o.w);
        if (c == (jt >> 2)) {
          const int e = jt & 3;
          ut = e == 0 ? u[4 * c] : e == 1 ? u[4 * c + 1] : e == 2 ? u[4 * c + 2] : u[4 * c + 3];
        }
      } else {
        u[4 * c] = u[4 * c + 1] = u[4 * c + 2] = u[4 * c + 3] = 0.5f;
      }
    }
    // Box-Muller with the hardware transcendentals: r = sqrt(-2 ln u) via
    // log2 (v_log_f32), and sin/cos of 2*pi*u straight from v_sin_f32 /
    // v_cos_f32, whose argument is in revolutions (u in (0, 1]). Their
    // ~1e-6 relative error only perturbs the direction, which any
    // rotation-invariant law serves.
    float dir[DP];
#pragma unroll
    for (int k = 0; k < DP / 2; ++k) {
      if (k < npairs) {
        const float r =
            __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u[2 * k]));
        dir[2 * k] = r * __builtin_amdgcn_cosf(u[2 * k + 1]);
        dir[2 * k + 1] = r * __builtin_amdgcn_sinf(u[2 * k + 1]);
      } else {
        dir[2 * k] = dir[2 * k + 1] = 0.f;
      }
    }
    // Chord of the line x + t dir: face f bounds t by its slack
    // lb = b_f - a_f x over ad = a_f dir, from above if ad > 0 and from below
    // if ad < 0. With the chain inside (lb > 0) the tightest bounds are
    // 1 / max_f(ad / lb) and 1 / min_f(ad / lb), so a face costs one
    // reciprocal (hardware, 1 ulp), a product and a max and a min, with no
    // compare or select; the two reciprocals of the extremes follow the
    // loop. A slack that rounding left at or below zero counts as 1e-30:
    // the face then stops the chain from moving further out (bound ~0 on
    // that side), as the reference's quotient lb / ad does. ad = 0 gives
    // ad / lb = 0 and no bound.
    float qmax = 0.f, qmin = 0.f;
#pragma unroll 2
    for (int f = 0; f < F; ++f) {
      const float *a = A + (size_t)f * stride;
      float ax = 0.f, ad = 0.f;
      // with b in the last slot that slot is skipped: x and dir are 0 there,
      // so the skipped terms were exact zeros
#pragma unroll
      for (int d = 0; d < (DIMC > 0 ? DIMC : BROW ? DP - 1 : DP); ++d) {
        if (LDS || d < dim) {
          ax = fmaf(a[d], x[d], ax);
          ad = fmaf(a[d], dir[d], ad);
        }
      }
      const float bf = BROW ? a[DP - 1] : b[f];
      const float q = ad * __builtin_amdgcn_rcpf(fmaxf(bf - ax, 1e-30f));
      qmax = fmaxf(qmax, q);
      qmin = fminf(qmin, q);
    }
    const float tmax = qmax > 0.f ? 1.f / qmax : 3.402823466e+38f;
    const float tmin = qmin < 0.f ? 1.f / qmin : -3.402823466e+38f;
    if (isinf(tmax) || isinf(tmin) || tmax < tmin) continue;
    // theta = u (0.999999 tmax - tmin) + tmin, in double as the reference
    const float theta = (float)((double)ut * (0.999999 * (double)tmax - (double)tmin) +
                                (double)tmin);
#pragma unroll
    for (int d = 0; d < DP; ++d) x[d] = d < dim ? fmaf(theta, dir[d], x[d]) : 0.f;
  }
}

// Kernels that sample take DPL = 2 * DP + (1 if the polytope is in LDS).
template <int DPL>
struct SamplerShape {
  static constexpr int DP = DPL >> 1;
  static constexpr bool LDS = (DPL & 1) != 0;
};

// ---------------------------------------------------------------------------
// EI-ZO device state (one slot per region)
// ---------------------------------------------------------------------------
struct EizoDev {
  float *A, *b;            // [R][cap][dim], [R][cap]
  int32_t *nf;             // [R]
  float *ls, *le;          // [R][dim]
  uint32_t *budget;        // [R]
  uint32_t *nsamp;         // [R] samples drawn by the current sampling phase
  uint32_t *pilot_col;     // [R] colliding samples of the last untested phase
  int32_t *active;         // [R] region ids of the current launch
  float *stage;            // [nactive][nblk] segments of 256*dim: samples ([dim][256]), then
                           // the colliding ones compacted in place ([count][dim])
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
  // per-round statistics [R][T][kStats] (null in the stage-wise entry
  // points); round k of region r at (r * T + k) * kStats
  int32_t *stats;
  int T;
  // ellipsoid metric of the face placement, dim x dim row-major FP64 (null:
  // the Euclidean metric of the reference)
  const double *metric;
  // Device-resident rounds (graph-captured): null when the host drives the
  // rounds. ctl = {next round, active regions, current round, M_k, test
  // threshold bits}; mtab/ttab hold M_k and the threshold of every round.
  uint32_t *ctl;
  const uint32_t *mtab;
  const float *ttab;
};

// Fields of a region's per-round statistics (EizoIterationInfo of
// cuda_edge_inflation_zero_order.h:37-47, plus what guided rounds add).
enum {
  kStatSamples = 0,     // samples of the unadaptive test (0: test skipped)
  kStatFirstCol = 1,    // collisions among the first M_k of them (-1: skipped)
  kStatCol = 2,         // collisions among all of them
  kStatUpdated = 3,     // particles bisected
  kStatFaces = 4,       // faces placed
  kStatFacesStart = 5,  // faces of the region when the round began
  kStatPilot = 6,       // pilot samples
  kStatPilotCol = 7,    // collisions among the pilot samples
  kStatPilotChecked = 8,  // pilot samples checked exactly (pruned pilots)
  kStats = 9
};

__device__ __forceinline__ int32_t *round_stats(const EizoDev &D, int r, uint32_t round) {
  return D.stats && round < (uint32_t)D.T ? D.stats + ((size_t)r * D.T + round) * kStats
                                          : nullptr;
}

// Projection parameter of p on the segment q1-q2 and the projected point
// (projectSamplesOntoLineSegmentsKernel, cuda_set_builder_utils.cu:25-58),
// as separately rounded float operations in the reference's order: every
// kernel that projects obtains the same bits, and so does a plain C
// restatement. Padding dimensions (zero in p, q1, q2) add exact zeros.
template <int DP>
__device__ __forceinline__ float seg_param(const float (&p)[DP], const float (&q1)[DP],
                                           const float (&q2)[DP]) {
#pragma clang fp contract(off)
  float a = 0.f, bb = 0.f;
#pragma unroll
  for (int d = 0; d < DP; ++d) {
    a = a + (p[d] - q1[d]) * (q2[d] - q1[d]);
    bb = bb + (q2[d] - q1[d]) * (q2[d] - q1[d]);
  }
  return fminf(fmaxf(a / bb, 0.f), 1.f);
}

__device__ __forceinline__ float seg_point(float t, float q1, float q2) {
#pragma clang fp contract(off)
  return (1.f - t) * q1 + t * q2;
}

// Projection of particle p onto the segment q1-q2 (its first `dim`
// coordinates written to proj_out right away, unless proj_out is null) and
// `steps` bisection steps between the projection (assumed free) and p
// (colliding), with FK + collision check inline (cuda_set_builder_utils.cu
// :25-67, 239-357). sc is the last colliding midpoint (p itself if none
// collides): it always equals the upper bound, so only the two bounds stay
// live across the steps (the projection goes to memory before them).
// Idle lanes pass idle = true.
// With S > 1, S consecutive waves of the workgroup hold the same 64
// particles; wave `sub` of the group checks its share of the pairs and the
// verdicts are OR-ed through `part` (one byte per thread) at every step, so
// all S waves follow the same bisection. Every thread of the workgroup must
// call this with the same `steps`.
// A wave-uniform pointer whose value the compiler cannot see through (an
// empty asm keeps it in scalar registers), so loads through it are not
// hoisted out of the loop that re-reads them.
template <class T>
__device__ __forceinline__ const T *opaque_uniform(const T *p) {
  asm volatile("" : "+s"(p));
  return p;
}

template <int VAR, int DP, int S = 1>
__device__ __forceinline__ void bisect_particle(const PlantImage &P, const VoxelArgs &V,
                                                const RobotGeomList &R, const float (&p)[DP],
                                                const float (&q1)[DP], const float (&q2)[DP],
                                                int steps, bool idle, float (&sc)[DP],
                                                float *proj_out, int dim, const float *tab,
                                                int sub = 0, uint8_t *part = nullptr) {
  const float t = seg_param<DP>(p, q1, q2);
  float lo[DP], hi[DP], mid[DP];
#pragma unroll
  for (int d = 0; d < DP; ++d) {
    lo[d] = seg_point(t, q1[d], q2[d]);
    hi[d] = p[d];
    if (proj_out && d < dim) proj_out[d] = lo[d];
  }
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int d = 0; d < DP; ++d) mid[d] = 0.5f * (hi[d] + lo[d]);
    // fast plants (32 slots with one wave per particle; 64 and 128 slots):
    // the plant's body and sphere records are read again in every step;
    // hoisted out of the loop they stay live across all steps and spill (to
    // scratch for 128 slots)
    PlantImage Ps = P;
    if constexpr ((VAR == kFast8x4 && S == 1) || VAR == kFast8x8 || VAR == kFast16x4 ||
                  VAR == kFast16x8) {
      Ps.bodies = opaque_uniform(P.bodies);
      Ps.slot_sphere = opaque_uniform(P.slot_sphere);
      Ps.self_active = opaque_uniform(P.self_active);
    }
    bool col = collides_variant<VAR, false, 3 | kCoherent | (S == 1 ? kGrouped : 0)>(
        Ps, QReg<DP>{mid}, V, R, idle, tab, sub, S);
    if constexpr (S > 1) {
      part[threadIdx.x] = col ? 1 : 0;
      __syncthreads();
      const int base = ((int)(threadIdx.x >> 6) / S) * S * 64 + (int)(threadIdx.x & 63);
      col = false;
#pragma unroll
      for (int k = 0; k < S; ++k) col |= part[base + k * 64] != 0;
      __syncthreads();
    }
#pragma unroll
    for (int d = 0; d < DP; ++d) {
      hi[d] = col ? mid[d] : hi[d];
      lo[d] = col ? lo[d] : mid[d];
    }
  }
#pragma unroll
  for (int d = 0; d < DP; ++d) sc[d] = hi[d];
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

static __global__ __launch_bounds__(kFaceBlock) void k_eizo_faces(EizoDev D, int dim, int U, int cap,
                                                           float margin, int max_faces,
                                                           int by_slot, uint32_t round) {
  if (by_slot && D.ctl) {
    if (blockIdx.x >= D.ctl[1]) return;  // slot past this round's regions
    round = D.ctl[2];
  }
  const int r = by_slot ? D.active[blockIdx.x] : blockIdx.x;
  if (D.status[r] != 0) return;
  const int u = D.u[r];
  if (u == 0) return;
  __shared__ double s_v[kFaceBlock / 64];
  __shared__ int s_i[kFaceBlock / 64];
  __shared__ double s_a[kDimMax];
  __shared__ double s_b;
  __shared__ int s_best, s_stop;
  __shared__ double s_M[kDimMax * kDimMax];
  double *dist = D.dist + (size_t)r * U;
  const float *scol = D.scol + (size_t)r * U * dim;
  const float *proj = D.proj + (size_t)r * U * dim;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (D.metric) {
    // particles are ordered by |scol - proj|_M = sqrt(e' M e)
    for (int k = threadIdx.x; k < dim * dim; k += kFaceBlock) s_M[k] = D.metric[k];
    __syncthreads();
    for (int i = threadIdx.x; i < u; i += kFaceBlock) {
      double e[kDimMax];
      for (int d = 0; d < dim; ++d)
        e[d] = (double)scol[(size_t)i * dim + d] - (double)proj[(size_t)i * dim + d];
      double q = 0.0;
      for (int a = 0; a < dim; ++a) {
        double me = 0.0;
        for (int c = 0; c < dim; ++c) me += s_M[a * dim + c] * e[c];
        q += e[a] * me;
      }
      dist[i] = sqrt(q);
    }
    __syncthreads();
  }
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
        for (int d = 0; d < dim; ++d) a[d] = (double)sc[d] - (double)pr[d];
        if (D.metric) {
          // normal M e: the hyperplane tangent at scol to the M-ellipsoid
          // centred at proj
          double e[kDimMax];
          for (int d = 0; d < dim; ++d) e[d] = a[d];
          for (int d = 0; d < dim; ++d) {
            double v = 0.0;
            for (int c = 0; c < dim; ++c) v += s_M[d * dim + c] * e[c];
            a[d] = v;
          }
        }
        for (int d = 0; d < dim; ++d) nrm += a[d] * a[d];
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
  if (threadIdx.x == 0) {
    D.faces_added[r] += added;
    if (int32_t *st = round_stats(D, r, round)) st[kStatFaces] = added;
  }
}

inline int fail(int code, const char *msg) {
  set_error(msg);
  return code;
}

}  // namespace csd
