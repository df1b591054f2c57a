// csd_device.h -- device-side building blocks shared by the gfx950 kernels:
// body-frame forward kinematics and the FP32 pair primitives.
//
// Comparison rule: the GPU rule of the reference
// (geometry/cuda_geometry_utilities.h:16-125): sphere-sphere and sphere-box
// are free iff squared distance >= squared radius sum; box-box is free iff a
// separating axis has lhs > rhs.
#pragma once
#include "csd_internal.h"

namespace csd {

// ---------------------------------------------------------------------------
// 3x4 row-major affine helpers: m[r*4+c], column 3 = translation
// ---------------------------------------------------------------------------
__device__ __forceinline__ void aff_identity(float *X) {
#pragma unroll
  for (int k = 0; k < 12; ++k) X[k] = (k == 0 || k == 5 || k == 10) ? 1.f : 0.f;
}

// C = A * B
__device__ __forceinline__ void aff_mul(const float *A, const float *B, float *C) {
  float T[12];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      T[r * 4 + c] = fmaf(A[r * 4 + 0], B[0 * 4 + c],
                          fmaf(A[r * 4 + 1], B[1 * 4 + c], A[r * 4 + 2] * B[2 * 4 + c]));
    T[r * 4 + 3] = fmaf(A[r * 4 + 0], B[3],
                        fmaf(A[r * 4 + 1], B[7], fmaf(A[r * 4 + 2], B[11], A[r * 4 + 3])));
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) C[k] = T[k];
}

// p' = X * p
__device__ __forceinline__ void aff_point(const float *X, float px, float py, float pz,
                                          float &ox, float &oy, float &oz) {
  ox = fmaf(X[0], px, fmaf(X[1], py, fmaf(X[2], pz, X[3])));
  oy = fmaf(X[4], px, fmaf(X[5], py, fmaf(X[6], pz, X[7])));
  oz = fmaf(X[8], px, fmaf(X[9], py, fmaf(X[10], pz, X[11])));
}

// M = C * J(v) for one body (see DevBody)
__device__ __forceinline__ void joint_local(const DevBody &bd, float v, float *M) {
  if (bd.type == CSD_JOINT_REVOLUTE) {
    float s, c;
    sincosf(v, &s, &c);
    const float oc = 1.0f - c;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        M[r * 4 + k] = fmaf(s, bd.CK[r * 3 + k], fmaf(oc, bd.CAm[r * 3 + k], bd.C[r * 4 + k]));
      M[r * 4 + 3] = bd.C[r * 4 + 3];
    }
  } else {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int k = 0; k < 3; ++k) M[r * 4 + k] = bd.C[r * 4 + k];
      M[r * 4 + 3] = fmaf(v, bd.CK[r], bd.C[r * 4 + 3]);
    }
  }
}

// Wave-uniform dynamic index into a register array, as a select chain: every
// array access has a compile-time index, so the arrays stay scalar VGPRs
// (no scratch, no whole-vector copies).
template <int N>
__device__ __forceinline__ float pick(const float *v, int j) {
  float r = v[0];
#pragma unroll
  for (int k = 1; k < N; ++k) r = (j == k) ? v[k] : r;
  return r;
}

// Body frames of one configuration held in VGPRs; the parent frame of a
// body is fetched with a select chain on the (wave-uniform) body index.
template <int NB>
struct BodyFrames {
  float m[NB][12];
  __device__ __forceinline__ void get(int b, float *X) const {
#pragma unroll
    for (int k = 0; k < 12; ++k) X[k] = m[0][k];
#pragma unroll
    for (int c = 1; c < NB; ++c)
#pragma unroll
      for (int k = 0; k < 12; ++k) X[k] = (b == c) ? m[c][k] : X[k];
  }
  __device__ __forceinline__ void set(int b, const float *X) {
#pragma unroll
    for (int k = 0; k < 12; ++k) m[b][k] = X[k];  // b is a compile-time index here
  }
};
#define CSD_BODY_FRAMES(NAME, NB) ::csd::BodyFrames<NB> NAME

// Configuration accessors: element j of the lane's configuration, either
// from HBM with strides or from a register array (EI-ZO kernels generate
// configurations in-lane).
struct QMem {
  const float *__restrict__ q;
  int64_t base, sj;
  __device__ __forceinline__ float operator()(int j) const { return q[base + (int64_t)j * sj]; }
};
// Configuration held in registers, read with a wave-uniform index j: as a
// vector value the read lowers to indexed register addressing
// (s_set_gpr_idx_on + v_mov), never to a private-memory copy.
template <int N>
struct QReg {
  typedef float vec __attribute__((ext_vector_type(N)));
  vec v;
  __device__ __forceinline__ explicit QReg(const float (&a)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = a[k];
  }
  __device__ __forceinline__ float operator()(int j) const { return v[j]; }
};

// The first min(dof, N) joint values of a configuration in registers, all
// loads issued before FK needs any of them (FK then waits for memory once,
// not once per joint). Fast-path plants have dof <= N bodies.
template <int N>
__device__ __forceinline__ QReg<N> preload_q(const QMem &qm, int dof) {
  float a[N];
#pragma unroll
  for (int j = 0; j < N; ++j) a[j] = j < dof ? qm(j) : 0.f;
  return QReg<N>(a);
}

// Forward kinematics of all bodies of one configuration.
template <int NB, class Q>
__device__ __forceinline__ void fk_bodies(const PlantImage &P, const Q &qa, BodyFrames<NB> &B) {
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b < P.num_bodies) {
      const DevBody &bd = P.bodies[b];
      const float v = qa(bd.q_idx);
      float M[12];
      joint_local(bd, v, M);
      const int par = bd.parent;
      if (par < 0) {
        B.set(b, M);
      } else {
        float X[12], Y[12];
        B.get(par, X);
        aff_mul(X, M, Y);
        B.set(b, Y);
      }
    }
  }
}


// ---------------------------------------------------------------------------
// pair primitives (hit == true means colliding)
// ---------------------------------------------------------------------------
// sphere centre vs world-fixed box given as inverse pose + half extents
// (DevObstacle box record); collision iff d^2 < r^2.
__device__ __forceinline__ bool hit_sphere_obox(float cx, float cy, float cz, float r2,
                                                const float *o) {
  const float lx = fmaf(o[0], cx, fmaf(o[1], cy, fmaf(o[2], cz, o[9])));
  const float ly = fmaf(o[3], cx, fmaf(o[4], cy, fmaf(o[5], cz, o[10])));
  const float lz = fmaf(o[6], cx, fmaf(o[7], cy, fmaf(o[8], cz, o[11])));
  // per-axis distance outside the box, clamped to [0, 1]: one v_sub with the
  // clamp modifier per axis. The decision d2 < r2 is unchanged as long as
  // r2 < 1 (an axis distance above 1 already exceeds r); the fast path only
  // takes robot spheres of radius below 1 m.
  const float ex = fminf(fmaxf(fabsf(lx) - o[12], 0.f), 1.f);
  const float ey = fminf(fmaxf(fabsf(ly) - o[13], 0.f), 1.f);
  const float ez = fminf(fmaxf(fabsf(lz) - o[14], 0.f), 1.f);
  const float d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
  return d2 < r2;
}

__device__ __forceinline__ bool hit_sphere_sphere(float ax, float ay, float az, float bx,
                                                  float by, float bz, float rsum2) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  return fmaf(dx, dx, fmaf(dy, dy, dz * dz)) < rsum2;
}

// point vs segment p0 + t d (inv_dd = 1/|d|^2), collision iff d^2 < rsum^2
__device__ __forceinline__ float seg_point_d2(float cx, float cy, float cz, float p0x,
                                              float p0y, float p0z, float dx, float dy,
                                              float dz, float inv_dd) {
  const float wx = cx - p0x, wy = cy - p0y, wz = cz - p0z;
  float t = fmaf(wx, dx, fmaf(wy, dy, wz * dz)) * inv_dd;
  t = fminf(fmaxf(t, 0.f), 1.f);
  const float ex = fmaf(t, dx, p0x) - cx, ey = fmaf(t, dy, p0y) - cy,
              ez = fmaf(t, dz, p0z) - cz;
  return fmaf(ex, ex, fmaf(ey, ey, ez * ez));
}

// sphere vs box posed by X (3x4), half extents h: reference order
// R^T (c - p), clamp, squared distance (cuda_geometry_utilities.h:31-55)
__device__ __forceinline__ float sphere_box_d2(float cx, float cy, float cz, const float *X,
                                               const float *h) {
  const float dx = cx - X[3], dy = cy - X[7], dz = cz - X[11];
  float d2 = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float l = fmaf(X[0 * 4 + i], dx, fmaf(X[1 * 4 + i], dy, X[2 * 4 + i] * dz));
    const float e = fmaxf(fabsf(l) - h[i], 0.f);
    d2 = fmaf(e, e, d2);
  }
  return d2;
}

// 15-axis SAT (cuda_geometry_utilities.h:67-125); true if separated
__device__ __forceinline__ bool box_box_free(const float *XA, const float *hA, const float *XB,
                                             const float *hB) {
  float ax[15][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ax[i][k] = XA[k * 4 + i];
      ax[3 + i][k] = XB[k * 4 + i];
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float *u = ax[i], *v = ax[3 + j];
      float *c = ax[6 + 3 * i + j];
      c[0] = u[1] * v[2] - u[2] * v[1];
      c[1] = u[2] * v[0] - u[0] * v[2];
      c[2] = u[0] * v[1] - u[1] * v[0];
    }
  const float px = XA[3] - XB[3], py = XA[7] - XB[7], pz = XA[11] - XB[11];
  bool sep = false;
#pragma unroll
  for (int i = 0; i < 15; ++i) {
    const float *n = ax[i];
    const float lhs = fabsf(px * n[0] + py * n[1] + pz * n[2]);
    float rhs = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      rhs += fabsf(hA[k] * (ax[k][0] * n[0] + ax[k][1] * n[1] + ax[k][2] * n[2]));
      rhs += fabsf(hB[k] * (ax[3 + k][0] * n[0] + ax[3 + k][1] * n[1] + ax[3 + k][2] * n[2]));
    }
    sep |= lhs > rhs;
  }
  return sep;
}

// squared distance between segments (Ericson, Real-Time Collision Detection,
// section 5.1.9), FP32
__device__ __forceinline__ float seg_seg_d2(const float *p1, const float *d1, const float *p2,
                                            const float *d2) {
  const float r[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const float a = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
  const float e = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
  const float f = d2[0] * r[0] + d2[1] * r[1] + d2[2] * r[2];
  const float c = d1[0] * r[0] + d1[1] * r[1] + d1[2] * r[2];
  const float b = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
  float s, t;
  if (a <= 1e-30f && e <= 1e-30f) {
    s = t = 0.f;
  } else if (a <= 1e-30f) {
    s = 0.f;
    t = fminf(fmaxf(f / e, 0.f), 1.f);
  } else if (e <= 1e-30f) {
    t = 0.f;
    s = fminf(fmaxf(-c / a, 0.f), 1.f);
  } else {
    const float den = a * e - b * b;
    s = den > 0.f ? fminf(fmaxf((b * f - c * e) / den, 0.f), 1.f) : 0.f;
    t = (b * s + f) / e;
    if (t < 0.f) {
      t = 0.f;
      s = fminf(fmaxf(-c / a, 0.f), 1.f);
    } else if (t > 1.f) {
      t = 1.f;
      s = fminf(fmaxf((b - c) / a, 0.f), 1.f);
    }
  }
  float d = 0.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float x = (p1[k] + d1[k] * s) - (p2[k] + d2[k] * t);
    d = fmaf(x, x, d);
  }
  return d;
}

// squared distance from segment a + t d (box frame) to a box of half extents
// h: f(t) is convex, so bisect on the sign of f'(t).
__device__ __forceinline__ float seg_box_local_d2(const float *a, const float *d,
                                                  const float *h) {
  auto f_and_df = [&](float t, float &df) {
    float f = 0.f;
    df = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float x = fmaf(t, d[i], a[i]);
      const float e = fmaxf(fabsf(x) - h[i], 0.f);
      const float sg = x >= 0.f ? 1.f : -1.f;
      f = fmaf(e, e, f);
      df = fmaf(e * sg, d[i], df);
    }
    return f;
  };
  float lo = 0.f, hi = 1.f, df;
  f_and_df(0.f, df);
  if (df >= 0.f) return f_and_df(0.f, df);
  f_and_df(1.f, df);
  if (df <= 0.f) return f_and_df(1.f, df);
  for (int it = 0; it < 32; ++it) {
    const float mid = 0.5f * (lo + hi);
    f_and_df(mid, df);
    if (df > 0.f) hi = mid; else lo = mid;
  }
  return f_and_df(0.5f * (lo + hi), df);
}

// ---------------------------------------------------------------------------
// generic geometry: world pose and dispatch
// ---------------------------------------------------------------------------
// world pose of compiled geometry g (full 3x4) from the body frames
template <int NB>
__device__ __forceinline__ void geom_world(const DevGeom &g, const BodyFrames<NB> &B, float *X) {
  if (g.body < 0) {
#pragma unroll
    for (int k = 0; k < 12; ++k) X[k] = g.G[k];
  } else {
    float Y[12];
    B.get(g.body, Y);
    aff_mul(Y, g.G, X);
  }
}

// capsule segment endpoints in world: p0 = c - hl z, d = 2 hl z
__device__ __forceinline__ void capsule_seg(const float *X, float hl, float *p0, float *dd) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p0[k] = X[k * 4 + 3] - hl * X[k * 4 + 2];
    dd[k] = 2.f * hl * X[k * 4 + 2];
  }
}

// collision test between two posed geometries of ranks ra <= rb
// (rank: sphere 0, box 1, capsule 2); dims as in DevGeom
__device__ __forceinline__ bool hit_posed(int ra, const float *XA, const float *dA, int rb,
                                          const float *XB, const float *dB) {
  if (ra == 0 && rb == 0) {
    const float rs = dA[0] + dB[0];
    return hit_sphere_sphere(XA[3], XA[7], XA[11], XB[3], XB[7], XB[11], rs * rs);
  }
  if (ra == 0 && rb == 1) return sphere_box_d2(XA[3], XA[7], XA[11], XB, dB) < dA[0] * dA[0];
  if (ra == 1 && rb == 1) return !box_box_free(XA, dA, XB, dB);
  if (ra == 0 && rb == 2) {
    float p0[3], dd[3];
    capsule_seg(XB, dB[1], p0, dd);
    const float n2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
    const float inv = n2 > 0.f ? 1.f / n2 : 0.f;
    const float rs = dA[0] + dB[0];
    return seg_point_d2(XA[3], XA[7], XA[11], p0[0], p0[1], p0[2], dd[0], dd[1], dd[2], inv) <
           rs * rs;
  }
  if (ra == 2 && rb == 2) {
    float p1[3], d1[3], p2[3], d2[3];
    capsule_seg(XA, dA[1], p1, d1);
    capsule_seg(XB, dB[1], p2, d2);
    const float rs = dA[0] + dB[0];
    return seg_seg_d2(p1, d1, p2, d2) < rs * rs;
  }
  // box (A, rank 1) vs capsule (B, rank 2)
  float p0[3], dd[3], a[3], dl[3];
  capsule_seg(XB, dB[1], p0, dd);
  const float wx = p0[0] - XA[3], wy = p0[1] - XA[7], wz = p0[2] - XA[11];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    a[i] = fmaf(XA[0 * 4 + i], wx, fmaf(XA[1 * 4 + i], wy, XA[2 * 4 + i] * wz));
    dl[i] = fmaf(XA[0 * 4 + i], dd[0], fmaf(XA[1 * 4 + i], dd[1], XA[2 * 4 + i] * dd[2]));
  }
  return seg_box_local_d2(a, dl, dA) < dB[0] * dB[0];
}

__device__ __forceinline__ int geom_rank(int type) {
  return type == CSD_SPHERE ? 0 : (type == CSD_BOX ? 1 : 2);
}

}  // namespace csd
