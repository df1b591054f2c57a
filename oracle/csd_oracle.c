/*
 * csd_oracle.c -- CPU restatement of csdecomp's FK / collision / EI-ZO path.
 *
 * TEST INFRASTRUCTURE ONLY: used as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg. The product
 * (csdecomp_amd/) never links or calls it.
 *
 * Built with -ffp-contract=off so that every float op is a separately
 * rounded IEEE operation, in the order the reference writes it.
 */
#include "csd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define M4(X, r, c) ((X)[(c)*4 + (r)]) /* column-major 4x4 */

/* ------------------------------------------------------------------------ */
/* small matrix helpers                                                      */
/* ------------------------------------------------------------------------ */
static void mat4_mul_f(const float *A, const float *B, float *C) {
  /* C = A*B, sum over k in increasing order (Eigen lazy product order) */
  float T[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      float s = M4(A, r, 0) * M4(B, 0, c);
      s = s + M4(A, r, 1) * M4(B, 1, c);
      s = s + M4(A, r, 2) * M4(B, 2, c);
      s = s + M4(A, r, 3) * M4(B, 3, c);
      T[c * 4 + r] = s;
    }
  memcpy(C, T, sizeof(T));
}
static void mat4_mul_d(const double *A, const double *B, double *C) {
  double T[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += M4(A, r, k) * M4(B, k, c);
      T[c * 4 + r] = s;
    }
  memcpy(C, T, sizeof(T));
}
static void mat4_identity_f(float *X) {
  memset(X, 0, 16 * sizeof(float));
  X[0] = X[5] = X[10] = X[15] = 1.0f;
}
static void mat4_identity_d(double *X) {
  memset(X, 0, 16 * sizeof(double));
  X[0] = X[5] = X[10] = X[15] = 1.0;
}

/* Rotation about a unit axis, in the term order of Eigen's
 * AngleAxis::toRotationMatrix (Rodrigues' formula). Writes the 3x3 block of a
 * column-major 4x4. */
static void axis_angle_f(float q, const float *a, float *X) {
  float s = sinf(q), c = cosf(q);
  float sx = s * a[0], sy = s * a[1], sz = s * a[2];
  float one_c = 1.0f - c;
  float cx = one_c * a[0], cy = one_c * a[1], cz = one_c * a[2];
  float t;
  t = cx * a[1];
  M4(X, 0, 1) = t - sz;
  M4(X, 1, 0) = t + sz;
  t = cx * a[2];
  M4(X, 0, 2) = t + sy;
  M4(X, 2, 0) = t - sy;
  t = cy * a[2];
  M4(X, 1, 2) = t - sx;
  M4(X, 2, 1) = t + sx;
  M4(X, 0, 0) = cx * a[0] + c;
  M4(X, 1, 1) = cy * a[1] + c;
  M4(X, 2, 2) = cz * a[2] + c;
}
static void axis_angle_d(double q, const double *a, double *X) {
  double s = sin(q), c = cos(q), one_c = 1.0 - c;
  M4(X, 0, 0) = one_c * a[0] * a[0] + c;
  M4(X, 1, 1) = one_c * a[1] * a[1] + c;
  M4(X, 2, 2) = one_c * a[2] * a[2] + c;
  M4(X, 0, 1) = one_c * a[0] * a[1] - s * a[2];
  M4(X, 1, 0) = one_c * a[0] * a[1] + s * a[2];
  M4(X, 0, 2) = one_c * a[0] * a[2] + s * a[1];
  M4(X, 2, 0) = one_c * a[0] * a[2] - s * a[1];
  M4(X, 1, 2) = one_c * a[1] * a[2] - s * a[0];
  M4(X, 2, 1) = one_c * a[1] * a[2] + s * a[0];
}

/* ------------------------------------------------------------------------ */
/* forward kinematics -- minimal_kinematic_tree.cpp:10-58                    */
/* ------------------------------------------------------------------------ */
void orc_fk_f32(const orc_plant *p, const float *q, float *X) {
  mat4_identity_f(X); /* root (link 0) is identity, :22 */
  for (int t = 0; t < p->num_joints; ++t) {
    int j = p->traversal[t];
    int parent = p->joint_parent_link[j], child = p->joint_child_link[j];
    float XJL[16];
    mat4_identity_f(XJL);
    float v = 0.0f;
    if (p->joint_type[j] != ORC_FIXED) v = q[p->joint_config_idx[j]];
    const float *axis = p->joint_axis + 3 * j;
    if (p->joint_type[j] == ORC_REVOLUTE) {
      axis_angle_f(v, axis, XJL);
    } else if (p->joint_type[j] == ORC_PRISMATIC) {
      M4(XJL, 0, 3) = M4(XJL, 0, 3) + axis[0] * v;
      M4(XJL, 1, 3) = M4(XJL, 1, 3) + axis[1] * v;
      M4(XJL, 2, 3) = M4(XJL, 2, 3) + axis[2] * v;
    }
    float T[16];
    /* X_W_L = X_W_PL * X_PL_J * X_J_L, evaluated left to right (:55) */
    mat4_mul_f(X + 16 * parent, p->joint_X_PL_J + 16 * j, T);
    mat4_mul_f(T, XJL, X + 16 * child);
  }
}

void orc_fk_f64(const orc_plant *p, const double *q, double *X) {
  mat4_identity_d(X);
  for (int t = 0; t < p->num_joints; ++t) {
    int j = p->traversal[t];
    int parent = p->joint_parent_link[j], child = p->joint_child_link[j];
    double XJL[16], XPJ[16], axis[3];
    mat4_identity_d(XJL);
    for (int k = 0; k < 16; ++k) XPJ[k] = p->joint_X_PL_J[16 * j + k];
    for (int k = 0; k < 3; ++k) axis[k] = p->joint_axis[3 * j + k];
    double v = 0.0;
    if (p->joint_type[j] != ORC_FIXED) v = q[p->joint_config_idx[j]];
    if (p->joint_type[j] == ORC_REVOLUTE) {
      axis_angle_d(v, axis, XJL);
    } else if (p->joint_type[j] == ORC_PRISMATIC) {
      for (int k = 0; k < 3; ++k) M4(XJL, k, 3) = axis[k] * v;
    }
    double T[16];
    mat4_mul_d(X + 16 * parent, XPJ, T);
    mat4_mul_d(T, XJL, X + 16 * child);
  }
}

/* ------------------------------------------------------------------------ */
/* FP32 pair primitives                                                      */
/* ------------------------------------------------------------------------ */
static float dot3f(const float *a, const float *b) {
  float s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}
static void cross3f(const float *a, const float *b, float *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* sphere-box: box-local centre, clamp, squared distance (shared by both
 * rules; collision_checker.cpp:170-199, cuda_geometry_utilities.h:31-55) */
static float sphere_box_d2_f(const float *c, const float *XB, const float *dims) {
  float d[3] = {c[0] - M4(XB, 0, 3), c[1] - M4(XB, 1, 3), c[2] - M4(XB, 2, 3)};
  float loc[3];
  for (int i = 0; i < 3; ++i) {
    /* R^T * d : row i of R^T is column i of R */
    float s = M4(XB, 0, i) * d[0];
    s = s + M4(XB, 1, i) * d[1];
    s = s + M4(XB, 2, i) * d[2];
    loc[i] = s;
  }
  float d2 = 0.0f;
  for (int i = 0; i < 3; ++i) {
    float h = dims[i] * 0.5f;
    float cl = fmaxf(-h, fminf(loc[i], h));
    float e = cl - loc[i];
    d2 = (i == 0) ? e * e : d2 + e * e;
  }
  return d2;
}

/* 15-axis SAT; returns 1 if a separating axis exists
 * (collision_checker.cpp:118-167, cuda_geometry_utilities.h:67-125) */
static int box_box_sat_free_f(const float *XA, const float *dA, const float *XB,
                              const float *dB) {
  float ax[15][3];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      ax[i][k] = M4(XA, k, i);
      ax[3 + i][k] = M4(XB, k, i);
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) cross3f(ax[i], ax[3 + j], ax[6 + 3 * i + j]);
  float pAB[3] = {M4(XA, 0, 3) - M4(XB, 0, 3), M4(XA, 1, 3) - M4(XB, 1, 3),
                  M4(XA, 2, 3) - M4(XB, 2, 3)};
  for (int i = 0; i < 15; ++i) {
    const double lhs = fabsf(dot3f(pAB, ax[i]));
    /* the reference sums six float terms in float, then widens */
    float r = fabsf((dA[0] / 2) * dot3f(ax[0], ax[i]));
    r = r + fabsf((dA[1] / 2) * dot3f(ax[1], ax[i]));
    r = r + fabsf((dA[2] / 2) * dot3f(ax[2], ax[i]));
    r = r + fabsf((dB[0] / 2) * dot3f(ax[3], ax[i]));
    r = r + fabsf((dB[1] / 2) * dot3f(ax[4], ax[i]));
    r = r + fabsf((dB[2] / 2) * dot3f(ax[5], ax[i]));
    const double rhs = r;
    if (lhs > rhs) return 1;
  }
  return 0;
}

/* ---- capsule helpers (no reference counterpart: csdecomp throws on
 * capsules, collision_checker.cpp:104-107). A capsule is the set of points
 * within radius r of the segment (0,0,-L/2)-(0,0,+L/2) of its frame. ---- */
static void capsule_ends_d(const double *X, double L, double *p0, double *p1) {
  for (int k = 0; k < 3; ++k) {
    p0[k] = M4(X, k, 3) - 0.5 * L * M4(X, k, 2);
    p1[k] = M4(X, k, 3) + 0.5 * L * M4(X, k, 2);
  }
}
/* squared distance point-segment */
static double pt_seg_d2(const double *c, const double *p0, const double *p1) {
  double d[3], w[3], dd = 0, wd = 0;
  for (int k = 0; k < 3; ++k) {
    d[k] = p1[k] - p0[k];
    w[k] = c[k] - p0[k];
    dd += d[k] * d[k];
    wd += w[k] * d[k];
  }
  double t = dd > 0 ? wd / dd : 0.0;
  if (t < 0) t = 0;
  if (t > 1) t = 1;
  double s = 0;
  for (int k = 0; k < 3; ++k) {
    double e = p0[k] + t * d[k] - c[k];
    s += e * e;
  }
  return s;
}
/* squared distance segment-segment (Ericson, Real-Time Collision Detection,
 * section 5.1.9, ClosestPtSegmentSegment) */
static double seg_seg_d2(const double *p1, const double *q1, const double *p2,
                         const double *q2) {
  double d1[3], d2[3], r[3];
  for (int k = 0; k < 3; ++k) {
    d1[k] = q1[k] - p1[k];
    d2[k] = q2[k] - p2[k];
    r[k] = p1[k] - p2[k];
  }
  double a = 0, e = 0, f = 0, c = 0, b = 0;
  for (int k = 0; k < 3; ++k) {
    a += d1[k] * d1[k];
    e += d2[k] * d2[k];
    f += d2[k] * r[k];
    c += d1[k] * r[k];
    b += d1[k] * d2[k];
  }
  double s, t;
  const double eps = 1e-300;
  if (a <= eps && e <= eps) {
    s = t = 0;
  } else if (a <= eps) {
    s = 0;
    t = f / e;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
  } else if (e <= eps) {
    t = 0;
    s = -c / a;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
  } else {
    double denom = a * e - b * b;
    s = denom > 0 ? (b * f - c * e) / denom : 0.0;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
    t = (b * s + f) / e;
    if (t < 0) {
      t = 0;
      s = -c / a;
      s = s < 0 ? 0 : (s > 1 ? 1 : s);
    } else if (t > 1) {
      t = 1;
      s = (b - c) / a;
      s = s < 0 ? 0 : (s > 1 ? 1 : s);
    }
  }
  double s2 = 0;
  for (int k = 0; k < 3; ++k) {
    double ee = (p1[k] + d1[k] * s) - (p2[k] + d2[k] * t);
    s2 += ee * ee;
  }
  return s2;
}

/* signed distance from a point (box frame) to a box with half extents h */
static double pt_box_sd_local(const double *l, const double *h) {
  double out2 = 0, inside = -1e300;
  int outside = 0;
  for (int i = 0; i < 3; ++i) {
    double e = fabs(l[i]) - h[i];
    if (e > 0) {
      out2 += e * e;
      outside = 1;
    }
    if (e > inside) inside = e;
  }
  return outside ? sqrt(out2) : inside;
}
static void to_local_d(const double *X, const double *w, double *l) {
  double d[3] = {w[0] - M4(X, 0, 3), w[1] - M4(X, 1, 3), w[2] - M4(X, 2, 3)};
  for (int i = 0; i < 3; ++i)
    l[i] = M4(X, 0, i) * d[0] + M4(X, 1, i) * d[1] + M4(X, 2, i) * d[2];
}
/* signed distance of a segment to a box: min over t of the (convex) signed
 * distance of p(t); golden-section search in double. */
static double seg_box_sd(const double *p0w, const double *p1w, const double *XB,
                         const double *h) {
  double a[3], b[3];
  to_local_d(XB, p0w, a);
  to_local_d(XB, p1w, b);
  double lo = 0.0, hi = 1.0;
  const double g = 0.6180339887498949;
  double x1 = hi - g * (hi - lo), x2 = lo + g * (hi - lo);
#define SEGPT(t, o)                                          \
  do {                                                       \
    for (int k = 0; k < 3; ++k) (o)[k] = a[k] + (t) * (b[k] - a[k]); \
  } while (0)
  double pt[3];
  SEGPT(x1, pt);
  double f1 = pt_box_sd_local(pt, h);
  SEGPT(x2, pt);
  double f2 = pt_box_sd_local(pt, h);
  for (int it = 0; it < 120; ++it) {
    if (f1 <= f2) {
      hi = x2;
      x2 = x1;
      f2 = f1;
      x1 = hi - g * (hi - lo);
      SEGPT(x1, pt);
      f1 = pt_box_sd_local(pt, h);
    } else {
      lo = x1;
      x1 = x2;
      f1 = f2;
      x2 = lo + g * (hi - lo);
      SEGPT(x2, pt);
      f2 = pt_box_sd_local(pt, h);
    }
  }
  double best = f1 < f2 ? f1 : f2;
  SEGPT(0.0, pt);
  double f0 = pt_box_sd_local(pt, h);
  SEGPT(1.0, pt);
  double fe = pt_box_sd_local(pt, h);
#undef SEGPT
  if (f0 < best) best = f0;
  if (fe < best) best = fe;
  return best;
}

/* ------------------------------------------------------------------------ */
/* pair dispatch                                                             */
/* ------------------------------------------------------------------------ */
static void geom_world_f(const orc_plant *p, int g, const float *X_W_L,
                         float *X) {
  mat4_mul_f(X_W_L, p->geom_X_L_B + 16 * g, X); /* X_W_L * X_L_B */
}
static void geom_world_d(const orc_plant *p, int g, const double *X_W_L,
                         double *X) {
  double XLB[16];
  for (int k = 0; k < 16; ++k) XLB[k] = p->geom_X_L_B[16 * g + k];
  mat4_mul_d(X_W_L, XLB, X);
}

static double pair_sd_world(int tA, const double *XA, const double *dA, int tB,
                            const double *XB, const double *dB);

/* FP32 capsule tests are computed in double from the FP32 geometry poses:
 * capsules have no reference code to mirror (see above). */
static int capsule_pair_free_f(int tA, const float *XA, const float *dA,
                               int tB, const float *XB, const float *dB) {
  double xa[16], xb[16], da[3], db[3];
  for (int k = 0; k < 16; ++k) {
    xa[k] = XA[k];
    xb[k] = XB[k];
  }
  for (int k = 0; k < 3; ++k) {
    da[k] = dA[k];
    db[k] = dB[k];
  }
  return pair_sd_world(tA, xa, da, tB, xb, db) >= 0.0;
}

int orc_pair_free_f32(const orc_plant *p, int gA, const float *X_W_LA, int gB,
                      const float *X_W_LB, int rule) {
  float XA[16], XB[16];
  geom_world_f(p, gA, X_W_LA, XA);
  geom_world_f(p, gB, X_W_LB, XB);
  int tA = p->geom_type[gA], tB = p->geom_type[gB];
  const float *dA = p->geom_dims + 3 * gA, *dB = p->geom_dims + 3 * gB;
  float posA[3] = {M4(XA, 0, 3), M4(XA, 1, 3), M4(XA, 2, 3)};
  float posB[3] = {M4(XB, 0, 3), M4(XB, 1, 3), M4(XB, 2, 3)};
  if (tA == ORC_CAPSULE || tB == ORC_CAPSULE)
    return capsule_pair_free_f(tA, XA, dA, tB, XB, dB);
  if (tA == ORC_SPHERE && tB == ORC_SPHERE) {
    float dx = posA[0] - posB[0], dy = posA[1] - posB[1], dz = posA[2] - posB[2];
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    float rs = dA[0] + dB[0];
    if (rule == ORC_RULE_CPU) return sqrtf(d2) > rs; /* :111-115 */
    return d2 >= rs * rs;                            /* cuda_geom :16-20 */
  }
  if (tA == ORC_BOX && tB == ORC_BOX) {
    if (rule == ORC_RULE_CPU) { /* bounding-sphere shortcut :118-123 */
      float maxlen = dA[0];
      for (int i = 0; i < 3; ++i) {
        if (dA[i] > maxlen) maxlen = dA[i];
        if (dB[i] > maxlen) maxlen = dB[i];
      }
      float dx = posA[0] - posB[0], dy = posA[1] - posB[1],
            dz = posA[2] - posB[2];
      float d2 = dx * dx;
      d2 = d2 + dy * dy;
      d2 = d2 + dz * dz;
      if ((double)sqrtf(d2) > 2 * sqrt(3.0) * maxlen) return 1;
    }
    return box_box_sat_free_f(XA, dA, XB, dB);
  }
  /* sphere-box */
  const float *Xs = (tA == ORC_SPHERE) ? XA : XB;
  const float *Xb = (tA == ORC_SPHERE) ? XB : XA;
  const float *ds = (tA == ORC_SPHERE) ? dA : dB;
  const float *db = (tA == ORC_SPHERE) ? dB : dA;
  float c[3] = {M4(Xs, 0, 3), M4(Xs, 1, 3), M4(Xs, 2, 3)};
  float d2 = sphere_box_d2_f(c, Xb, db);
  float r2 = ds[0] * ds[0];
  if (rule == ORC_RULE_CPU) return d2 > r2; /* :198 */
  return d2 >= r2;                          /* cuda_geom :54 */
}

int orc_geom_voxel_free_f32(const orc_plant *p, int g, const float *X_W_L,
                            const float *vox, float voxel_radius, int rule) {
  float X[16];
  geom_world_f(p, g, X_W_L, X);
  int t = p->geom_type[g];
  const float *d = p->geom_dims + 3 * g;
  if (t == ORC_CAPSULE) {
    double xc[16], dc[3], xv[16], dv[3] = {voxel_radius, -1, -1};
    for (int k = 0; k < 16; ++k) xc[k] = X[k];
    for (int k = 0; k < 3; ++k) dc[k] = d[k];
    memset(xv, 0, sizeof(xv));
    xv[0] = xv[5] = xv[10] = xv[15] = 1;
    for (int k = 0; k < 3; ++k) M4(xv, k, 3) = vox[k];
    return pair_sd_world(ORC_CAPSULE, xc, dc, ORC_SPHERE, xv, dv) >= 0.0;
  }
  if (t == ORC_SPHERE) {
    float dx = M4(X, 0, 3) - vox[0], dy = M4(X, 1, 3) - vox[1],
          dz = M4(X, 2, 3) - vox[2];
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    float rs = d[0] + voxel_radius;
    if (rule == ORC_RULE_CPU) return sqrtf(d2) > rs; /* :223-227 */
    return d2 >= rs * rs;
  }
  float d2 = sphere_box_d2_f(vox, X, d);
  float r2 = voxel_radius * voxel_radius;
  if (rule == ORC_RULE_CPU) return d2 > r2; /* :248-250 */
  return d2 >= r2;
}

/* ------------------------------------------------------------------------ */
/* FP64 signed distances                                                     */
/* ------------------------------------------------------------------------ */
static double box_box_sd(const double *XA, const double *dA, const double *XB,
                         const double *dB) {
  /* max over the 15 SAT axes (normalised) of projected gap. Exact
   * penetration depth when overlapping; a lower bound of the distance when
   * separated (the band test only needs |sd| >= 1e-5 to be certain). */
  double ax[15][3];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      ax[i][k] = M4(XA, k, i);
      ax[3 + i][k] = M4(XB, k, i);
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double *c = ax[6 + 3 * i + j];
      const double *u = ax[i], *v = ax[3 + j];
      c[0] = u[1] * v[2] - u[2] * v[1];
      c[1] = u[2] * v[0] - u[0] * v[2];
      c[2] = u[0] * v[1] - u[1] * v[0];
    }
  double pAB[3];
  for (int k = 0; k < 3; ++k) pAB[k] = M4(XA, k, 3) - M4(XB, k, 3);
  double best = -1e300;
  for (int i = 0; i < 15; ++i) {
    double n = sqrt(ax[i][0] * ax[i][0] + ax[i][1] * ax[i][1] +
                    ax[i][2] * ax[i][2]);
    if (n < 1e-9) continue;
    double u[3] = {ax[i][0] / n, ax[i][1] / n, ax[i][2] / n};
    double lhs = fabs(pAB[0] * u[0] + pAB[1] * u[1] + pAB[2] * u[2]);
    double rhs = 0;
    for (int k = 0; k < 3; ++k) {
      rhs += 0.5 * dA[k] * fabs(ax[k][0] * u[0] + ax[k][1] * u[1] + ax[k][2] * u[2]);
      rhs += 0.5 * dB[k] *
             fabs(ax[3 + k][0] * u[0] + ax[3 + k][1] * u[1] + ax[3 + k][2] * u[2]);
    }
    double gap = lhs - rhs;
    if (gap > best) best = gap;
  }
  return best;
}

static double pair_sd_world(int tA, const double *XA, const double *dA, int tB,
                            const double *XB, const double *dB) {
  if (tA == ORC_BOX && tB == ORC_BOX) return box_box_sd(XA, dA, XB, dB);
  if (tA == ORC_SPHERE && tB == ORC_SPHERE) {
    double s = 0;
    for (int k = 0; k < 3; ++k) {
      double e = M4(XA, k, 3) - M4(XB, k, 3);
      s += e * e;
    }
    return sqrt(s) - dA[0] - dB[0];
  }
  if ((tA == ORC_SPHERE && tB == ORC_BOX) || (tA == ORC_BOX && tB == ORC_SPHERE)) {
    const double *Xs = tA == ORC_SPHERE ? XA : XB, *Xb = tA == ORC_SPHERE ? XB : XA;
    const double *ds = tA == ORC_SPHERE ? dA : dB, *db = tA == ORC_SPHERE ? dB : dA;
    double c[3] = {M4(Xs, 0, 3), M4(Xs, 1, 3), M4(Xs, 2, 3)}, l[3];
    double h[3] = {0.5 * db[0], 0.5 * db[1], 0.5 * db[2]};
    to_local_d(Xb, c, l);
    return pt_box_sd_local(l, h) - ds[0];
  }
  /* capsule cases */
  if (tB == ORC_CAPSULE && tA != ORC_CAPSULE) {
    const double *t1 = XA, *t2 = dA;
    int tt = tA;
    XA = XB;
    dA = dB;
    tA = tB;
    XB = t1;
    dB = t2;
    tB = tt;
  }
  double a0[3], a1[3];
  capsule_ends_d(XA, dA[1], a0, a1);
  if (tB == ORC_SPHERE) {
    double c[3] = {M4(XB, 0, 3), M4(XB, 1, 3), M4(XB, 2, 3)};
    return sqrt(pt_seg_d2(c, a0, a1)) - dA[0] - dB[0];
  }
  if (tB == ORC_CAPSULE) {
    double b0[3], b1[3];
    capsule_ends_d(XB, dB[1], b0, b1);
    return sqrt(seg_seg_d2(a0, a1, b0, b1)) - dA[0] - dB[0];
  }
  /* capsule-box */
  double h[3] = {0.5 * dB[0], 0.5 * dB[1], 0.5 * dB[2]};
  return seg_box_sd(a0, a1, XB, h) - dA[0];
}

double orc_pair_sd_f64(const orc_plant *p, int gA, const double *X_W_LA, int gB,
                       const double *X_W_LB) {
  double XA[16], XB[16], dA[3], dB[3];
  geom_world_d(p, gA, X_W_LA, XA);
  geom_world_d(p, gB, X_W_LB, XB);
  for (int k = 0; k < 3; ++k) {
    dA[k] = p->geom_dims[3 * gA + k];
    dB[k] = p->geom_dims[3 * gB + k];
  }
  return pair_sd_world(p->geom_type[gA], XA, dA, p->geom_type[gB], XB, dB);
}

double orc_geom_voxel_sd_f64(const orc_plant *p, int g, const double *X_W_L,
                             const double *vox, double voxel_radius) {
  double X[16], d[3], XV[16], dv[3] = {voxel_radius, -1, -1};
  geom_world_d(p, g, X_W_L, X);
  for (int k = 0; k < 3; ++k) d[k] = p->geom_dims[3 * g + k];
  mat4_identity_d(XV);
  for (int k = 0; k < 3; ++k) M4(XV, k, 3) = vox[k];
  return pair_sd_world(p->geom_type[g], X, d, ORC_SPHERE, XV, dv);
}

/* ------------------------------------------------------------------------ */
/* per-config drivers                                                        */
/* ------------------------------------------------------------------------ */
#define ORC_MAX_LINKS 128

/* checkCollisionFreeVoxels semantics: early exit, no printing
 * (collision_checker.cpp:55-89); num_voxels = 0 gives checkCollisionFree */
static int check_one(const orc_plant *p, const float *q, const float *voxels,
                     int32_t nv, float vr, const int32_t *rg, int32_t nrg,
                     int rule) {
  float X[ORC_MAX_LINKS * 16];
  orc_fk_f32(p, q, X);
  for (int k = 0; k < p->num_pairs; ++k) {
    int a = p->pairs[2 * k], b = p->pairs[2 * k + 1];
    if (!orc_pair_free_f32(p, a, X + 16 * p->geom_link[a], b,
                           X + 16 * p->geom_link[b], rule))
      return 0;
  }
  for (int r = 0; r < nrg; ++r) {
    int g = rg[r];
    for (int v = 0; v < nv; ++v)
      if (!orc_geom_voxel_free_f32(p, g, X + 16 * p->geom_link[g],
                                   voxels + 3 * v, vr, rule))
        return 0;
  }
  return 1;
}

static int set_threads(int nthreads) {
#ifdef _OPENMP
  if (omp_in_parallel()) return 1;  /* inside orc_eizo_batch's region loop */
  if (nthreads > 0) return nthreads;
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

void orc_check_batch(const orc_plant *p, const float *q, int64_t n, int rule,
                     uint8_t *out, int nthreads) {
  int nt = set_threads(nthreads);
  const int dof = p->num_config_vars;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < n; ++i)
    out[i] = (uint8_t)check_one(p, q + i * dof, NULL, 0, 0.f, NULL, 0, rule);
}

void orc_check_voxels_batch(const orc_plant *p, const float *q, int64_t n,
                            const float *voxels, int32_t num_voxels,
                            float voxel_radius, const int32_t *robot_geoms,
                            int32_t num_robot_geoms, int rule, uint8_t *out,
                            int nthreads) {
  int nt = set_threads(nthreads);
  const int dof = p->num_config_vars;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < n; ++i)
    out[i] = (uint8_t)check_one(p, q + i * dof, voxels, num_voxels,
                                voxel_radius, robot_geoms, num_robot_geoms,
                                rule);
}

/* checkCollisionFreeVoxelsWithoutSelfCollisionCuda semantics
 * (cuda_collision_checker.cu:109-132,310-375): per (config, voxel), 1 iff no
 * robot geometry touches the voxel. */
void orc_voxel_flags_batch(const orc_plant *p, const float *q, int64_t n,
                           const float *voxels, int32_t num_voxels,
                           float voxel_radius, const int32_t *robot_geoms,
                           int32_t num_robot_geoms, int rule, uint8_t *out,
                           int nthreads) {
  int nt = set_threads(nthreads);
  const int dof = p->num_config_vars;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < n; ++i) {
    float X[ORC_MAX_LINKS * 16];
    orc_fk_f32(p, q + i * dof, X);
    for (int v = 0; v < num_voxels; ++v) {
      uint8_t f = 1;
      for (int r = 0; r < num_robot_geoms && f; ++r) {
        int g = robot_geoms[r];
        f = (uint8_t)orc_geom_voxel_free_f32(p, g, X + 16 * p->geom_link[g],
                                             voxels + 3 * v, voxel_radius, rule);
      }
      out[i * num_voxels + v] = f;
    }
  }
}

void orc_min_sd_batch(const orc_plant *p, const float *q, int64_t n,
                      const float *voxels, int32_t num_voxels,
                      float voxel_radius, const int32_t *robot_geoms,
                      int32_t num_robot_geoms, double *out, int nthreads) {
  int nt = set_threads(nthreads);
  const int dof = p->num_config_vars;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < n; ++i) {
    double qd[64], X[ORC_MAX_LINKS * 16];
    for (int k = 0; k < dof; ++k) qd[k] = q[i * dof + k];
    orc_fk_f64(p, qd, X);
    double best = 1e300;
    for (int k = 0; k < p->num_pairs; ++k) {
      int a = p->pairs[2 * k], b = p->pairs[2 * k + 1];
      double s = orc_pair_sd_f64(p, a, X + 16 * p->geom_link[a], b,
                                 X + 16 * p->geom_link[b]);
      if (s < best) best = s;
    }
    for (int r = 0; r < num_robot_geoms; ++r) {
      int g = robot_geoms[r];
      for (int v = 0; v < num_voxels; ++v) {
        double vx[3] = {voxels[3 * v], voxels[3 * v + 1], voxels[3 * v + 2]};
        double s = orc_geom_voxel_sd_f64(p, g, X + 16 * p->geom_link[g], vx,
                                         voxel_radius);
        if (s < best) best = s;
      }
    }
    out[i] = best;
  }
}

void orc_fk_error_batch(const orc_plant *p, const float *q, int64_t n,
                        double *out, int nthreads) {
  int nt = set_threads(nthreads);
  const int dof = p->num_config_vars, L = p->num_links;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < n; ++i) {
    float Xf[ORC_MAX_LINKS * 16];
    double qd[64], Xd[ORC_MAX_LINKS * 16];
    for (int k = 0; k < dof; ++k) qd[k] = q[i * dof + k];
    orc_fk_f32(p, q + i * dof, Xf);
    orc_fk_f64(p, qd, Xd);
    for (int l = 0; l < L; ++l) {
      double e = 0;
      for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r) {
          double d = fabs((double)Xf[16 * l + 4 * c + r] - Xd[16 * l + 4 * c + r]);
          if (d > e) e = d;
        }
      out[i * L + l] = e;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* edges                                                                     */
/* ------------------------------------------------------------------------ */
static float norm_diff_f(const float *a, const float *b, int n) {
  float s = 0.f;
  for (int k = 0; k < n; ++k) {
    float d = b[k] - a[k];
    s = s + d * d;
  }
  return sqrtf(s);
}

int32_t orc_edge_num_steps(const float *q1, const float *q2, int32_t dof,
                           float step) {
  return (int32_t)ceilf(norm_diff_f(q1, q2, dof) / step);
}

/* collision_checker.cpp:41-53: t = k / n for k in [0, n) -- q1 included,
 * q2 excluded, every interior point checked with the CPU driver. */
int orc_check_edge(const orc_plant *p, const float *q1, const float *q2,
                   float step, int rule) {
  const int dof = p->num_config_vars;
  int n = orc_edge_num_steps(q1, q2, dof, step);
  float q[64];
  for (int k = 0; k < n; ++k) {
    float t = (float)(k / (1.0 * n));
    for (int d = 0; d < dof; ++d) q[d] = t * q2[d] + (1 - t) * q1[d];
    if (!check_one(p, q, NULL, 0, 0.f, NULL, 0, rule)) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                         */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* 24-bit uniform in (0,1): exact in float, never 0 or 1 */
static float u01(uint32_t x) {
  return (float)(x >> 8) * (1.0f / 16777216.0f) + (1.0f / 33554432.0f);
}

/* ------------------------------------------------------------------------ */
/* hit-and-run -- cuda_hit_and_run_sampling.cu:74-112                        */
/* ------------------------------------------------------------------------ */
#define ORC_MAX_DIM 64
/* one chain of `mixing_steps` steps from x (in place); its randomness comes
 * from counters (chain, step, stream, round*16 + call) */
static void hnr_chain(const float *A, const float *b, int32_t F, int32_t dim,
                      float *x, uint32_t chain, int32_t mixing_steps,
                      uint64_t seed, uint32_t stream, uint32_t round) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const int npairs = (dim + 1) / 2;
  const int nuni = 2 * npairs + 1; /* normals + one for theta */
  const int ncalls = (nuni + 3) / 4;
  {
    const uint32_t i = chain;
    float dir[ORC_MAX_DIM + 1], u[32];
    for (int s = 0; s < mixing_steps; ++s) {
      for (int c = 0; c < ncalls; ++c) {
        uint32_t ctr[4] = {i, (uint32_t)s, stream, round * 16u + (uint32_t)c};
        uint32_t o[4];
        orc_philox4x32_10(ctr, key, o);
        for (int k = 0; k < 4; ++k) u[4 * c + k] = u01(o[k]);
      }
      /* Box-Muller on consecutive uniform pairs */
      for (int k = 0; k < npairs; ++k) {
        float r = sqrtf(-2.0f * logf(u[2 * k]));
        float ang = 6.2831853071795864769f * u[2 * k + 1];
        dir[2 * k] = r * cosf(ang);
        dir[2 * k + 1] = r * sinf(ang);
      }
      float theta_max = FLT_MAX, theta_min = -FLT_MAX;
      for (int f = 0; f < F; ++f) {
        float ax = 0.f, ad = 0.f;
        for (int d = 0; d < dim; ++d) {
          ax = ax + A[f * dim + d] * x[d];
          ad = ad + A[f * dim + d] * dir[d];
        }
        float lb = b[f] - ax;
        if (ad < 0.0f) {
          float t = lb / ad;
          if (t > theta_min) theta_min = t;
        } else if (ad > 0.0f) {
          float t = lb / ad;
          if (t < theta_max) theta_max = t;
        }
      }
      if (isinf(theta_max) || isinf(theta_min) || theta_max < theta_min)
        continue; /* failed step: the chain stays where it is */
      float theta = (float)((double)u[2 * npairs] *
                                (0.999999 * (double)theta_max - (double)theta_min) +
                            (double)theta_min);
      for (int d = 0; d < dim; ++d) x[d] = x[d] + theta * dir[d];
    }
  }
}

void orc_hit_and_run(const float *A, const float *b, int32_t F, int32_t dim,
                     const float *start, int64_t num_chains,
                     int64_t chain_offset, int32_t mixing_steps, uint64_t seed,
                     uint32_t stream, uint32_t round, float *out) {
  for (int64_t i = 0; i < num_chains; ++i) {
    float *x = out + i * dim;
    for (int d = 0; d < dim; ++d) x[d] = start[d];
    hnr_chain(A, b, F, dim, x, (uint32_t)(chain_offset + i), mixing_steps, seed,
              stream, round);
  }
}

/* chains spread along the segment ls -> le: chain c of `total` starts at
 * fraction (c + 1/2) / total (fmaf(f, le - ls, ls)) */
void orc_hit_and_run_segment(const float *A, const float *b, int32_t F,
                             int32_t dim, const float *ls, const float *le,
                             int64_t num_chains, int64_t chain_offset,
                             int64_t total, int32_t mixing_steps, uint64_t seed,
                             uint32_t stream, uint32_t round, float *out) {
  for (int64_t i = 0; i < num_chains; ++i) {
    const uint32_t c = (uint32_t)(chain_offset + i);
    const float f = ((float)c + 0.5f) / (float)total;
    float *x = out + i * dim;
    for (int d = 0; d < dim; ++d) x[d] = fmaf(f, le[d] - ls[d], ls[d]);
    hnr_chain(A, b, F, dim, x, c, mixing_steps, seed, stream, round);
  }
}

/* ------------------------------------------------------------------------ */
/* EI-ZO stages                                                              */
/* ------------------------------------------------------------------------ */
/* cuda_set_builder_utils.h:198-203 -- float log, float ceil */
uint32_t orc_num_samples_unadaptive(float epsilon, float delta, float tau) {
  return (uint32_t)ceilf(2 * logf(1 / delta) / (epsilon * tau * tau));
}
/* cuda_edge_inflation_zero_order.cu:113-114 */
float orc_delta_k(float delta, int32_t it) {
  return (float)(6 * delta / (M_PI * M_PI * (it + 1) * (it + 1)));
}

/* cuda_set_builder_utils.cu:25-67 */
void orc_project_onto_segment(const float *p, const float *q1, const float *q2,
                              int32_t dim, float *proj, float *dist) {
  float a = 0.f, b = 0.f;
  for (int d = 0; d < dim; ++d) {
    a = a + (p[d] - q1[d]) * (q2[d] - q1[d]);
    b = b + (q2[d] - q1[d]) * (q2[d] - q1[d]);
  }
  float t = fminf(fmaxf(a / b, 0.0f), 1.0f);
  float s = 0.f;
  for (int d = 0; d < dim; ++d) {
    proj[d] = (1 - t) * q1[d] + t * q2[d];
    float e = proj[d] - p[d];
    s = s + e * e;
  }
  if (dist) *dist = sqrtf(s);
}

/* bisection loop, cuda_edge_inflation_zero_order.cu:215-259 with the
 * kernels of cuda_set_builder_utils.cu:239-319 */
void orc_bisect(const orc_plant *p, const float *particle, const float *proj,
                int32_t dim, int32_t steps, const float *voxels,
                int32_t num_voxels, float voxel_radius,
                const int32_t *robot_geoms, int32_t num_robot_geoms, int rule,
                float *scol) {
  float lo[ORC_MAX_DIM], hi[ORC_MAX_DIM], mid[ORC_MAX_DIM];
  for (int d = 0; d < dim; ++d) {
    lo[d] = proj[d];
    hi[d] = particle[d];
    scol[d] = particle[d];
  }
  for (int s = 0; s < steps; ++s) {
    for (int d = 0; d < dim; ++d) mid[d] = (float)(0.5 * (hi[d] + lo[d]));
    int free_ = check_one(p, mid, voxels, num_voxels, voxel_radius, robot_geoms,
                          num_robot_geoms, rule);
    if (!free_) {
      for (int d = 0; d < dim; ++d) {
        hi[d] = mid[d];
        scol[d] = mid[d];
      }
    } else {
      for (int d = 0; d < dim; ++d) lo[d] = mid[d];
    }
  }
}

/* round a double to the nearest float at or above it */
static float round_up_f(double v) {
  float f = (float)v;
  if ((double)f < v) f = nextafterf(f, INFINITY);
  return f;
}

/* greedy face placement: cuda_edge_inflation_zero_order.cu:268-306 with
 * maybePlaceFace / checkParticlesMadeRedundantByLastFace
 * (cuda_set_builder_utils.cu:359-424), placement and redundancy in FP64.
 * With a metric M (SPD, dim x dim row-major; NULL = identity, the
 * reference), particles are taken in ascending |e|_M = sqrt(e' M e),
 * e = scol - proj, and the face normal is M e normalised to unit length:
 * the hyperplane tangent at scol to the M-ellipsoid centred at proj. */
int32_t orc_place_faces_metric(float *A, float *b, int32_t *num_faces,
                               int32_t cap, int32_t dim, const float *scol,
                               const float *proj, int32_t u, const float *ls,
                               const float *le, float margin,
                               int32_t max_faces, const double *M) {
  double *dist = (double *)malloc(sizeof(double) * (u > 0 ? u : 1));
  for (int i = 0; i < u; ++i) {
    double e[ORC_MAX_DIM], s = 0;
    for (int d = 0; d < dim; ++d)
      e[d] = (double)scol[i * dim + d] - (double)proj[i * dim + d];
    if (M) {
      for (int r = 0; r < dim; ++r) {
        double me = 0;
        for (int c = 0; c < dim; ++c) me += M[r * dim + c] * e[c];
        s += e[r] * me;
      }
    } else {
      for (int d = 0; d < dim; ++d) s += e[d] * e[d];
    }
    dist[i] = sqrt(s);
  }
  int added = 0;
  for (; added < max_faces; ++added) {
    int best = -1;
    double bd = DBL_MAX;
    for (int i = 0; i < u; ++i)
      if (dist[i] < bd) { /* strict: first index on ties */
        bd = dist[i];
        best = i;
      }
    if (best < 0) break; /* every particle redundant */
    if (*num_faces >= cap) {
      free(dist);
      return -1;
    }
    const float *sc = scol + best * dim, *pr = proj + best * dim;
    double e[ORC_MAX_DIM], a[ORC_MAX_DIM], nrm = 0;
    for (int d = 0; d < dim; ++d) e[d] = (double)sc[d] - (double)pr[d];
    for (int r = 0; r < dim; ++r) {
      if (M) {
        a[r] = 0;
        for (int c = 0; c < dim; ++c) a[r] += M[r * dim + c] * e[c];
      } else {
        a[r] = e[r];
      }
      nrm += a[r] * a[r];
    }
    nrm = sqrt(nrm);
    float af[ORC_MAX_DIM];
    for (int d = 0; d < dim; ++d) {
      af[d] = (float)(a[d] / nrm);
      a[d] = af[d];
    }
    double bb = -(double)margin, v1 = 0, v2 = 0;
    for (int d = 0; d < dim; ++d) {
      bb += a[d] * sc[d];
      v1 += a[d] * ls[d];
      v2 += a[d] * le[d];
    }
    double relax = v2 - bb > v1 - bb ? v2 - bb : v1 - bb;
    if (relax > 0) bb += relax;
    int f = *num_faces;
    for (int d = 0; d < dim; ++d) A[f * dim + d] = af[d];
    b[f] = round_up_f(bb);
    *num_faces = f + 1;
    dist[best] = DBL_MAX;
    for (int i = 0; i < u; ++i) {
      if (dist[i] == DBL_MAX) continue;
      double v = -bb;
      for (int d = 0; d < dim; ++d) v += a[d] * scol[i * dim + d];
      if (v > 0) dist[i] = DBL_MAX;
    }
  }
  free(dist);
  return added;
}

int32_t orc_place_faces(float *A, float *b, int32_t *num_faces, int32_t cap,
                        int32_t dim, const float *scol, const float *proj,
                        int32_t u, const float *ls, const float *le,
                        float margin, int32_t max_faces) {
  return orc_place_faces_metric(A, b, num_faces, cap, dim, scol, proj, u, ls, le,
                                margin, max_faces, NULL);
}

/* region editing (cuda_edit_regions.cu:236-293) for one region: the
 * reference's walk over its collisions sorted by distance, each tested for
 * strict containment in the region as edited so far (A.x - b < 0 on every
 * face, evaluated in FP64 on the stored FP32 faces), with the face
 * a = (scol - proj)/|scol - proj| (left zero for a zero difference),
 * b = a.scol - margin, raised by max(a.ls - b, a.le - b) + 2e-6 when that is
 * positive, stored rounded up. Ties in distance keep the lower index. */
typedef struct {
  double d;
  int32_t i;
} orc_key;

static int key_cmp(const void *x, const void *y) {
  const orc_key *a = (const orc_key *)x, *b = (const orc_key *)y;
  if (a->d < b->d) return -1;
  if (a->d > b->d) return 1;
  return (a->i > b->i) - (a->i < b->i);
}

int32_t orc_edit_region(float *A, float *b, int32_t *num_faces, int32_t cap,
                        int32_t dim, const float *scol, const float *proj,
                        int32_t u, const float *ls, const float *le,
                        float margin) {
  orc_key *keys = (orc_key *)malloc(sizeof(orc_key) * (u > 0 ? u : 1));
  for (int i = 0; i < u; ++i) {
    double s = 0;
    for (int d = 0; d < dim; ++d) {
      double e = (double)scol[i * dim + d] - (double)proj[i * dim + d];
      s += e * e;
    }
    keys[i].d = sqrt(s);
    keys[i].i = i;
  }
  qsort(keys, (size_t)u, sizeof(orc_key), key_cmp);
  int32_t added = 0;
  for (int k = 0; k < u; ++k) {
    const float *sc = scol + keys[k].i * dim, *pr = proj + keys[k].i * dim;
    int inside = 1;
    for (int f = 0; f < *num_faces && inside; ++f) {
      double v = -(double)b[f];
      for (int d = 0; d < dim; ++d) v += (double)A[f * dim + d] * (double)sc[d];
      inside = v < 0.0;
    }
    if (!inside) continue;
    if (*num_faces >= cap) {
      free(keys);
      return -1;
    }
    double a[ORC_MAX_DIM], nrm = 0;
    for (int d = 0; d < dim; ++d) {
      a[d] = (double)sc[d] - (double)pr[d];
      nrm += a[d] * a[d];
    }
    nrm = sqrt(nrm);
    const int f = *num_faces;
    double bb = -(double)margin, v1 = 0, v2 = 0;
    for (int d = 0; d < dim; ++d) {
      float af = nrm > 0 ? (float)(a[d] / nrm) : 0.f;
      A[f * dim + d] = af;
      bb += (double)af * sc[d];
      v1 += (double)af * ls[d];
      v2 += (double)af * le[d];
    }
    double relax = v1 - bb > v2 - bb ? v1 - bb : v2 - bb;
    if (relax > 0) bb += relax + 2e-6;
    b[f] = round_up_f(bb);
    *num_faces = f + 1;
    ++added;
  }
  free(keys);
  return added;
}

/* distance field (geometry/cuda_distance_field.h:76-131): occupancy by
 * voxel round((p - origin) / voxel) per axis; the squared distance in voxels
 * from every voxel to the nearest occupied one, by definition (a scan over
 * all occupied voxels), 0x3fffffff when none is occupied. */
void orc_voxelize(const float *points, int64_t m, const float *origin,
                  float voxel, int32_t G, uint8_t *occ) {
  for (int64_t i = 0; i < m; ++i) {
    int c[3];
    int inside = 1;
    for (int k = 0; k < 3; ++k) {
      c[k] = (int)roundf((points[3 * i + k] - origin[k]) / voxel);
      inside &= c[k] >= 0 && c[k] < G;
    }
    if (inside) occ[((int64_t)c[2] * G + c[1]) * G + c[0]] = 1;
  }
}

void orc_edt_brute(const uint8_t *occ, int32_t G, int32_t *d2, int threads) {
  const int64_t n = (int64_t)G * G * G;
  int64_t ns = 0;
  for (int64_t i = 0; i < n; ++i) ns += occ[i] != 0;
  int32_t *sites = (int32_t *)malloc(sizeof(int32_t) * 3 * (ns > 0 ? ns : 1));
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i)
    if (occ[i]) {
      sites[3 * k] = (int32_t)(i % G);
      sites[3 * k + 1] = (int32_t)((i / G) % G);
      sites[3 * k + 2] = (int32_t)(i / ((int64_t)G * G));
      ++k;
    }
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % G), y = (int)((i / G) % G), z = (int)(i / ((int64_t)G * G));
    int32_t best = 0x3fffffff;
    for (int64_t s = 0; s < ns; ++s) {
      const int dx = x - sites[3 * s], dy = y - sites[3 * s + 1], dz = z - sites[3 * s + 2];
      const int32_t v = dx * dx + dy * dy + dz * dz;
      if (v < best) best = v;
    }
    d2[i] = best;
  }
  free(sites);
}

/* one EI-ZO region (cuda_edge_inflation_zero_order.cu:109-306), with the
 * build's deterministic choices: colliding samples are taken in sample-index
 * order, the sampler draws round-dependent Philox streams, failed
 * hit-and-run steps keep the chain in place. */
int32_t orc_eizo(const orc_plant *p, const int32_t *robot_geoms,
                 int32_t num_robot_geoms, const float *voxels,
                 int32_t num_voxels, float voxel_radius, float *A, float *b,
                 int32_t *num_faces, int32_t cap, const float *ls,
                 const float *le, const orc_eizo_options *opt, uint32_t stream,
                 int rule, int spread_starts, orc_eizo_info *info) {
  const int dim = p->num_config_vars;
  info->iterations = 0;
  info->converged = 0;
  info->samples_drawn = 0;
  info->faces_added = 0;
  for (uint32_t it = 0; it < opt->max_iterations; ++it) {
    float dk = orc_delta_k(opt->delta, (int32_t)it);
    uint32_t M = orc_num_samples_unadaptive(opt->epsilon, dk, opt->tau);
    uint32_t N = M > opt->num_particles ? M : opt->num_particles;
    float *samples = (float *)malloc(sizeof(float) * (size_t)N * dim);
    uint8_t *flags = (uint8_t *)malloc(N);
    if (spread_starts)
      orc_hit_and_run_segment(A, b, *num_faces, dim, ls, le, N, 0, N,
                              (int32_t)opt->mixing_steps, opt->seed, stream, it,
                              samples);
    else
      orc_hit_and_run(A, b, *num_faces, dim, ls, N, 0, (int32_t)opt->mixing_steps,
                      opt->seed, stream, it, samples);
    info->samples_drawn += N;
    orc_check_voxels_batch(p, samples, N, voxels, num_voxels, voxel_radius,
                           robot_geoms, num_robot_geoms, rule, flags, 0);
    uint32_t ncol_first = 0;
    for (uint32_t i = 0; i < M; ++i) ncol_first += flags[i] ? 0 : 1;
    info->iterations = (int32_t)it + 1;
    float thresh = (1 - opt->tau) * opt->epsilon * M;
    if (ncol_first <= thresh) {
      info->converged = 1;
      free(samples);
      free(flags);
      break;
    }
    uint32_t u = 0;
    float *scol = (float *)malloc(sizeof(float) * opt->num_particles * dim);
    float *proj = (float *)malloc(sizeof(float) * opt->num_particles * dim);
    for (uint32_t i = 0; i < N && u < opt->num_particles; ++i) {
      if (flags[i]) continue;
      orc_project_onto_segment(samples + (size_t)i * dim, ls, le, dim,
                               proj + u * dim, NULL);
      orc_bisect(p, samples + (size_t)i * dim, proj + u * dim, dim,
                 (int32_t)opt->bisection_steps, voxels, num_voxels,
                 voxel_radius, robot_geoms, num_robot_geoms, rule,
                 scol + u * dim);
      ++u;
    }
    int32_t added = orc_place_faces(A, b, num_faces, cap, dim, scol, proj,
                                    (int32_t)u, ls, le, opt->configuration_margin,
                                    (int32_t)opt->max_hyperplanes_per_iteration);
    free(scol);
    free(proj);
    free(samples);
    free(flags);
    if (added < 0) return -1;
    info->faces_added += added;
  }
  return 0;
}

/* Regions of orc_eizo side by side: region r inflates ls[r]-le[r] (rows of
 * [R][dim]) from the domain (A_dom, b_dom, F faces) into A_out[r] ([cap][dim])
 * and b_out[r], with stream r; one region per OpenMP thread (dynamic
 * schedule), each region's checks on its own thread. status[r] is orc_eizo's
 * return code. No reference counterpart: the reference has no CPU EI-ZO;
 * this is the CPU restatement run on every host core. */
void orc_eizo_batch(const orc_plant *p, const int32_t *robot_geoms,
                    int32_t num_robot_geoms, const float *A_dom, const float *b_dom,
                    int32_t F, int32_t cap, int32_t R, const float *ls, const float *le,
                    const orc_eizo_options *opt, int rule, int spread_starts,
                    float *A_out, float *b_out, int32_t *nf_out, int32_t *status,
                    orc_eizo_info *infos, int nthreads) {
  const int dim = p->num_config_vars;
  const int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int32_t r = 0; r < R; ++r) {
    float *A = A_out + (size_t)r * cap * dim;
    float *b = b_out + (size_t)r * cap;
    memcpy(A, A_dom, sizeof(float) * (size_t)F * dim);
    memcpy(b, b_dom, sizeof(float) * (size_t)F);
    nf_out[r] = F;
    status[r] = orc_eizo(p, robot_geoms, num_robot_geoms, NULL, 0, 0.f, A, b, &nf_out[r], cap,
                         ls + (size_t)r * dim, le + (size_t)r * dim, opt, (uint32_t)r, rule,
                         spread_starts, &infos[r]);
  }
}
