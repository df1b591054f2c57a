// cpu_checker.cpp -- the reference's CPU collision API evaluated on the host:
// link transforms (minimal_kinematic_tree.cpp:10-58), pair tests
// (collision_checker.cpp:91-203), geometry-voxel tests (:205-253) and the
// per-configuration drivers (:9-89), with the reference's CPU comparison
// rules (strict ">" for "free", the bounding-sphere shortcut before the box
// SAT). One configuration at a time on the calling thread, as in the
// reference: these names serve callers that loop over single configurations
// on the CPU. The batched data-parallel path is the GPU's
// (checkCollisionFreeCuda and the C-ABI); nothing here is called by it.
//
// Capsules have no CPU test in the reference (it throws); here a capsule is
// the segment c +- hl z of its frame swept by radius r (dimensions (r, hl)),
// tested in FP32 with the same strict rule. Cylinders throw as in the
// reference.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "csdecomp.h"

namespace csdecomp {
namespace {

// C = A * B for column-major 4x4 (element (i, k) at [4k + i]); every
// coefficient accumulates k = 0..3 in order, unfused
void mul4(const float *A, const float *B, float *C) {
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) {
      float s = A[i] * B[4 * j];
      s = s + A[4 + i] * B[4 * j + 1];
      s = s + A[8 + i] * B[4 * j + 2];
      s = s + A[12 + i] * B[4 * j + 3];
      C[4 * j + i] = s;
    }
}

float dot3(const float *a, const float *b) {
  float s = a[0] * b[0];
  s = s + a[1] * b[1];
  return s + a[2] * b[2];
}

// column k of the rotation part of a column-major 4x4
const float *col(const Matrix4f &X, int k) { return X.data() + 4 * k; }

// joint motion X_J_L(q): Rodrigues rotation about a unit axis (revolute) or a
// translation along the axis (prismatic)
Matrix4f joint_motion(const MinimalJoint &j, float q) {
  Matrix4f X = identity4();
  if (j.type == REVOLUTE) {
    const float s = std::sin(q), c = std::cos(q), oc = 1.f - c;
    const float a0 = j.axis[0], a1 = j.axis[1], a2 = j.axis[2];
    const float w0 = oc * a0, w1 = oc * a1, w2 = oc * a2;  // (1 - cos) axis
    const float s0 = s * a0, s1 = s * a1, s2 = s * a2;     // sin axis
    X[0] = w0 * a0 + c;
    X[5] = w1 * a1 + c;
    X[10] = w2 * a2 + c;
    X[4] = w0 * a1 - s2;  // (0, 1)
    X[1] = w0 * a1 + s2;  // (1, 0)
    X[8] = w0 * a2 + s1;  // (0, 2)
    X[2] = w0 * a2 - s1;  // (2, 0)
    X[9] = w1 * a2 - s0;  // (1, 2)
    X[6] = w1 * a2 + s0;  // (2, 1)
  } else if (j.type == PRISMATIC) {
    X[12] = 0.f + j.axis[0] * q;
    X[13] = 0.f + j.axis[1] * q;
    X[14] = 0.f + j.axis[2] * q;
  }
  return X;
}

// world pose of a geometry: X_W_L * X_L_B
Matrix4f geometry_pose(const CollisionGeometry &g, const Matrix4f &X_W_L) {
  Matrix4f X;
  mul4(X_W_L.data(), g.X_L_B.data(), X.data());
  return X;
}

void translation(const Matrix4f &X, float *p) {
  p[0] = X[12];
  p[1] = X[13];
  p[2] = X[14];
}

float norm3(const float *d) { return std::sqrt(dot3(d, d)); }

// squared distance from point c to a box posed by X with full edge lengths
// `dims`: the point in box coordinates R^T (c - p), clamped to the box
float point_box_d2(const float *c, const Matrix4f &X, const Vector3f &dims) {
  float w[3] = {c[0] - X[12], c[1] - X[13], c[2] - X[14]};
  float d2 = 0.f;
  for (int i = 0; i < 3; ++i) {
    const float l = dot3(col(X, i), w);
    const float h = dims[i] * 0.5f;
    const float e = std::max(-h, std::min(l, h)) - l;
    d2 = i == 0 ? e * e : d2 + e * e;
  }
  return d2;
}

bool box_box_free(const Matrix4f &XA, const Vector3f &dA, const Matrix4f &XB,
                  const Vector3f &dB) {
  float pA[3], pB[3], d[3];
  translation(XA, pA);
  translation(XB, pB);
  for (int k = 0; k < 3; ++k) d[k] = pA[k] - pB[k];
  const float maxlen = std::max(*std::max_element(dA.begin(), dA.end()),
                                *std::max_element(dB.begin(), dB.end()));
  // bounding-sphere shortcut, compared in double
  if ((double)norm3(d) > 2 * std::sqrt(3.0) * maxlen) return true;
  // 15 candidate separating axes: A's three, B's three, their nine cross
  // products (A_i x B_j, i major)
  float axes[15][3];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      axes[i][k] = col(XA, i)[k];
      axes[3 + i][k] = col(XB, i)[k];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const float *u = axes[i], *v = axes[3 + j];
      float *c = axes[6 + 3 * i + j];
      c[0] = u[1] * v[2] - u[2] * v[1];
      c[1] = u[2] * v[0] - u[0] * v[2];
      c[2] = u[0] * v[1] - u[1] * v[0];
    }
  for (const auto &n : axes) {
    const float proj = std::fabs(dot3(d, n));
    float reach = std::fabs((dA[0] / 2) * dot3(axes[0], n));
    reach = reach + std::fabs((dA[1] / 2) * dot3(axes[1], n));
    reach = reach + std::fabs((dA[2] / 2) * dot3(axes[2], n));
    reach = reach + std::fabs((dB[0] / 2) * dot3(axes[3], n));
    reach = reach + std::fabs((dB[1] / 2) * dot3(axes[4], n));
    reach = reach + std::fabs((dB[2] / 2) * dot3(axes[5], n));
    if ((double)proj > (double)reach) return true;  // separating plane found
  }
  return false;
}

// capsule axis segment in world: p0 = c - hl z, d = 2 hl z
void capsule_segment(const Matrix4f &X, float hl, float *p0, float *d) {
  for (int k = 0; k < 3; ++k) {
    p0[k] = X[12 + k] - hl * X[8 + k];
    d[k] = 2.f * hl * X[8 + k];
  }
}

float point_segment_d2(const float *c, const float *p0, const float *d) {
  const float w[3] = {c[0] - p0[0], c[1] - p0[1], c[2] - p0[2]};
  const float dd = dot3(d, d);
  float t = dd > 0.f ? dot3(w, d) / dd : 0.f;
  t = std::min(std::max(t, 0.f), 1.f);
  float e[3];
  for (int k = 0; k < 3; ++k) e[k] = p0[k] + t * d[k] - c[k];
  return dot3(e, e);
}

// closest points of two segments p1 + s d1, p2 + t d2, s, t in [0, 1]
// (clamped minimisation of the quadratic; Ericson, Real-Time Collision
// Detection, 5.1.9)
float segment_segment_d2(const float *p1, const float *d1, const float *p2, const float *d2) {
  const float r[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const float a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
  float s = 0.f, t = 0.f;
  if (a > 0.f || e > 0.f) {
    if (a <= 0.f) {
      t = std::min(std::max(f / e, 0.f), 1.f);
    } else {
      const float c = dot3(d1, r);
      if (e <= 0.f) {
        s = std::min(std::max(-c / a, 0.f), 1.f);
      } else {
        const float b = dot3(d1, d2), den = a * e - b * b;
        s = den > 0.f ? std::min(std::max((b * f - c * e) / den, 0.f), 1.f) : 0.f;
        t = (b * s + f) / e;
        if (t < 0.f) {
          t = 0.f;
          s = std::min(std::max(-c / a, 0.f), 1.f);
        } else if (t > 1.f) {
          t = 1.f;
          s = std::min(std::max((b - c) / a, 0.f), 1.f);
        }
      }
    }
  }
  float x[3];
  for (int k = 0; k < 3; ++k) x[k] = (p1[k] + s * d1[k]) - (p2[k] + t * d2[k]);
  return dot3(x, x);
}

// squared distance from a segment to a box: the squared distance of the
// segment's point to the box is convex in the segment parameter, so a
// bisection on the sign of its derivative finds the minimum
float segment_box_d2(const float *p0w, const float *dw, const Matrix4f &X, const Vector3f &dims) {
  float a[3], d[3], h[3];
  const float w[3] = {p0w[0] - X[12], p0w[1] - X[13], p0w[2] - X[14]};
  for (int i = 0; i < 3; ++i) {
    a[i] = dot3(col(X, i), w);
    d[i] = dot3(col(X, i), dw);
    h[i] = dims[i] * 0.5f;
  }
  auto eval = [&](float t, float *slope) {
    float f = 0.f, g = 0.f;
    for (int i = 0; i < 3; ++i) {
      const float x = a[i] + t * d[i];
      const float e = std::max(std::fabs(x) - h[i], 0.f);
      f += e * e;
      g += (x >= 0.f ? e : -e) * d[i];
    }
    if (slope) *slope = g;
    return f;
  };
  float g0, g1;
  const float f0 = eval(0.f, &g0);
  if (g0 >= 0.f) return f0;
  const float f1 = eval(1.f, &g1);
  if (g1 <= 0.f) return f1;
  float lo = 0.f, hi = 1.f;
  for (int it = 0; it < 40; ++it) {
    const float mid = 0.5f * (lo + hi);
    float g;
    eval(mid, &g);
    (g > 0.f ? hi : lo) = mid;
  }
  return eval(0.5f * (lo + hi), nullptr);
}

void reject_cylinder(ShapePrimitive t) {
  if (t == CYLINDER) throw std::runtime_error("Cylinders are not supported.");
}

// capsule against a sphere, box or capsule (posed); true if free
bool capsule_pair_free(const CollisionGeometry &cap, const Matrix4f &XC,
                       const CollisionGeometry &o, const Matrix4f &XO) {
  float p0[3], d[3];
  capsule_segment(XC, cap.dimensions[1], p0, d);
  if (o.type == SPHERE) {
    float c[3];
    translation(XO, c);
    const float rs = cap.dimensions[0] + o.dimensions[0];
    return point_segment_d2(c, p0, d) > rs * rs;
  }
  if (o.type == BOX)
    return segment_box_d2(p0, d, XO, o.dimensions) > cap.dimensions[0] * cap.dimensions[0];
  float q0[3], e[3];
  capsule_segment(XO, o.dimensions[1], q0, e);
  const float rs = cap.dimensions[0] + o.dimensions[0];
  return segment_segment_d2(p0, d, q0, e) > rs * rs;
}

// transforms of all links for configuration q, [L] column-major 4x4
void link_transforms(const std::vector<float> &q, const MinimalKinematicTree &t,
                     std::vector<Matrix4f> &X) {
  if ((int)q.size() != t.num_configuration_variables)
    throw std::runtime_error("configuration has the wrong number of variables");
  X.assign(t.num_links, identity4());  // the root link is the world frame
  for (int s = 0; s < t.num_joints; ++s) {
    const MinimalJoint &j = t.joints[t.joint_traversal_sequence[s]];
    const float v = j.type == FIXED ? 0.f : q[t.joint_idx_to_config_idx[j.index]];
    const Matrix4f XJL = joint_motion(j, v);
    Matrix4f T;
    mul4(X[j.parent_link].data(), j.X_PL_J.data(), T.data());
    mul4(T.data(), XJL.data(), X[j.child_link].data());
  }
}

const CollisionGeometry &geometry(const MinimalPlant &p, int g) {
  if (g < 0 || g >= (int)p.scene_collision_geometries.size())
    throw std::runtime_error("collision pair names a geometry outside the scene");
  return p.scene_collision_geometries[g];
}

bool geometry_voxel_free(const CollisionGeometry &g, const Matrix4f &XG, const float *vox,
                         float voxel_radius) {
  reject_cylinder(g.type);
  if (g.type == SPHERE) {
    const float d[3] = {XG[12] - vox[0], XG[13] - vox[1], XG[14] - vox[2]};
    return norm3(d) > g.dimensions[0] + voxel_radius;
  }
  if (g.type == BOX) return point_box_d2(vox, XG, g.dimensions) > voxel_radius * voxel_radius;
  float p0[3], d[3];
  capsule_segment(XG, g.dimensions[1], p0, d);
  const float rs = g.dimensions[0] + voxel_radius;
  return point_segment_d2(vox, p0, d) > rs * rs;
}

}  // namespace

std::vector<Matrix4f> computeLinkFrameToWorldTransformsMinimal(const std::vector<float> &q,
                                                                const MinimalKinematicTree &tree) {
  std::vector<Matrix4f> X;
  link_transforms(q, tree, X);
  return X;
}

bool pairCollisionFree(const CollisionGeometry &a, const Matrix4f &X_W_LA,
                       const CollisionGeometry &b, const Matrix4f &X_W_LB) {
  reject_cylinder(a.type);
  reject_cylinder(b.type);
  const Matrix4f XA = geometry_pose(a, X_W_LA), XB = geometry_pose(b, X_W_LB);
  if (a.type == CAPSULE) return capsule_pair_free(a, XA, b, XB);
  if (b.type == CAPSULE) return capsule_pair_free(b, XB, a, XA);
  if (a.type == SPHERE && b.type == SPHERE) {
    const float d[3] = {XA[12] - XB[12], XA[13] - XB[13], XA[14] - XB[14]};
    return norm3(d) > a.dimensions[0] + b.dimensions[0];
  }
  if (a.type == BOX && b.type == BOX) return box_box_free(XA, a.dimensions, XB, b.dimensions);
  const bool a_sphere = a.type == SPHERE;
  const CollisionGeometry &sphere = a_sphere ? a : b, &box = a_sphere ? b : a;
  const Matrix4f &XS = a_sphere ? XA : XB, &XBox = a_sphere ? XB : XA;
  float c[3];
  translation(XS, c);
  return point_box_d2(c, XBox, box.dimensions) > sphere.dimensions[0] * sphere.dimensions[0];
}

bool geomToVoxelCollisionFree(const CollisionGeometry &geom, const Matrix4f &X_W_L,
                              const Vector3f &p_W_voxel, float voxel_radius) {
  return geometry_voxel_free(geom, geometry_pose(geom, X_W_L), p_W_voxel.data(), voxel_radius);
}

bool checkCollisionFree(const std::vector<float> &q, const MinimalPlant &plant) {
  std::vector<Matrix4f> X;
  link_transforms(q, plant.kin_tree, X);
  bool free = true;
  // every pair is tested and every colliding one reported (no early exit)
  for (int k = 0; k < plant.num_collision_pairs; ++k) {
    const int ia = plant.collision_pairs_flat[2 * k], ib = plant.collision_pairs_flat[2 * k + 1];
    const CollisionGeometry &a = geometry(plant, ia), &b = geometry(plant, ib);
    if (!pairCollisionFree(a, X[a.link_index], b, X[b.link_index])) {
      std::printf("geoms colliding: (%d, %d) , links: (%d,%d) \n", ia, ib, a.link_index,
                  b.link_index);
      free = false;
    }
  }
  std::fflush(stdout);
  return free;
}

bool checkCollisionFreeVoxels(const std::vector<float> &q, const MatrixXf &voxels,
                              float voxel_radius, const MinimalPlant &plant,
                              const std::vector<GeometryIndex> &robot_geoms) {
  if (voxels.cols > 0 && voxels.rows != 3)
    throw std::runtime_error("voxels must be a 3 x V matrix");
  std::vector<Matrix4f> X;
  link_transforms(q, plant.kin_tree, X);
  for (int k = 0; k < plant.num_collision_pairs; ++k) {
    const CollisionGeometry &a = geometry(plant, plant.collision_pairs_flat[2 * k]);
    const CollisionGeometry &b = geometry(plant, plant.collision_pairs_flat[2 * k + 1]);
    if (!pairCollisionFree(a, X[a.link_index], b, X[b.link_index])) return false;
  }
  for (GeometryIndex g : robot_geoms) {
    const CollisionGeometry &rg = geometry(plant, g);
    const Matrix4f XG = geometry_pose(rg, X[rg.link_index]);
    for (int v = 0; v < voxels.cols; ++v)
      if (!geometry_voxel_free(rg, XG, &voxels.data[(size_t)3 * v], voxel_radius)) return false;
  }
  return true;
}

bool checkEdgeCollisionFree(const std::vector<float> &q1, const std::vector<float> &q2,
                            const MinimalPlant &plant, float step_size) {
  const int dof = plant.kin_tree.num_configuration_variables;
  if ((int)q1.size() != dof || (int)q2.size() != dof)
    throw std::runtime_error("edge endpoints have the wrong size");
  float len2 = 0.f;
  for (int d = 0; d < dof; ++d) {
    const float e = q2[d] - q1[d];
    len2 = d == 0 ? e * e : len2 + e * e;
  }
  // ceil(|q2 - q1| / step) points at t = k / n, k = 0..n-1: q1 is checked,
  // q2 is not
  const int n = (int)std::ceil(std::sqrt(len2) / step_size);
  std::vector<float> q(dof), X;
  for (int k = 0; k < n; ++k) {
    const float t = (float)(k / (1.0 * n));
    for (int d = 0; d < dof; ++d) q[d] = t * q2[d] + (1 - t) * q1[d];
    std::vector<Matrix4f> T;
    link_transforms(q, plant.kin_tree, T);
    for (int p = 0; p < plant.num_collision_pairs; ++p) {
      const CollisionGeometry &a = geometry(plant, plant.collision_pairs_flat[2 * p]);
      const CollisionGeometry &b = geometry(plant, plant.collision_pairs_flat[2 * p + 1]);
      if (!pairCollisionFree(a, T[a.link_index], b, T[b.link_index])) return false;
    }
  }
  return true;
}

}  // namespace csdecomp
