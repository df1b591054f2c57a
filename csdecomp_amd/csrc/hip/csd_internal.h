// csd_internal.h -- device image of a compiled plant and shared helpers of
// the gfx950 kernels. Host and device code both include this file.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <string>
#include <vector>

#include "../../../include/csdecomp_amd.h"

namespace csd {

// ---------------------------------------------------------------------------
// capacities of the compiled image (runtime values, not recompile-time)
// ---------------------------------------------------------------------------
constexpr int kMaxBodies = 64;     // non-fixed joints per plant
constexpr int kMaxDof = 32;        // configuration variables
constexpr int kFastMaxSlots = 128; // sphere slots of the fast kernel
constexpr int kGenericBlock = 64;  // one wave per block in the generic kernel

// A body is the rigid group of links moved by one non-fixed joint.
// X_body = X_parent_body * C * J(q), 3x4 row-major [R | t].
// For a revolute joint with unit axis a, J(q) = c I + (1-c) a a^T + s [a]x,
// so C_R J_R = C_R + s CK + (1-c) CAm with CK = C_R [a]x, CAm = C_R a a^T - C_R
// (both constant). For a prismatic joint CK[0..2] = C_R a.
struct DevBody {
  int32_t parent;  // parent body, -1 = world-fixed (C then includes its pose)
  int32_t type;    // CSD_JOINT_REVOLUTE / CSD_JOINT_PRISMATIC
  int32_t q_idx;   // configuration index
  int32_t pad;
  float C[12];
  float CK[9];
  float CAm[9];
  float pad2[2];
};

// Geometry of the generic kernel: pose G relative to its body (body >= 0)
// or to the world (body == -1), 3x4 row-major.
struct DevGeom {
  int32_t type;
  int32_t body;
  int32_t pad[2];
  float dims[4];  // box: half extents; sphere: r; capsule: r, half length
  float G[12];
};

// Static obstacle record of the fast kernel (16 floats).
//  box:     [0..8] R^T row-major, [9..11] -R^T p, [12..14] half extents
//  sphere:  [0..2] centre, [3] radius
//  capsule: [0..2] p0, [3..5] d = p1 - p0, [6] 1/|d|^2, [7] radius
struct DevObstacle {
  float v[16];
};

// Link record for standalone FK output: X_W_L = X_body * O (body >= 0) or O
// (world-fixed link, body == -1).
struct DevLink {
  int32_t body;
  int32_t pad[3];
  float O[12];
};

struct PlantImage {
  int32_t num_links;
  int32_t dof;
  int32_t num_bodies;
  int32_t num_geoms;
  int32_t num_pairs;
  int32_t always_colliding;  // a world-fixed pair collides at every config
  // fast path (every moving geometry is a sphere)
  int32_t fast;          // 1 if the fast kernel applies
  int32_t kmax;          // sphere slots per body
  int32_t nslots;        // num_bodies * kmax
  int32_t num_obstacles; // static obstacles paired with moving spheres
  int32_t num_self_pairs; // moving sphere x moving sphere pairs
  int32_t pad;
  // device arrays
  const DevBody *bodies;
  const DevGeom *geoms;
  const int32_t *pairs;        // [P][2] compiled geometry indices
  const int32_t *pair_types;   // [P] (typeA*4+typeB) after canonical order
  const DevLink *links;
  const float4 *slot_sphere;   // [nslots] local centre + radius (r<0: empty)
  const DevObstacle *obstacles;
  const int32_t *obstacle_type;   // [num_obstacles]
  // squared contact distance of (obstacle, sphere slot): box r^2, sphere
  // and capsule (r + r_o)^2, all rounded in FP32 as the device would;
  // -1 where the slot is not paired with the obstacle (never a hit)
  const float *obstacle_r2;       // [num_obstacles][nslots]
  // self collision of the fast path as body-pair blocks: bit bb of
  // self_active[ba] says bodies ba < bb have sphere pairs; self_r2 holds
  // (rA+rB)^2 per (ba, bb, ka, kb), 0 for absent pairs (never a hit)
  const uint64_t *self_active;    // [nbt]
  const float *self_r2;           // [nbt][nbt][kmax][kmax]
  const int32_t *geom_slot;       // [G] slot of moving sphere geometry, -1 else
};

}  // namespace csd

struct csd_plant {
  csd::PlantImage img;      // host copy of the image header (device pointers)
  void *arena = nullptr;    // one device allocation holding all arrays
  int device = 0;
  // host-side copies used for argument checks
  std::vector<int32_t> geom_type, geom_link, geom_body, geom_slot;
  int32_t num_configuration_variables = 0;
  int32_t num_links = 0;
};

namespace csd {
// thread-local error message for csd_last_error
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define CSD_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return ::csd::hip_fail(_e, #call); \
  } while (0)
}  // namespace csd
