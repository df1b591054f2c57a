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
  int32_t q_identity;     // 1 if body b is driven by configuration variable b
  // device arrays
  const DevBody *bodies;
  const DevGeom *geoms;
  const int32_t *pairs;        // [P][2] compiled geometry indices
  const int32_t *pair_types;   // [P] (typeA*4+typeB) after canonical order
  const DevLink *links;
  const float4 *slot_sphere;   // [nslots] local centre + radius (r<0: empty)
  const DevObstacle *obstacles;
  const int32_t *obstacle_type;   // [num_obstacles]
  // bit s of obstacle_mask[2o + s/64]: sphere slot s is paired with
  // obstacle o (all 64-byte records and masks together fit the scalar cache)
  const uint64_t *obstacle_mask;  // [num_obstacles][2]
  // plants of at most 64 sphere slots also get contact-distance tables:
  // obstacle_r2[o][s] = r^2 (box) or (r + r_o)^2 (sphere, capsule) in FP32 as
  // the device computes it, -1 if slot s is not paired with obstacle o
  const float *obstacle_r2;       // [num_obstacles][nslots] (nslots <= 64)
  // self collision of the fast path as body-pair blocks: bit bb of
  // self_active[ba] says bodies ba < bb have sphere pairs; bit ka*kmax+kb
  // of self_bits[ba][bb] says slots (ba, ka) and (bb, kb) form a pair
  const uint64_t *self_active;    // [nbt]
  const uint64_t *self_bits;      // [nbt][nbt]
  const float *self_r2;           // [nbt][nbt][kmax][kmax] (nslots <= 64), 0 if absent
  const int32_t *geom_slot;       // [G] slot of moving sphere geometry, -1 else
  // bounds for culling obstacle tests. body_bound[b]: radius about the
  // centre of body b's first sphere slot that encloses all of its spheres;
  // obstacle_bound[o]: world centre + radius of a sphere enclosing obstacle
  // o. Both radii are inflated by kBoundInflate (relative) + kBoundPad
  // (metres), far beyond FP32 rounding of the centres and of the exact
  // tests, so a culled test can never have reported a hit.
  const float *body_bound;        // [nbt]
  const float4 *obstacle_bound;   // [num_obstacles]
};

constexpr double kBoundInflate = 1e-3;
constexpr double kBoundPad = 5e-4;

// Screen of the fast checker: a subset of the plant's pair tests that finds
// most collisions of a workload early. Configurations the screen leaves
// undecided get the full check. Pair tests are the full check's own (same
// contact distances, same arithmetic), so the screened result is the
// unscreened result. The screen is a program in device memory that each
// workgroup copies to LDS (16-byte words, layout for NB bodies of KM slots,
// NS = NB * KM):
//   uint64 self_active[NB]       bit bb: body pair (ba, bb) has screened pairs
//   uint64 self_rows[NB][NB]     bit ka: sphere ka of body ba has screened
//                                pairs with body bb
//   float  self_r2[NB][NB][KM][KM]  contact distances of screened pairs
//                                (the plant's), -1 for the others
//   then num_obst records of {float ov[16]; float r2[NS]; uint64 bodies;
//   int32 type; int32 pad}: the obstacle, its contact distances with the
//   screened slots (-1 for the others) and the bodies owning them
constexpr int kScreenMaxObst = 32;
struct ScreenImage {
  int32_t num_obst;    // obstacles with screened slots
  int32_t num_tests;   // screened pair tests in total
  int32_t prog_words;  // program size in 16-byte words
  int32_t pad;
  const uint4 *prog;
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
  // waves sharing one configuration in the checker, tuned per batch-size
  // class (index: bit length of n); 0 = not tuned
  int8_t check_split[64] = {0};
  // screened checking, chosen by the autotuner per batch-size class:
  // check_range[c] > 0 is the configuration range of one survivor workgroup
  // (0 = unscreened); the screen itself lives in screen_arena
  int32_t check_range[64] = {0};
  csd::ScreenImage screen{};
  void *screen_arena = nullptr;
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
