// csd_collide.h -- "is this configuration in collision?" as device functions,
// shared by the batched checker, the EI-ZO sampler/bisection kernels and the
// edge validator. One lane evaluates one configuration.
#pragma once
#include "csd_device.h"

namespace csd {

constexpr int kMaxRobotGeoms = 128;

// robot geometries checked against voxels, passed by value (kernarg segment)
struct RobotGeomList {
  int32_t n;
  int32_t pad[3];
  int32_t ids[kMaxRobotGeoms];
};

struct VoxelArgs {
  const float *vox;  // [V][3] device
  int32_t num;
  float radius;
  uint64_t mask0, mask1;  // fast path: robot sphere slots
};

// Keeps a wave-uniform value in a vector register: the slot radii are read
// by every pair test, and held in scalar registers they would spill.
__device__ __forceinline__ float in_vgpr(float x) {
  float v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
  return v;
}

// Wave-level early exit. ANY = false: every lane of the wave has found a
// collision (independent configurations). ANY = true: some lane has (the
// wave evaluates points of one edge and only "any collision" matters; the
// per-lane results are then incomplete by design).
template <bool ANY>
__device__ __forceinline__ bool wave_done(bool col) {
  if constexpr (ANY) return __any(col);
  else return __all(col);
}

// Bounding-sphere culling: true if, in every lane still undecided (col
// false), the centres a and b lie farther apart than rsum, the sum of two
// bounding radii about them. A body's bound is taken about the centre of its
// first sphere slot (PlantImage::body_bound), an obstacle's about its own
// centre (PlantImage::obstacle_bound). No sphere of the body can then touch
// the obstacle or the other body's spheres in any such lane -- both radii are
// inflated far beyond FP32 rounding -- and the wave skips those exact tests,
// so every verdict is unchanged. Lanes already colliding count as apart.
__device__ __forceinline__ bool wave_apart(float ax, float ay, float az, float bx, float by,
                                           float bz, float rsum, bool col) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  return __all(col || fmaf(dx, dx, fmaf(dy, dy, dz * dz)) > rsum * rsum);
}

// a body (first-slot centre x, y, z; bound rb) against an obstacle's bound
__device__ __forceinline__ bool wave_clear(float x, float y, float z, float rb,
                                           const float4 &ob, bool col) {
  return wave_apart(x, y, z, ob.x, ob.y, ob.z, rb + ob.w, col);
}

// PH bit 2 of the collide functions: the wave's configurations are close
// together in configuration space, so bounding-sphere culls are likely to
// hold for the whole wave
constexpr int kCoherent = 4;
// PH bit 3: check 32-slot plants in body groups (collides_fast_groups), which
// keeps fewer centres live and culls obstacles in chunks. It pays for EI-ZO's
// sample checks and single-wave bisection; edge points (any-collision exits)
// and bisection split over several waves measured faster without it.
constexpr int kGrouped = 8;

// Obstacle tests of bodies b0 .. b0 + KB - 1 (sphere centres cx[bl * KM + k])
// against obstacles o0, o0 + nsub, ..., culled in chunks of 32 obstacles:
// first the bounding-sphere culls of the whole chunk -- independent of each
// other, so their bound loads overlap instead of each obstacle waiting for
// its own -- give a mask of the (body, obstacle) blocks left per body; then
// only the obstacles with a block left have their records read and tested.
// row(o, b, rb) fills rb[KM] with body b's contact distances against o;
// bounds[o] is obstacle o's bounding sphere (staged in LDS or the plant's).
template <int KB, int KM, bool ANY, class Row>
__device__ __forceinline__ bool obstacles_culled(const PlantImage &P, int b0,
                                                 const float (&cx)[KB * KM],
                                                 const float (&cy)[KB * KM],
                                                 const float (&cz)[KB * KM], bool col, int o0,
                                                 int nsub, const float4 *bounds,
                                                 const Row &row) {
  for (int c0 = o0; c0 < P.num_obstacles; c0 += 32 * nsub) {
    if (wave_done<ANY>(col)) break;
    uint32_t need[KB];
#pragma unroll
    for (int bl = 0; bl < KB; ++bl) need[bl] = 0;
#pragma unroll 8
    for (int j = 0; j < 32; ++j) {
      const int o = c0 + j * nsub;
      if (o >= P.num_obstacles) continue;
      const float4 ob = bounds[o];
#pragma unroll
      for (int bl = 0; bl < KB; ++bl)
        if (b0 + bl < P.num_bodies &&
            !wave_clear(cx[bl * KM], cy[bl * KM], cz[bl * KM], P.body_bound[b0 + bl], ob, col))
          need[bl] |= 1u << j;
    }
    uint32_t todo = 0;
#pragma unroll
    for (int bl = 0; bl < KB; ++bl) todo |= need[bl];
    while (todo != 0) {
      if (wave_done<ANY>(col)) break;
      const int j = __builtin_ctz(todo);
      todo &= todo - 1;
      const int o = c0 + j * nsub;
      const int t = P.obstacle_type[o];
      float ov[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) ov[k] = P.obstacles[o].v[k];
      if (t == CSD_BOX) {
#pragma unroll
        for (int bl = 0; bl < KB; ++bl) {
          if (!((need[bl] >> j) & 1)) continue;
          float rb[KM];
          row(o, b0 + bl, rb);
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = bl * KM + k;
            col |= hit_sphere_obox(cx[s], cy[s], cz[s], rb[k], ov);
          }
        }
      } else if (t == CSD_SPHERE) {
#pragma unroll
        for (int bl = 0; bl < KB; ++bl) {
          if (!((need[bl] >> j) & 1)) continue;
          float rb[KM];
          row(o, b0 + bl, rb);
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = bl * KM + k;
            col |= hit_sphere_sphere(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], rb[k]);
          }
        }
      } else {  // capsule
#pragma unroll
        for (int bl = 0; bl < KB; ++bl) {
          if (!((need[bl] >> j) & 1)) continue;
          float rb[KM];
          row(o, b0 + bl, rb);
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = bl * KM + k;
            col |= seg_point_d2(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], ov[3], ov[4], ov[5],
                                ov[6]) < rb[k];
          }
        }
      }
    }
  }
  return col;
}

// Chain FK of the fast path fused with the sphere centres: only the current
// body frame is live, and every register index is a compile-time constant.
// RAD: also return the slot radii (kept in vector registers).
template <int NB, int KM, bool RAD, class Q>
__device__ __forceinline__ void fast_spheres(const PlantImage &P, const Q &qa,
                                             float (&cx)[NB * KM], float (&cy)[NB * KM],
                                             float (&cz)[NB * KM], float (&rad)[NB * KM]) {
  float X[12];
  aff_identity(X);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b < P.num_bodies) {
      const DevBody &bd = P.bodies[b];
      float M[12];
      // with q_identity the joint value has a compile-time index (a plain
      // register read for configurations held in registers)
      joint_local(bd, P.q_identity ? qa(b) : qa(bd.q_idx), M);
      if (bd.parent < 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) X[k] = M[k];
      } else {
        float Y[12];
        aff_mul(X, M, Y);
#pragma unroll
        for (int k = 0; k < 12; ++k) X[k] = Y[k];
      }
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        const float4 sp = P.slot_sphere[b * KM + k];
        aff_point(X, sp.x, sp.y, sp.z, cx[b * KM + k], cy[b * KM + k], cz[b * KM + k]);
        if constexpr (RAD) rad[b * KM + k] = in_vgpr(sp.w);
      }
    } else {
#pragma unroll
      for (int k = 0; k < KM; ++k)
        cx[b * KM + k] = cy[b * KM + k] = cz[b * KM + k] = rad[b * KM + k] = 0.f;
    }
  }
}

// Fast path: every moving geometry is a sphere (PlantImage::fast).
// `col` enters as the lane's initial state (true = nothing left to decide,
// e.g. an idle tail lane) and is returned OR-ed with the result.
// `tab`: for 64 sphere slots, the plant's contact-distance tables staged in
// LDS by the kernel (stage_tables); unused otherwise. (sub, nsub): this call
// tests only part `sub` of `nsub` of the work -- the body-pair blocks and
// obstacles dealt to the parts in turn, voxels v = sub mod nsub -- so that
// nsub waves can share one configuration (the caller ORs the parts).
// PH selects the pair classes tested: bit 0 the moving x moving (self)
// pairs, bit 1 the obstacles and voxels; a two-pass checker runs the cheap
// self pairs first and the obstacles only for configurations still free.
// Bit 2 (kCoherent) says the wave's configurations lie close together
// (samples of one region, points of one edge): body-pair blocks are then
// culled too, and each body's contact distances are read just before its
// obstacle tests instead of all of them per obstacle. For batches of
// unrelated configurations both cost more than they save (measured).
template <int NB, int KM, bool ANY = false, int PH = 3, class Q>
__device__ __forceinline__ bool collides_fast(const PlantImage &P, const Q &qa,
                                              const VoxelArgs &V, bool col, const float *tab,
                                              int sub = 0, int nsub = 1) {
  constexpr int NS = NB * KM;
  // Contact distances of the pair tests, by plant size:
  //  * up to 32 sphere slots: per-pair tables read into scalar registers (no
  //    extra vector work per pair; the tables fit the scalar cache);
  //  * 64 slots: the same tables, staged in LDS once per workgroup;
  //  * 128 slots: formed from the slot radii and pair masks.
  constexpr bool kTab = NS <= 32;
  constexpr bool kLds = NS == 64;
  constexpr bool kMask = !kTab && !kLds;
  constexpr bool kCoh = (PH & kCoherent) != 0;
  float cx[NS], cy[NS], cz[NS], rad[NS];  // sphere centres (world) and radii
  fast_spheres<NB, KM, kMask>(P, qa, cx, cy, cz, rad);
  col = col || P.always_colliding;

  // moving sphere x moving sphere (self collision) as body-pair blocks. With
  // tables, self_r2[ba][bb][ka][kb] holds (rA + rB)^2 (0 for absent pairs,
  // never a hit); otherwise bit ka*KM+kb of self_bits[ba][bb] marks a pair,
  // (rA + rB)^2 is formed in FP32 and absent pairs get -1.
  // (Loops over bodies keep a constant trip count with uniform guards
  // instead of `break`, so they unroll fully for 16 bodies too and every
  // register index stays static.)
  // Parts: the active body-pair blocks, then the obstacles, are dealt to the
  // nsub parts in turn (rr: the part that takes the next one), so that every
  // part gets a similar number of tests.
  int rr = 0;
#pragma unroll
  for (int ba = 0; ba < NB; ++ba) {
    if constexpr (!(PH & 1)) continue;
    const uint64_t act = P.self_active[ba];
    if (act == 0 || wave_done<ANY>(col)) continue;
#pragma unroll
    for (int bb = ba + 1; bb < NB; ++bb) {
      if ((act >> bb) & 1) {
        const bool mine = rr == sub;
        if (++rr == nsub) rr = 0;
        if (!mine) continue;
        if (kCoh && wave_apart(cx[ba * KM], cy[ba * KM], cz[ba * KM], cx[bb * KM], cy[bb * KM],
                               cz[bb * KM], P.body_bound[ba] + P.body_bound[bb], col))
          continue;
        if constexpr (kTab || kLds) {
          const float *r2 = (kLds ? tab : P.self_r2) + ((size_t)ba * NB + bb) * (KM * KM);
#pragma unroll
          for (int ka = 0; ka < KM; ++ka)
#pragma unroll
            for (int kb = 0; kb < KM; ++kb) {
              const int sa = ba * KM + ka, sb = bb * KM + kb;
              col |= hit_sphere_sphere(cx[sa], cy[sa], cz[sa], cx[sb], cy[sb], cz[sb],
                                       r2[ka * KM + kb]);
            }
        } else {
          const uint64_t bits = P.self_bits[ba * NB + bb];
#pragma unroll
          for (int ka = 0; ka < KM; ++ka)
#pragma unroll
            for (int kb = 0; kb < KM; ++kb) {
              const int sa = ba * KM + ka, sb = bb * KM + kb;
              const float rs = rad[sa] + rad[sb];
              const float r2 = ((bits >> (ka * KM + kb)) & 1) ? rs * rs : -1.f;
              col |= hit_sphere_sphere(cx[sa], cy[sa], cz[sa], cx[sb], cy[sb], cz[sb], r2);
            }
        }
      }
    }
  }

  // the part that takes obstacle 0: obstacle o goes to part (rr + o) mod nsub
  const int o0 = sub >= rr ? sub - rr : sub - rr + nsub;

  // moving sphere x world-fixed obstacle, tables: the record and its row of
  // contact distances obstacle_r2[o][s] (box r^2, sphere/capsule
  // (r + r_o)^2, -1 if the slot is not paired with o) are read once per
  // obstacle with a few wide loads (scalar loads for one part; with parts
  // derived from the wave index the compiler cannot prove o uniform and
  // issues vector loads, which all lanes share through the L1 -- forcing
  // scalar loads there measured slower); the slot loop has no branch. With
  // kCoherent the row is read body by body, for the bodies not culled. (The
  // chunked culls of obstacles_culled take this path, with all 32 centres
  // live, to 256 VGPRs and one wave per SIMD, so it culls obstacle by
  // obstacle.)
  if constexpr (kTab && (PH & 2) && kCoh) {
    for (int o = o0; o < P.num_obstacles; o += nsub) {
      if (wave_done<ANY>(col)) break;
      const int t = P.obstacle_type[o];
      const float4 ob = P.obstacle_bound[o];
      float ov[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) ov[k] = P.obstacles[o].v[k];
      const float *orow = P.obstacle_r2 + (size_t)o * NS;
      auto skip = [&](int b) {
        return b >= P.num_bodies ||
               wave_clear(cx[b * KM], cy[b * KM], cz[b * KM], P.body_bound[b], ob, col);
      };
      if (t == CSD_BOX) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
          float rb[KM];
#pragma unroll
          for (int k = 0; k < KM; ++k) rb[k] = orow[b * KM + k];
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= hit_sphere_obox(cx[s], cy[s], cz[s], rb[k], ov);
          }
        }
      } else if (t == CSD_SPHERE) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
          float rb[KM];
#pragma unroll
          for (int k = 0; k < KM; ++k) rb[k] = orow[b * KM + k];
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= hit_sphere_sphere(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], rb[k]);
          }
        }
      } else {  // capsule
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
          float rb[KM];
#pragma unroll
          for (int k = 0; k < KM; ++k) rb[k] = orow[b * KM + k];
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= seg_point_d2(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], ov[3], ov[4], ov[5],
                                ov[6]) < rb[k];
          }
        }
      }
    }
  }
  if constexpr (kTab && (PH & 2) && !kCoh) {
    for (int o = o0; o < P.num_obstacles; o += nsub) {
      if (wave_done<ANY>(col)) break;
      const int t = P.obstacle_type[o];
      const float4 ob = P.obstacle_bound[o];
      float ov[16], r2[NS];
#pragma unroll
      for (int k = 0; k < 16; ++k) ov[k] = P.obstacles[o].v[k];
#pragma unroll
      for (int s = 0; s < NS; ++s) r2[s] = P.obstacle_r2[(size_t)o * NS + s];
      auto skip = [&](int b) {
        return b >= P.num_bodies ||
               wave_clear(cx[b * KM], cy[b * KM], cz[b * KM], P.body_bound[b], ob, col);
      };
      if (t == CSD_BOX) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= hit_sphere_obox(cx[s], cy[s], cz[s], r2[s], ov);
          }
        }
      } else if (t == CSD_SPHERE) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= hit_sphere_sphere(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], r2[s]);
          }
        }
      } else {  // capsule
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= seg_point_d2(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], ov[3], ov[4], ov[5],
                                ov[6]) < r2[s];
          }
        }
      }
    }
  }

  // moving sphere x world-fixed obstacle, LDS tables: the record in scalar
  // registers, each body's row of contact distances read from LDS just before
  // its tests (one 16-byte read per 4 slots).
  if constexpr (kLds && (PH & 2)) {
    const float *orow = tab + NB * NB * KM * KM;
    for (int o = o0; o < P.num_obstacles; o += nsub) {
      if (wave_done<ANY>(col)) break;
      const int t = P.obstacle_type[o];
      const float4 ob = P.obstacle_bound[o];
      float ov[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) ov[k] = P.obstacles[o].v[k];
      const float *r2 = orow + (size_t)o * NS;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (b >= P.num_bodies ||
            wave_clear(cx[b * KM], cy[b * KM], cz[b * KM], P.body_bound[b], ob, col))
          continue;
        float rb[KM];
#pragma unroll
        for (int k = 0; k < KM; ++k) rb[k] = r2[b * KM + k];
        if (t == CSD_BOX) {
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= hit_sphere_obox(cx[s], cy[s], cz[s], rb[k], ov);
          }
        } else if (t == CSD_SPHERE) {
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= hit_sphere_sphere(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], rb[k]);
          }
        } else {
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            col |= seg_point_d2(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], ov[3], ov[4], ov[5],
                                ov[6]) < rb[k];
          }
        }
      }
    }
  }

  // moving sphere x world-fixed obstacle, masks. Per obstacle only its 64-byte record and two mask words are read (scalar
  // loads, issued one obstacle ahead so their latency hides behind the
  // current obstacle's tests). Contact distances are formed in FP32 as
  // r * r (box) and (r + r_o) * (r + r_o) (sphere, capsule); a slot not
  // paired with the obstacle gets -1, which never reports a hit, so there is
  // no per-slot branch. Slots of absent bodies are skipped per body.
  if (kMask && (PH & 2) && o0 < P.num_obstacles) {
    int tn = P.obstacle_type[o0];
    float ovn[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) ovn[k] = P.obstacles[o0].v[k];
    uint64_t m0n = P.obstacle_mask[2 * o0], m1n = P.obstacle_mask[2 * o0 + 1];
    for (int o = o0; o < P.num_obstacles; o += nsub) {
      if (wave_done<ANY>(col)) break;
      const int t = tn;
      float ov[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) ov[k] = ovn[k];
      const uint64_t m0 = m0n, m1 = m1n;
      const int onext = o + nsub < P.num_obstacles ? o + nsub : o;
      tn = P.obstacle_type[onext];
#pragma unroll
      for (int k = 0; k < 16; ++k) ovn[k] = P.obstacles[onext].v[k];
      m0n = P.obstacle_mask[2 * onext];
      m1n = P.obstacle_mask[2 * onext + 1];
      auto paired = [&](int s) { return s < 64 ? ((m0 >> s) & 1) : ((m1 >> (s - 64)) & 1); };
      const float4 ob = P.obstacle_bound[o];
      auto skip = [&](int b) {
        return b >= P.num_bodies ||
               wave_clear(cx[b * KM], cy[b * KM], cz[b * KM], P.body_bound[b], ob, col);
      };
      if (t == CSD_BOX) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            const float r2 = paired(s) ? rad[s] * rad[s] : -1.f;
            col |= hit_sphere_obox(cx[s], cy[s], cz[s], r2, ov);
          }
        }
      } else if (t == CSD_SPHERE) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            const float rs = rad[s] + ov[3];
            const float r2 = paired(s) ? rs * rs : -1.f;
            col |= hit_sphere_sphere(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], r2);
          }
        }
      } else {  // capsule
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (skip(b)) continue;
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const int s = b * KM + k;
            const float rs = rad[s] + ov[7];
            const float r2 = paired(s) ? rs * rs : -1.f;
            col |= seg_point_d2(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], ov[3], ov[4], ov[5],
                                ov[6]) < r2;
          }
        }
      }
    }
  }

  // robot spheres x voxel spheres (executeCollisionFreeVoxelsKernel); the
  // per-slot contact distances are formed once, -1 for slots not checked
  if ((PH & 2) && V.num > 0) {
    float vr2[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const bool on = s < 64 ? ((V.mask0 >> s) & 1) : ((V.mask1 >> (s - 64)) & 1);
      const float rs = V.radius + P.slot_sphere[s].w;
      vr2[s] = on ? rs * rs : -1.f;
    }
    for (int v = sub; v < V.num; v += nsub) {
      if (wave_done<ANY>(col)) break;
      const float vx = V.vox[3 * v], vy = V.vox[3 * v + 1], vz = V.vox[3 * v + 2];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        col |= hit_sphere_sphere(vx, vy, vz, cx[s], cy[s], cz[s], vr2[s]);
    }
  }
  return col;
}

// 64 sphere slots (tables in LDS) checked in G groups of NB / G bodies, so
// that only one group's sphere centres are live at a time (192 / G
// registers instead of 192). Group g continues the chain FK into its bodies
// [g NB/G, (g+1) NB/G): their centres and the self pairs among them; then
// the bodies of the earlier groups are walked again one at a time (the same
// FK arithmetic, so the same centres) for the pairs between them and group
// g; then group g's obstacles and voxels. Every test is collides_fast's own,
// so the verdict is the same; only the order of the tests differs. (sub,
// nsub) and PH as in collides_fast.
template <int NB, int KM, int G, bool ANY = false, int PH = 3, class Q>
__device__ __forceinline__ bool collides_fast_groups(const PlantImage &P, const Q &qa,
                                                     const VoxelArgs &V, bool col,
                                                     const float *tab, int sub = 0,
                                                     int nsub = 1) {
  constexpr int KG = NB / G, NSG = KG * KM, NS = NB * KM;
  constexpr bool kCoh = (PH & kCoherent) != 0;
  static_assert(NS <= 64 && KG * G == NB, "the table plants, whole groups");
  col = col || P.always_colliding;
  // contact-distance tables and obstacle bounds: staged in LDS (tab) for 64
  // slots, the plant's own arrays (same layouts) for 32
  const float *stab = NS == 64 ? tab : P.self_r2;
  const float *orow = NS == 64 ? tab + NB * NB * KM * KM : P.obstacle_r2;
  const float4 *obnd = NS == 64 ? reinterpret_cast<const float4 *>(orow + (size_t)P.num_obstacles * NS)
                                : P.obstacle_bound;
  int rr = 0;  // the part that takes the next body-pair block
  auto mine = [&]() {
    const bool m = rr == sub;
    if (++rr == nsub) rr = 0;
    return m;
  };
  float X[12];  // the chain's current body frame
  aff_identity(X);
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int b0 = g * KG;
    float cx[NSG], cy[NSG], cz[NSG];
#pragma unroll
    for (int bl = 0; bl < KG; ++bl) {
      const int b = b0 + bl;
      if (b < P.num_bodies) {
        const DevBody &bd = P.bodies[b];
        float M[12];
        joint_local(bd, P.q_identity ? qa(b) : qa(bd.q_idx), M);
        if (bd.parent < 0) {
#pragma unroll
          for (int k = 0; k < 12; ++k) X[k] = M[k];
        } else {
          float Y[12];
          aff_mul(X, M, Y);
#pragma unroll
          for (int k = 0; k < 12; ++k) X[k] = Y[k];
        }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          const float4 sp = P.slot_sphere[b * KM + k];
          aff_point(X, sp.x, sp.y, sp.z, cx[bl * KM + k], cy[bl * KM + k], cz[bl * KM + k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < KM; ++k) cx[bl * KM + k] = cy[bl * KM + k] = cz[bl * KM + k] = 0.f;
      }
    }
    if constexpr (PH & 1) {
      // self pairs inside this group
#pragma unroll
      for (int bla = 0; bla < KG; ++bla) {
        const int ba = b0 + bla;
        const uint64_t act = P.self_active[ba];
        if (act == 0 || wave_done<ANY>(col)) continue;
        // constant trip counts (uniform guards), so that every loop unrolls
#pragma unroll
        for (int blb = 0; blb < KG; ++blb) {
          const int bb = b0 + blb;
          if (blb <= bla || !((act >> bb) & 1) || !mine()) continue;
          if (kCoh && wave_apart(cx[bla * KM], cy[bla * KM], cz[bla * KM], cx[blb * KM],
                                 cy[blb * KM], cz[blb * KM], P.body_bound[ba] + P.body_bound[bb],
                                 col))
            continue;
          const float *r2 = stab + ((size_t)ba * NB + bb) * (KM * KM);
#pragma unroll
          for (int ka = 0; ka < KM; ++ka)
#pragma unroll
            for (int kb = 0; kb < KM; ++kb) {
              const int sa = bla * KM + ka, sb = blb * KM + kb;
              col |= hit_sphere_sphere(cx[sa], cy[sa], cz[sa], cx[sb], cy[sb], cz[sb],
                                       r2[ka * KM + kb]);
            }
        }
      }
      // pairs between the earlier groups and this one: those bodies again,
      // one at a time
      if (g > 0) {
        float W[12];
        aff_identity(W);
#pragma unroll
        for (int ba = 0; ba < NB - KG; ++ba) {
          if (ba >= b0 || ba >= P.num_bodies) continue;
          const DevBody &bd = P.bodies[ba];
          float M[12];
          joint_local(bd, P.q_identity ? qa(ba) : qa(bd.q_idx), M);
          if (bd.parent < 0) {
#pragma unroll
            for (int k = 0; k < 12; ++k) W[k] = M[k];
          } else {
            float Y[12];
            aff_mul(W, M, Y);
#pragma unroll
            for (int k = 0; k < 12; ++k) W[k] = Y[k];
          }
          const uint64_t act = (P.self_active[ba] >> b0) & ((1ull << KG) - 1);
          if (act == 0 || wave_done<ANY>(col)) continue;
          float ax[KM], ay[KM], az[KM];
#pragma unroll
          for (int k = 0; k < KM; ++k) {
            const float4 sp = P.slot_sphere[ba * KM + k];
            aff_point(W, sp.x, sp.y, sp.z, ax[k], ay[k], az[k]);
          }
#pragma unroll
          for (int blb = 0; blb < KG; ++blb) {
            if (!((act >> blb) & 1) || !mine()) continue;
            if (kCoh && wave_apart(ax[0], ay[0], az[0], cx[blb * KM], cy[blb * KM], cz[blb * KM],
                                   P.body_bound[ba] + P.body_bound[b0 + blb], col))
              continue;
            const float *r2 = stab + ((size_t)ba * NB + b0 + blb) * (KM * KM);
#pragma unroll
            for (int ka = 0; ka < KM; ++ka)
#pragma unroll
              for (int kb = 0; kb < KM; ++kb) {
                const int sb = blb * KM + kb;
                col |= hit_sphere_sphere(ax[ka], ay[ka], az[ka], cx[sb], cy[sb], cz[sb],
                                         r2[ka * KM + kb]);
              }
          }
        }
      }
    }
    // this group's spheres x world-fixed obstacles in culled chunks: the
    // record in scalar registers, each body's row of contact distances from
    // LDS. (Culling obstacle by obstacle, each iteration waited for its own
    // bound and record loads: dual-arm bisection took 15.8 ms per call at 64
    // segments, 11% VALU-busy; in chunks it takes 9.1 ms.)
    if constexpr (PH & 2) {
      const int o0 = sub >= rr ? sub - rr : sub - rr + nsub;
      col = obstacles_culled<KG, KM, ANY>(P, b0, cx, cy, cz, col, o0, nsub, obnd,
                                          [&](int o, int b, float *rb) {
                                            const float *r2 = orow + (size_t)o * NS + b * KM;
#pragma unroll
                                            for (int k = 0; k < KM; ++k) rb[k] = r2[k];
                                          });
      // this group's spheres x voxel spheres
      if (V.num > 0) {
        float vr2[NSG];
#pragma unroll
        for (int sl = 0; sl < NSG; ++sl) {
          const int s = b0 * KM + sl;
          const bool on = (V.mask0 >> s) & 1;
          const float rs = V.radius + P.slot_sphere[s].w;
          vr2[sl] = on ? rs * rs : -1.f;
        }
        for (int v = sub; v < V.num; v += nsub) {
          if (wave_done<ANY>(col)) break;
          const float vx = V.vox[3 * v], vy = V.vox[3 * v + 1], vz = V.vox[3 * v + 2];
#pragma unroll
          for (int sl = 0; sl < NSG; ++sl)
            col |= hit_sphere_sphere(vx, vy, vz, cx[sl], cy[sl], cz[sl], vr2[sl]);
        }
      }
    }
  }
  return col;
}

// Smallest signed distance over the plant's collision pairs in FP32 (fast
// path plants): sphere-sphere |c_a - c_b| - r_a - r_b, sphere-box distance of
// the centre to the box (negative inside) - r, sphere-capsule distance to the
// segment - r_a - r_c. Voxels are not included. A plant with a colliding
// world-fixed pair reports -inf. Used to label training data for the
// signed-distance surrogate.
template <int NB, int KM, class Q>
__device__ __forceinline__ float min_signed_distance_fast(const PlantImage &P, const Q &qa) {
  constexpr int NS = NB * KM;
  float cx[NS], cy[NS], cz[NS], rad[NS];
  fast_spheres<NB, KM, true>(P, qa, cx, cy, cz, rad);
  float sd = P.always_colliding ? -INFINITY : INFINITY;
#pragma unroll
  for (int ba = 0; ba < NB; ++ba) {
    const uint64_t act = P.self_active[ba];
    if (act == 0) continue;
#pragma unroll
    for (int bb = ba + 1; bb < NB; ++bb) {
      if ((act >> bb) & 1) {
        const uint64_t bits = P.self_bits[ba * NB + bb];
#pragma unroll
        for (int ka = 0; ka < KM; ++ka)
#pragma unroll
          for (int kb = 0; kb < KM; ++kb) {
            const int sa = ba * KM + ka, sb = bb * KM + kb;
            const float dx = cx[sa] - cx[sb], dy = cy[sa] - cy[sb], dz = cz[sa] - cz[sb];
            const float d = sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz))) - (rad[sa] + rad[sb]);
            sd = ((bits >> (ka * KM + kb)) & 1) ? fminf(sd, d) : sd;
          }
      }
    }
  }
  for (int o = 0; o < P.num_obstacles; ++o) {
    const int t = P.obstacle_type[o];
    float ov[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) ov[k] = P.obstacles[o].v[k];
    const uint64_t m0 = P.obstacle_mask[2 * o], m1 = P.obstacle_mask[2 * o + 1];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b >= P.num_bodies) continue;
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        const int s = b * KM + k;
        float d;
        if (t == CSD_BOX) {
          const float lx = fmaf(ov[0], cx[s], fmaf(ov[1], cy[s], fmaf(ov[2], cz[s], ov[9])));
          const float ly = fmaf(ov[3], cx[s], fmaf(ov[4], cy[s], fmaf(ov[5], cz[s], ov[10])));
          const float lz = fmaf(ov[6], cx[s], fmaf(ov[7], cy[s], fmaf(ov[8], cz[s], ov[11])));
          const float qx = fabsf(lx) - ov[12], qy = fabsf(ly) - ov[13], qz = fabsf(lz) - ov[14];
          const float ex = fmaxf(qx, 0.f), ey = fmaxf(qy, 0.f), ez = fmaxf(qz, 0.f);
          d = sqrtf(fmaf(ex, ex, fmaf(ey, ey, ez * ez))) +
              fminf(fmaxf(qx, fmaxf(qy, qz)), 0.f) - rad[s];
        } else if (t == CSD_SPHERE) {
          const float dx = cx[s] - ov[0], dy = cy[s] - ov[1], dz = cz[s] - ov[2];
          d = sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz))) - (rad[s] + ov[3]);
        } else {
          d = sqrtf(seg_point_d2(cx[s], cy[s], cz[s], ov[0], ov[1], ov[2], ov[3], ov[4], ov[5],
                                 ov[6])) -
              (rad[s] + ov[7]);
        }
        const bool on = s < 64 ? ((m0 >> s) & 1) : ((m1 >> (s - 64)) & 1);
        sd = on ? fminf(sd, d) : sd;
      }
    }
  }
  return sd;
}

// Generic path: any sphere / box / capsule pair, full 3x4 geometry poses.
template <int NB, bool ANY = false, class Q>
__device__ __forceinline__ bool collides_generic(const PlantImage &P, const Q &qa,
                                                 const VoxelArgs &V, const RobotGeomList &R,
                                                 bool col) {
  CSD_BODY_FRAMES(B, NB);
  fk_bodies<NB>(P, qa, B);
  col = col || P.always_colliding;
  for (int p = 0; p < P.num_pairs; ++p) {
    if ((p & 7) == 0 && wave_done<ANY>(col)) break;
    const DevGeom &ga = P.geoms[P.pairs[2 * p]];
    const DevGeom &gb = P.geoms[P.pairs[2 * p + 1]];
    float XA[12], XB[12];
    geom_world<NB>(ga, B, XA);
    geom_world<NB>(gb, B, XB);
    col |= hit_posed(geom_rank(ga.type), XA, ga.dims, geom_rank(gb.type), XB, gb.dims);
  }
  for (int r = 0; r < R.n; ++r) {
    if (wave_done<ANY>(col)) break;
    const DevGeom &g = P.geoms[R.ids[r]];
    float X[12];
    geom_world<NB>(g, B, X);
    const int rk = geom_rank(g.type);
    for (int v = 0; v < V.num; ++v) {
      float Y[12];
      aff_identity(Y);
      Y[3] = V.vox[3 * v];
      Y[7] = V.vox[3 * v + 1];
      Y[11] = V.vox[3 * v + 2];
      const float vd[4] = {V.radius, 0.f, 0.f, 0.f};
      col |= hit_posed(0, Y, vd, rk, X, g.dims);
    }
  }
  return col;
}

// Kernel-variant selector shared by every launcher: which template the plant
// needs. 0..3 fast <8,4>,<8,8>,<16,4>,<16,8>; 4,5 generic <8>,<16>.
enum CheckVariant { kFast8x4 = 0, kFast8x8, kFast16x4, kFast16x8, kGeneric8, kGeneric16 };

__host__ __device__ constexpr int variant_slots(int var) {
  return var == kFast8x4 ? 32 : (var == kFast8x8 || var == kFast16x4) ? 64
                                : var == kFast16x8 ? 128 : 0;
}

// LDS bytes of the contact-distance tables a 64-slot variant stages
constexpr size_t kCheckLdsMax = 56 * 1024;
inline size_t check_lds_bytes(const PlantImage &P, int var) {
  if (variant_slots(var) != 64) return 0;
  const size_t nbt = P.nslots / P.kmax;
  return (nbt * nbt * P.kmax * P.kmax + (size_t)P.num_obstacles * (64 + 4)) * sizeof(float);
}

// the fast shape of a fast-path plant
inline int fast_variant(const PlantImage &P) {
  const int nbt = P.nslots / P.kmax;
  if (nbt == 8) return P.kmax == 4 ? kFast8x4 : kFast8x8;
  return P.kmax == 4 ? kFast16x4 : kFast16x8;
}

// the checker a launch uses: the fast shape unless generic geometry is
// needed or its staged tables would not fit in LDS
inline int check_variant(const PlantImage &P, bool need_generic) {
  if (P.fast && !need_generic) {
    const int v = fast_variant(P);
    if (check_lds_bytes(P, v) <= kCheckLdsMax) return v;
  }
  return P.num_bodies <= 8 ? kGeneric8 : kGeneric16;
}

// Copies the contact-distance tables of a 64-slot variant into LDS (self
// pairs, then obstacle rows, then the obstacles' bounding spheres as float4)
// and returns them; every thread of the block must call it. Other variants
// need nothing (nullptr).
template <int VAR>
__device__ __forceinline__ const float *stage_tables(const PlantImage &P, float *smem) {
  if constexpr (variant_slots(VAR) == 64) {
    constexpr int NB = VAR == kFast8x8 ? 8 : 16, KM = VAR == kFast8x8 ? 8 : 4;
    constexpr int ns = NB * NB * KM * KM;
    const int no = P.num_obstacles * 64;
    for (int e = threadIdx.x; e < ns; e += blockDim.x) smem[e] = P.self_r2[e];
    for (int e = threadIdx.x; e < no; e += blockDim.x) smem[ns + e] = P.obstacle_r2[e];
    const float *ob = reinterpret_cast<const float *>(P.obstacle_bound);
    for (int e = threadIdx.x; e < 4 * P.num_obstacles; e += blockDim.x) smem[ns + no + e] = ob[e];
    __syncthreads();
    return smem;
  }
  return nullptr;
}

// dispatch a device-side call on the variant (VAR is a compile-time value
// inside the generic lambda-free macro body)
// (sub, nsub), PH: see collides_fast; the generic checker ignores the split
// (part 0 does all of the work, other parts report no collision) and PH
// (it always tests every pair).
// body groups of the 64-slot layouts (collides_fast_groups): two bodies
// (8 or 16 spheres) per group. Plain dual-arm EI-ZO, 64 segments: 92.8
// regions/s with all 64 centres live (one wave per SIMD), 144.9 with 2
// groups; 256 segments: 157.3 / 156.9 / 160.4 regions/s with 2 / 4 / 8
// groups of the 16 x 4 layout (check 203 / 144 / 120 VGPRs, bisection
// 235 / 181 / 156).
constexpr int kGroups16x4 = 8, kGroups8x8 = 4;
// 32 slots with kGrouped: 4 groups of two bodies, so that the chunked
// obstacle culls fit in registers
constexpr int kGroups8x4 = 4;

template <int VAR, bool ANY = false, int PH = 3, class Q>
__device__ __forceinline__ bool collides_variant(const PlantImage &P, const Q &qa,
                                                 const VoxelArgs &V, const RobotGeomList &R,
                                                 bool col, const float *tab, int sub = 0,
                                                 int nsub = 1) {
  if constexpr (VAR == kFast8x4 && (PH & kGrouped))
    return collides_fast_groups<8, 4, kGroups8x4, ANY, PH>(P, qa, V, col, tab, sub, nsub);
  else if constexpr (VAR == kFast8x4) return collides_fast<8, 4, ANY, PH>(P, qa, V, col, tab, sub, nsub);
  else if constexpr (VAR == kFast8x8) return collides_fast_groups<8, 8, kGroups8x8, ANY, PH>(P, qa, V, col, tab, sub, nsub);
  else if constexpr (VAR == kFast16x4) return collides_fast_groups<16, 4, kGroups16x4, ANY, PH>(P, qa, V, col, tab, sub, nsub);
  else if constexpr (VAR == kFast16x8) return collides_fast<16, 8, ANY, PH>(P, qa, V, col, tab, sub, nsub);
  else if (sub != 0) return col;
  else if constexpr (VAR == kGeneric8) return collides_generic<8, ANY>(P, qa, V, R, col);
  else return collides_generic<16, ANY>(P, qa, V, R, col);
}

// collides_variant on configuration i of a strided array; fast shapes load
// the joint values into registers first (one memory wait before FK)
template <int VAR, bool ANY = false, int PH = 3>
__device__ __forceinline__ bool collides_at(const PlantImage &P, const QMem &qm,
                                            const VoxelArgs &V, const RobotGeomList &R,
                                            bool col, const float *tab, int sub = 0,
                                            int nsub = 1) {
  if constexpr (VAR == kFast8x4 || VAR == kFast8x8)
    return collides_variant<VAR, ANY, PH>(P, preload_q<8>(qm, P.dof), V, R, col, tab, sub, nsub);
  else if constexpr (VAR == kFast16x4 || VAR == kFast16x8)
    return collides_variant<VAR, ANY, PH>(P, preload_q<16>(qm, P.dof), V, R, col, tab, sub, nsub);
  else
    return collides_variant<VAR, ANY, PH>(P, qm, V, R, col, tab, sub, nsub);
}

// host: launch KERNEL<VAR> for the runtime variant
#define CSD_DISPATCH_VARIANT(var, KERNEL, grid, block, shmem, stream, ...)                 \
  do {                                                                                    \
    switch (var) {                                                                        \
      case ::csd::kFast8x4:                                                               \
        hipLaunchKernelGGL((KERNEL<::csd::kFast8x4>), grid, block, shmem, stream,        \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      case ::csd::kFast8x8:                                                               \
        hipLaunchKernelGGL((KERNEL<::csd::kFast8x8>), grid, block, shmem, stream,        \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      case ::csd::kFast16x4:                                                              \
        hipLaunchKernelGGL((KERNEL<::csd::kFast16x4>), grid, block, shmem, stream,       \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      case ::csd::kFast16x8:                                                              \
        hipLaunchKernelGGL((KERNEL<::csd::kFast16x8>), grid, block, shmem, stream,       \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      case ::csd::kGeneric8:                                                              \
        hipLaunchKernelGGL((KERNEL<::csd::kGeneric8>), grid, block, shmem, stream,       \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      default:                                                                            \
        hipLaunchKernelGGL((KERNEL<::csd::kGeneric16>), grid, block, shmem, stream,      \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
    }                                                                                     \
  } while (0)

// host: launch KERNEL<VAR, X> for the runtime variant (X: a compile-time
// integer of the caller, e.g. a padded dimension)
#define CSD_DISPATCH_VARIANT_X(var, KERNEL, X, grid, block, shmem, stream, ...)             \
  do {                                                                                    \
    switch (var) {                                                                        \
      case ::csd::kFast8x4:                                                               \
        hipLaunchKernelGGL((KERNEL<::csd::kFast8x4, X>), grid, block, shmem, stream,     \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      case ::csd::kFast8x8:                                                               \
        hipLaunchKernelGGL((KERNEL<::csd::kFast8x8, X>), grid, block, shmem, stream,     \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      case ::csd::kFast16x4:                                                              \
        hipLaunchKernelGGL((KERNEL<::csd::kFast16x4, X>), grid, block, shmem, stream,    \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      case ::csd::kFast16x8:                                                              \
        hipLaunchKernelGGL((KERNEL<::csd::kFast16x8, X>), grid, block, shmem, stream,    \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      case ::csd::kGeneric8:                                                              \
        hipLaunchKernelGGL((KERNEL<::csd::kGeneric8, X>), grid, block, shmem, stream,    \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
      default:                                                                            \
        hipLaunchKernelGGL((KERNEL<::csd::kGeneric16, X>), grid, block, shmem, stream,   \
                           __VA_ARGS__);                                                  \
        break;                                                                            \
    }                                                                                     \
  } while (0)

// host: robot geometry list and voxel slot masks; returns false if a robot
// geometry is not a fast-path sphere slot
bool fill_robot_list(const csd_plant *pl, const int32_t *robot, int32_t nr, RobotGeomList &R,
                     VoxelArgs &V, int *status);

}  // namespace csd
