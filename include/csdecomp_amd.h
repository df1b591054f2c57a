/*
 * csdecomp_amd.h -- C-ABI of the MI355X (gfx950) implementation of
 * csdecomp's data-parallel hot path: batched forward kinematics, collision
 * checking of sphere/box/capsule geometry, hit-and-run sampling and EI-ZO
 * edge inflation.
 *
 * Conventions
 *  - Every entry point returns an int status (CSD_OK == 0); csd_status_string
 *    names it. Nothing in this library calls exit().
 *  - Pointers documented "device" are HBM pointers the caller owns; "host"
 *    pointers are read synchronously before the call returns.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *    Kernels are launched asynchronously on it unless documented otherwise.
 *  - Configurations are addressed with two strides: element (config n, joint
 *    j) lives at q[n * stride_n + j * stride_j]. Structure-of-arrays
 *    (stride_n = 1, stride_j = N) gives coalesced reads; the reference's
 *    (dim, N) column-major Eigen layout is stride_n = dim, stride_j = 1.
 *  - 4x4 transforms are column-major (Eigen default), as in the reference.
 */
#ifndef CSDECOMP_AMD_H
#define CSDECOMP_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CSD_OK = 0,
  CSD_ERR_INVALID_ARGUMENT = 1,
  CSD_ERR_CAPACITY = 2,      /* a fixed capacity was exceeded */
  CSD_ERR_HIP = 3,           /* a HIP runtime call failed */
  CSD_ERR_NO_DEVICE = 4,     /* no gfx950 device visible */
  CSD_ERR_UNSUPPORTED = 5,   /* geometry / option combination unsupported */
  CSD_ERR_OUT_OF_MEMORY = 6, /* hipMalloc failed */
};

enum { CSD_JOINT_FIXED = 0, CSD_JOINT_REVOLUTE = 1, CSD_JOINT_PRISMATIC = 2 };
enum { CSD_BOX = 0, CSD_SPHERE = 1, CSD_CYLINDER = 2, CSD_CAPSULE = 3 };
enum { CSD_GROUP_ROBOT = 0, CSD_GROUP_WORLD = 1, CSD_GROUP_VOXELS = 2 };

/* Host-side description of a plant. Field-for-field the content of the
 * reference's MinimalPlant (plant/minimal_plant.h:16-39, MinimalKinematicTree
 * plant/minimal_kinematic_tree.h:42-53, CollisionGeometry
 * plant/collision_geometry.h:27-35) without the compile-time capacities. */
typedef struct csd_plant_desc {
  int32_t num_links;
  int32_t num_joints;
  int32_t num_configuration_variables;
  const int32_t *joint_type;        /* [J] CSD_JOINT_* */
  const int32_t *joint_parent_link; /* [J] */
  const int32_t *joint_child_link;  /* [J] */
  const int32_t *joint_config_idx;  /* [J], -1 for fixed joints */
  const int32_t *traversal;         /* [J] joint_traversal_sequence */
  const float *joint_X_PL_J;        /* [J][16] column-major */
  const float *joint_axis;          /* [J][3] unit axis */
  int32_t num_geometries;
  const int32_t *geom_type;  /* [G] CSD_BOX / CSD_SPHERE / CSD_CAPSULE */
  const int32_t *geom_link;  /* [G] link index */
  const int32_t *geom_group; /* [G] CSD_GROUP_* */
  const float *geom_dims;    /* [G][3] box: full edge lengths; sphere: (r,-1,-1);
                                capsule: (radius, length, radius), axis = z */
  const float *geom_X_L_B;   /* [G][16] column-major */
  int32_t num_collision_pairs;
  const int32_t *collision_pairs; /* [P][2] geometry indices (A, B) */
} csd_plant_desc;

typedef struct csd_plant csd_plant; /* opaque, device resident */

/* ---- library ---- */
const char *csd_status_string(int status);
/* Selects the device and checks it is gfx950. Replaces the implicit CUDA
 * context setup of utils/cuda_utilities.h:28-64 (CudaPtr). */
int csd_init(int device);
/* Last error message recorded on this thread (for exceptions). */
const char *csd_last_error(void);

/* Device memory and stream helpers, so host code above this boundary never
 * calls HIP directly. Replace CudaPtr (utils/cuda_utilities.h:28-64); errors
 * are returned, never exit()ed (the reference's CUDACHECKERROR,
 * utils/cuda_utilities.h:11-20, terminates the process). */
int csd_malloc(void **ptr, size_t bytes);
int csd_free(void *ptr);
int csd_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int csd_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int csd_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int csd_memset(void *dst, int value, size_t bytes, void *stream);
int csd_stream_synchronize(void *stream);
int csd_stream_create(void **stream);
int csd_stream_destroy(void *stream);

/* ---- plant ----
 * Replaces the per-call `CudaPtr<const MinimalPlant>` upload
 * (plant/cuda_collision_checker.cu:455-462): the plant is compiled once into
 * a compact device image (fixed joints folded, world-fixed geometry
 * pre-transformed, pairs grouped by type). */
int csd_plant_create(const csd_plant_desc *desc, csd_plant **out);
/* Waits for the device (kernels may still read the image), then frees. */
int csd_plant_destroy(csd_plant *plant);
int csd_plant_info(const csd_plant *plant, int32_t *num_links,
                   int32_t *num_configuration_variables,
                   int32_t *num_collision_pairs, int32_t *fast_path);

/* ---- forward kinematics ----
 * Replaces executeForwardKinematicsKernel (plant/cuda_forward_kinematics.h:27-30)
 * and computeForwardKinematicsCuda (:53-55). transforms (device):
 * [N][num_links][16] column-major, as the reference's transforms_flat. */
int csd_forward_kinematics(const csd_plant *plant, const float *q, int64_t n,
                           int64_t stride_n, int64_t stride_j,
                           float *transforms, void *stream);

/* ---- collision checking ----
 * Replaces executeCollisionFreeKernel + poolPairCollisionFreeResultsKernel
 * (plant/cuda_collision_checker.cu:16-85,182-209) and checkCollisionFreeCuda
 * (:437-487): FK and all pair tests fused, no per-pair intermediates.
 * out (device): [N] uint8, 1 = collision free. */
int csd_check_collision_free(const csd_plant *plant, const float *q, int64_t n,
                             int64_t stride_n, int64_t stride_j, uint8_t *out,
                             void *stream);

/* Autotuner of csd_check_collision_free (no counterpart in the reference):
 * times the kernel on the caller's batch in six shapes and keeps the fastest
 * for this plant and batch-size class (same bit length of n):
 *   1, 2, 4    -- one pass, that many waves sharing each group of 64
 *                 configurations (each wave taking part of the pair tests;
 *                 small batches otherwise leave SIMDs idle);
 *   11, 12, 14 -- two passes: FK + robot self pairs with 1, 2 or 4 waves per
 *                 64 configurations, then obstacles and voxels only for the
 *                 configurations still free, compacted per workgroup (pays
 *                 when most configurations collide with the robot itself);
 *   20         -- screened (plants of at most 4 sphere slots per body): the
 *                 pair tests that hit most often in an even sample of up to
 *                 4,096 of the caller's configurations (greedy set cover)
 *                 run first for every configuration; the configurations
 *                 they leave undecided get the full check, compacted per
 *                 workgroup with 8 waves sharing each 64. The flags are the
 *                 unscreened ones bit for bit; only the speed depends on
 *                 the sample. Checks with voxels stay unscreened.
 * out receives the flags. Synchronizes `stream`. *split (may be NULL)
 * reports the choice; if it holds one of the seven shapes on entry, that
 * shape is recorded without timing (20 still samples q for its screen).
 * Generic-path plants always use shape 1. */
int csd_tune_check(const csd_plant *plant, const float *q, int64_t n,
                   int64_t stride_n, int64_t stride_j, uint8_t *out, void *stream,
                   int32_t *split);

/* Replaces executeCollisionFreeVoxelsKernel (cuda_collision_checker.cu:255-307)
 * / checkCollisionFreeVoxelsCuda (:377-435). voxels (device) [V][3] centres,
 * robot_geoms (HOST array, at most 128 entries) [R] geometry indices checked
 * against the voxels; it is passed to the kernel by value. */
int csd_check_collision_free_voxels(const csd_plant *plant, const float *q,
                                    int64_t n, int64_t stride_n,
                                    int64_t stride_j, const float *voxels,
                                    int32_t num_voxels, float voxel_radius,
                                    const int32_t *robot_geoms,
                                    int32_t num_robot_geoms, uint8_t *out,
                                    void *stream);

/* Replaces checkCollisionFreeVoxelsWithoutSelfCollisionCuda
 * (cuda_collision_checker.cu:310-375): out (device) [N][V] uint8, 1 iff no
 * robot geometry touches voxel v at config n. */
int csd_voxel_flags(const csd_plant *plant, const float *q, int64_t n,
                    int64_t stride_n, int64_t stride_j, const float *voxels,
                    int32_t num_voxels, float voxel_radius,
                    const int32_t *robot_geoms, int32_t num_robot_geoms,
                    uint8_t *out, void *stream);

/* ---- hit-and-run sampling ----
 * Replaces executeUniformSampleInHPolyhedronKernel
 * (geometry/cuda_hit_and_run_sampling.h:25-29). Polytopes (device): A
 * [num_polytopes][face_stride][dim] row-major, b [num_polytopes][face_stride],
 * num_faces [num_polytopes] (device int32). starts (device)
 * [num_polytopes][dim]. out (device) [num_polytopes][num_samples][dim].
 * Randomness: Philox4x32-10 keyed by seed; chain i of polytope p uses
 * counters (i, step, p, round*16 + k). */
int csd_hit_and_run(const float *A, const float *b, const int32_t *num_faces,
                    int32_t face_stride, int32_t dim, int32_t num_polytopes,
                    const float *starts, int64_t num_samples,
                    int32_t mixing_steps, uint64_t seed, uint32_t round,
                    float *out, void *stream);

/* The EI-ZO round's sampler on its own (no counterpart in the reference,
 * whose chains all start at line_start, cuda_edge_inflation_zero_order.cu
 * :122-131): as csd_hit_and_run, but chain i of polytope p starts at
 * fraction (i + 1/2) / num_samples of the segment segment_start[p] ->
 * segment_end[p] (device [num_polytopes][dim] each) and draws counters
 * (i, step, p + stream_offset, round*16 + k). With p the region's index in
 * a csd_eizo_run batch, stream_offset 0, round k and num_samples the
 * round's sample count max(M_k, budget), it yields that round's samples bit
 * for bit. */
int csd_hit_and_run_segments(const float *A, const float *b, const int32_t *num_faces,
                             int32_t face_stride, int32_t dim, int32_t num_polytopes,
                             const float *segment_start, const float *segment_end,
                             int64_t num_samples, int32_t mixing_steps, uint64_t seed,
                             uint32_t stream_offset, uint32_t round, float *out, void *stream);

/* ---- EI-ZO ----
 * Options: field-for-field EizoOptions
 * (planning/cuda_edge_inflation_zero_order.h:18-32). */
typedef struct csd_eizo_options {
  float configuration_margin;
  float tau;
  float delta;
  float epsilon;
  uint32_t bisection_steps;
  uint32_t num_particles;
  uint32_t max_hyperplanes_per_iteration;
  uint32_t max_iterations;
  uint32_t mixing_steps;
  uint64_t seed;
} csd_eizo_options;

/* Per-region result summary (EizoInfo, cuda_edge_inflation_zero_order.h:37-54
 * aggregated per region). */
typedef struct csd_eizo_region_info {
  int32_t iterations;
  int32_t converged;        /* 1: the unadaptive test certified the region */
  int32_t num_faces;
  int32_t faces_added;
  int32_t capacity_reached; /* 1: stopped because face_capacity was full */
  int32_t pad;
  int64_t samples_drawn;
  int64_t collisions_found;
} csd_eizo_region_info;

typedef struct csd_eizo csd_eizo; /* persistent device workspace */

/* Replaces EizoGpuMemory / CudaEdgeInflator's constructor
 * (cuda_edge_inflation_zero_order.cu:448-494, cuda_polytope_builder.cu:69-108):
 * allocates once for up to max_regions concurrent regions of at most
 * face_capacity faces. robot_geoms is a host array. */
int csd_eizo_create(const csd_plant *plant, const int32_t *robot_geoms,
                    int32_t num_robot_geoms, const csd_eizo_options *options,
                    int32_t max_regions, int32_t face_capacity,
                    int32_t max_particles, int32_t max_voxels,
                    csd_eizo **out);
int csd_eizo_destroy(csd_eizo *ws);

/* Inflates num_regions edges at once (executeEizo,
 * cuda_edge_inflation_zero_order.cu:94-339, batched). All arrays are host
 * arrays; the call is synchronous.
 *  domain_A [domain_faces][dim], domain_b [domain_faces]
 *  line_start / line_end [num_regions][dim]
 *  particle_budgets [num_regions] (NULL = options.num_particles each)
 *  voxels [num_voxels][3]
 * Outputs: out_A [num_regions][face_capacity][dim], out_b
 * [num_regions][face_capacity], infos [num_regions]. A region whose faces
 * fill face_capacity stops there (infos[r].capacity_reached = 1); the other
 * regions continue. */
int csd_eizo_run(csd_eizo *ws, const float *domain_A, const float *domain_b,
                 int32_t domain_faces, const float *line_start,
                 const float *line_end, int32_t num_regions,
                 const uint32_t *particle_budgets, const float *voxels,
                 int32_t num_voxels, float voxel_radius, float *out_A,
                 float *out_b, csd_eizo_region_info *infos);

/* One round of one region in the last csd_eizo_run: the fields of
 * EizoIterationInfo (planning/cuda_edge_inflation_zero_order.h:37-47,
 * filled at cuda_edge_inflation_zero_order.cu:181-195,309-336) plus what the
 * batched and guided rounds add. */
typedef struct csd_eizo_round_info {
  int32_t iteration_index;
  int32_t num_samples_drawn;         /* samples of the unadaptive test (0: skipped) */
  float collisions_threshold;        /* (1 - tau) epsilon M_k */
  int32_t num_collisions_unadaptive; /* among the first M_k samples (-1: skipped) */
  int32_t num_collisions_found;      /* among all samples of the test */
  int32_t num_collisions_updated;    /* particles bisected */
  int32_t num_faces_placed;
  int32_t num_faces_at_start;        /* faces of the region when the round began */
  int32_t pilot_samples;             /* csd_eizo_set_guidance pilot */
  int32_t pilot_collisions;          /* found by the exact check */
  float iteration_time_ms;           /* wall time of the round, shared by the batch */
  int32_t pilot_checked;             /* pilot samples checked exactly (all unless pruned) */
} csd_eizo_round_info;

/* Per-round record of `region` in the last csd_eizo_run (replaces the
 * reference's EizoInfo::iteration_info). *count receives the region's
 * rounds; out (host, capacity entries, may be NULL to only count) receives
 * them. */
int csd_eizo_get_rounds(const csd_eizo *ws, int32_t region, csd_eizo_round_info *out,
                        int32_t capacity, int32_t *count);

/* The particles of `region` in its last round of the last csd_eizo_run, as
 * EizoGpuMemory holds them for C++ callers after executeEizo
 * (cuda_edge_inflation_zero_order.h:59-161): the colliding samples taken
 * (particles), their bisected boundary points (scol) and their projections
 * onto the segment (proj), each [count][dim] on the host (any may be NULL;
 * all NULL only counts). count is 0 when the region's last round certified
 * it. */
int csd_eizo_last_particles(const csd_eizo *ws, int32_t region, float *particles, float *scol,
                            float *proj, int32_t capacity, int32_t *count);

struct csd_mlp;

/* Guidance for later csd_eizo_run calls (no counterpart in the reference;
 * all off by default, which is the reference's round structure):
 *  pilot_samples > 0: each round first checks pilot_samples samples of an
 *    independent random stream; a region whose pilot collision fraction
 *    exceeds pilot_factor * epsilon skips that round's unadaptive test and
 *    takes its particles from the pilot. The test, when run, draws its own
 *    samples, so certification keeps the reference's guarantee.
 *  surrogate != NULL: particles are ranked by the surrogate's predicted
 *    boundary distance (signed distance at line_points points between each
 *    particle's projection and the particle) and only the keep_particles (1..1024) best
 *    of the first 1024 are bisected; the exact checker still judges every
 *    bisection step and every sample. */
/* The workspace holds the surrogate until the guidance is changed or the
 * workspace destroyed; csd_mlp_destroy refuses a model held this way. */
int csd_eizo_set_guidance(csd_eizo *ws, uint32_t pilot_samples, float pilot_factor,
                          const struct csd_mlp *surrogate, int32_t keep_particles,
                          int32_t line_points);

/* Surrogate-pruned pilot for later csd_eizo_run calls (no counterpart in the
 * reference). With enable != 0, and while guidance sets both a pilot and a
 * surrogate, each round's pilot samples are first predicted by the
 * surrogate and only those whose prediction is below predicted_below (in
 * the surrogate's output units) get the exact collision check; the others
 * count as free. The pilot then decides from the collisions the exact check
 * confirmed, and only those become particles: a wrong prediction can make a
 * round run its unadaptive test needlessly or draw fewer particles, never
 * certify a region (the test always checks all its samples exactly).
 * enable = 0 (the default) checks every pilot sample. */
int csd_eizo_set_pruning(csd_eizo *ws, int32_t enable, float predicted_below);

/* Ellipsoid metric of the face placement for later csd_eizo_run calls (no
 * counterpart in the reference, whose faces are placed in the Euclidean
 * metric, cuda_set_builder_utils.cu:359-424). metric (host, dim x dim
 * row-major, symmetric positive definite; NULL restores the Euclidean
 * metric, the default): each round takes its particles in ascending
 * |scol - proj|_M = sqrt(e' M e), e = scol - proj, and places the face with
 * the unit normal M e / |M e|, the hyperplane tangent at scol to the
 * M-ellipsoid centred at proj; b = a.scol - margin, relaxed to keep the
 * segment inside as before. In FP64 on the device. CSD_ERR_INVALID_ARGUMENT
 * if the matrix is not symmetric positive definite. */
int csd_eizo_set_metric(csd_eizo *ws, const float *metric);

/* Sampler tile of later csd_eizo_run calls (no counterpart in the
 * reference, whose uniformSampleInHPolyhedronKernel re-reads the polytope
 * from global memory in every thread, cuda_hit_and_run_sampling.cu:17-112).
 * One lane runs one hit-and-run chain; a workgroup stages its region's
 * polytope in LDS once. 256-thread workgroups stage up to 64 KB, 1024-thread
 * workgroups (16 waves sharing one copy) up to 160 KB; larger polytopes are
 * read from global memory. The samples are the same bit for bit whatever
 * the tile. csd_eizo_set_sampler_tile: 1024-thread workgroups from
 * wide_from faces on when both tiles fit (0: always; the default,
 * INT32_MAX: only when the polytope exceeds 64 KB).
 * csd_eizo_tune_sampler: times both tiles on synthetic polytopes of this
 * workspace's dimension at several face counts and sets wide_from to the
 * smallest count from which the wide tile is faster at every larger count
 * tested (returned in *wide_from, may be NULL). Synchronous; it overwrites
 * the workspace's region state, so call it before csd_eizo_run. */
int csd_eizo_set_sampler_tile(csd_eizo *ws, int32_t wide_from);
int csd_eizo_tune_sampler(csd_eizo *ws, int32_t *wide_from);

/* Device-resident rounds for later csd_eizo_run calls (no counterpart in the
 * reference, whose executeEizo launches every stage from the host and reads
 * counters back between them, cuda_edge_inflation_zero_order.cu:109-306).
 * rounds_per_sync 1..64: a kernel builds each round's region list on the
 * device, one round is captured as a hipGraph and replayed, and the host
 * reads the region status every rounds_per_sync rounds. Guided rounds
 * (csd_eizo_set_guidance, csd_eizo_set_pruning) are device-resident too: the
 * pilot's decision picks the test phase's regions on the device, and the
 * surrogate's predictions and ranking run over the device-built slot list.
 * Results are those of host-driven rounds bit for bit. 0 (the default)
 * restores host-driven rounds. */
int csd_eizo_set_graph(csd_eizo *ws, int32_t rounds_per_sync);

/* EI-ZO stage kernels exposed for stage-wise parity tests (device arrays).
 * Replaces executeProjectSamplesOntoLineSegmentsKernel and the bisection
 * kernels (planning/cuda_set_builder_utils.h:25-131), fused: per particle,
 * project onto its segment then run `steps` bisection steps with the full
 * FK + collision check inline. */
int csd_bisect(const csd_plant *plant, const float *particles,
               const float *line_start, const float *line_end,
               const uint32_t *segment_idx, int32_t num_particles,
               int32_t steps, const float *voxels, int32_t num_voxels,
               float voxel_radius, const int32_t *robot_geoms,
               int32_t num_robot_geoms, float *scol, float *proj,
               void *stream);

/* Greedy FP64 face placement (maybePlaceFace +
 * checkParticlesMadeRedundantByLastFace, cuda_set_builder_utils.cu:359-455,
 * loop of cuda_edge_inflation_zero_order.cu:268-306) for one region, one
 * workgroup. A/b/num_faces device; returns faces added in *added (device). */
int csd_place_faces(float *A, float *b, int32_t *num_faces, int32_t capacity,
                    int32_t dim, const float *scol, const float *proj,
                    int32_t num_particles, const float *line_start,
                    const float *line_end, float margin, int32_t max_faces,
                    int32_t *added, void *stream);

/* csd_place_faces without allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It places
 * the same faces by the same kernels, but keeps its scratch (the per-particle
 * FP64 distances and four counters) in a fixed device array of the library,
 * one per GPU, allocated with the library (520 KiB); so num_particles is at
 * most 65536 (CSD_ERR_CAPACITY above). Fully asynchronous on `stream`,
 * nothing is read back: *added (device) receives the faces added and
 * *status (device, may be NULL) 0, or 2 when the face capacity ran out
 * (csd_place_faces returns CSD_ERR_CAPACITY then). Every call on a GPU uses
 * that GPU's one scratch array: calls that overlap on one GPU (on different
 * streams or threads) race and must be serialised by the caller. */
int csd_place_faces_noalloc(float *A, float *b, int32_t *num_faces, int32_t capacity,
                            int32_t dim, const float *scol, const float *proj,
                            int32_t num_particles, const float *line_start,
                            const float *line_end, float margin, int32_t max_faces,
                            int32_t *added, int32_t *status, void *stream);

/* ---- region editing ----
 * Replaces EditRegionsCuda (planning/cuda_edit_regions.cu:44-300; declared
 * planning/cuda_edit_regions.h:58-67) below its host-side argument checks
 * and its grouping of the collisions by segment. All arrays are device
 * arrays:
 *   collisions [C][dim], segment_idx [C]: the colliding configurations,
 *     grouped by the segment / region they belong to: those of region r are
 *     [region_offsets[r], region_offsets[r+1]) and carry segment_idx r (the
 *     bisection reads segment_idx, the face placement the offsets; their
 *     agreement is not checked on the device);
 *   region_offsets [R+1]: non-decreasing, region_offsets[R] <= C;
 *   line_start, line_end [R][dim]: segment r lies in region r (the caller
 *     checks that both ends are inside, as the reference does);
 *   A [R][capacity][dim], b [R][capacity], num_faces [R]: the regions,
 *     edited in place;
 *   scol, proj [C][dim] (out): bisected collisions and their projections;
 *   status [R] (out): faces added, or -1 more than 4096 collisions for the
 *     region (region untouched), -2 face capacity reached (faces placed
 *     until then kept), -3 num_faces outside [0, capacity] (untouched),
 *     -4 offsets outside [0, C] or decreasing (untouched).
 * Per collision: projection onto its segment and bisection_steps bisection
 * steps with the full check (plant pairs and, with voxels, robot geometries
 * against voxel spheres). Per region, in FP64 on the stored FP32 faces: its
 * collisions whose bisected point is strictly inside the region, in ascending
 * |scol - proj| (lower index on ties), each add the face a = (scol -
 * proj)/|scol - proj|, b = a.scol - margin, raised by max(a.ls - b, a.le - b)
 * + 2e-6 when that is positive, if the point is still strictly inside the
 * region edited so far. b is stored rounded up. robot_geoms is a host
 * array. Asynchronous on stream. */
int csd_edit_regions(const csd_plant *plant, const float *collisions,
                     const uint32_t *segment_idx, const int32_t *region_offsets,
                     int32_t num_collisions, const float *line_start,
                     const float *line_end, int32_t num_regions, float *A,
                     float *b, int32_t *num_faces, int32_t capacity,
                     const float *voxels, int32_t num_voxels, float voxel_radius,
                     const int32_t *robot_geoms, int32_t num_robot_geoms,
                     float margin, int32_t bisection_steps, float *scol,
                     float *proj, int32_t *status, void *stream);

/* Offline voxel collision maps. Replaces BuildCollisionMap and
 * BuildEdgeCollisionMap (planning/roadmap_builder.cpp:326-679), which check
 * every (configuration, voxel) pair with
 * checkCollisionFreeVoxelsWithoutSelfCollisionCuda in chunks.
 * Grid of RoadmapOptions (planning/roadmap.h:55-75): nx = ceil(size_x /
 * (2 res)) etc., voxel id ix + iy nx + iz nx ny, centre as
 * GetCollisionVoxelCenter (roadmap_builder.cpp:709-741), voxel spheres of
 * radius sqrt(3) res. Items are the nodes (edges == NULL) or the edges
 * ([E][2] node indices; an edge's configurations are t q2 + (1 - t) q1,
 * t = (k + 1) / n, k < n = max(0, floor(|q2 - q1| / step) - 2), the
 * roadmap rule of roadmap_builder.cpp:526-541).
 * Output, per voxel, the items at which some robot geometry touches the
 * voxel sphere: offsets (device [V+1] int64, always written) and items
 * (device [total] int32, item indices in unspecified order within a voxel;
 * may be NULL to only count). *total (host) receives the number of entries;
 * CSD_ERR_CAPACITY if items is given with capacity < *total (offsets and
 * *total are still valid). nodes (device [N][dim]); robot_geoms, map_center,
 * map_size are host arrays. Grids of at most 524,288 voxels. Synchronous. */
int csd_voxel_collision_map(const csd_plant *plant, const float *nodes,
                            int32_t num_nodes, const int32_t *edges,
                            int64_t num_edges, float step,
                            const int32_t *robot_geoms, int32_t num_robot_geoms,
                            const float *map_center, const float *map_size,
                            float resolution, int64_t *offsets, int32_t *items,
                            int64_t capacity, int64_t *total, void *stream);

/* ---- distance field ----
 * Replaces DistanceField (geometry/cuda_distance_field.h:133-172): exact
 * Euclidean distance from each voxel of a G^3 grid (voxel i at origin +
 * i * voxel_size) to the nearest occupied voxel. A point maps to voxel
 * round((p - origin) / voxel_size) per axis; points outside the grid are
 * ignored by set_occupancy, and get_distance returns 0 for them (as the
 * reference does) and +inf inside a field with no occupied voxel.
 * grid_size 1..1024 (the reference requires a multiple of 32). Memory: G^3
 * bytes plus three G^3 int32 grids. Points are device arrays [M][3]. */
typedef struct csd_distance_field csd_distance_field;
int csd_distance_field_create(const float *origin /* host [3] */, int32_t grid_size,
                              float voxel_size, csd_distance_field **out);
int csd_distance_field_destroy(csd_distance_field *field);
/* resetOccupancy / setOccupancy (cuda_distance_field.cu:315-330) */
int csd_distance_field_reset_occupancy(csd_distance_field *field, void *stream);
int csd_distance_field_set_occupancy(csd_distance_field *field,
                                     const float *points, int64_t num_points,
                                     void *stream);
/* computeEDT (cuda_distance_field.cu:304-313) */
int csd_distance_field_compute(csd_distance_field *field, void *stream);
/* getDistance (cuda_distance_field.cu:332-349): out (device) [M] */
int csd_distance_field_get_distance(const csd_distance_field *field,
                                    const float *points, int64_t num_points,
                                    float *out, void *stream);
/* Device pointer to the squared distances in voxel units, [z][y][x] int32
 * (0x3fffffff where no voxel is occupied); valid until the next compute. */
int csd_distance_field_squared(const csd_distance_field *field, const int32_t **d2);

/* ---- edge validation and occupancy ----
 * Configurations along each edge are generated in-kernel, never stored.
 * rule 0 -- checkEdgeCollisionFree (plant/collision_checker.cpp:41-53):
 *   n = ceil(|q2-q1| / step), configurations q1 + (k/n)(q2-q1), k < n
 *   (q2 itself is not checked).
 * rule 1 -- VisibilityGraph (planning/cuda_visibility_graph.cu:58-172):
 *   n = (uint)(|q2-q1| / step + 2) points: q1, q2 and t = (k+1)/(n-1),
 *   k < n-2.
 * rule 2 -- RoadmapBuilder::BuildRoadmap (planning/roadmap_builder.cpp
 *   :262-280): n = max(0, floor(|q2-q1| / step) - 2) points t = (k+1)/n,
 *   k < n (q2 is checked, q1 is not; no points: free).
 * nodes (device) [num_nodes][dim], edges (device) [E][2] node indices,
 * out (device) [E] uint8 (1 = collision free). */
int csd_validate_edges(const csd_plant *plant, const float *nodes,
                       int32_t num_nodes, const int32_t *edges,
                       int64_t num_edges, float step, int32_t rule,
                       uint8_t *out, void *stream);

/* Swept occupancy on a regular voxel grid (voxel centre / id convention of
 * planning/roadmap_builder.cpp:681-741, voxel radius sqrt(3)*resolution):
 * grid (device) [nz][ny][nx] uint8 is OR-ed with 1 for every voxel touched
 * by a robot geometry at any of the configurations. */
int csd_occupancy(const csd_plant *plant, const float *q, int64_t n,
                  int64_t stride_n, int64_t stride_j, const int32_t *robot_geoms,
                  int32_t num_robot_geoms, const float *map_center,
                  const float *map_size, float resolution, uint8_t *grid,
                  void *stream);

/* Brute-force k nearest neighbours (squared L2 in joint space); replaces
 * ComputeJointDistancesKernel + host sort (planning/drm_cuda_utils.cu:6-39,
 * planning/roadmap_builder.cpp:160-200). nodes (device) [N][dim];
 * out_idx (device) [N][k] sorted by distance then index. */
int csd_knn(const float *nodes, int32_t num_nodes, int32_t dim, int32_t k,
            int32_t *out_idx, float *out_d2, void *stream);

/* Paged k nearest neighbours for the nodes queries[0..Q) (device [Q]):
 * row i of out_idx / out_d2 (device [Q][k]) lists, in (squared distance,
 * index) order, the k nearest neighbours of node queries[i] that come after
 * (after_d2[i], after_idx[i]) in that order (both NULL: from the start), so
 * repeated calls walk the full sorted list of ComputeJointDistancesKernel +
 * sort (roadmap_builder.cpp:176-203) page by page. -1 / +inf pad short
 * pages. */
int csd_knn_page(const float *nodes, int32_t num_nodes, int32_t dim,
                 const int32_t *queries, int32_t num_queries, int32_t k,
                 const float *after_d2, const int32_t *after_idx,
                 int32_t *out_idx, float *out_d2, void *stream);

/* ---- signed-distance surrogate ----
 * No counterpart in the reference (its checkers return booleans only). An MLP
 *   x = 2 (q - lower) / (upper - lower) - 1   (in_dim <= 32, zero-padded)
 *   a_l = bf16(SiLU(W_l a_{l-1} + b_l)), l = 1..4, width 256
 *   y = w_5 . a_4 + b_5
 * with BF16 weights/activations and FP32 accumulation on the MFMA units, FP32
 * master weights and Adam (beta 0.9 / 0.999, eps 1e-8). Parameters are one
 * FP32 vector: W1 [256][32], b1 [256], W2 [256][256], b2, W3, b3, W4, b4,
 * w5 [256], b5. EI-ZO uses it to steer sampling; the exact checker still
 * judges every sample. */
typedef struct csd_mlp csd_mlp;
int csd_mlp_create(int32_t in_dim, const float *lower, const float *upper,
                   uint64_t seed, csd_mlp **out);  /* lower/upper: host [in_dim] */
/* CSD_ERR_INVALID_ARGUMENT (model kept) while an EI-ZO workspace holds the
 * model as its surrogate; waits for the device before freeing. */
int csd_mlp_destroy(csd_mlp *model);
int csd_mlp_num_params(const csd_mlp *model, int64_t *count);
int csd_mlp_get_params(const csd_mlp *model, float *host_params);
int csd_mlp_set_params(csd_mlp *model, const float *host_params);
/* y for n configurations (device q addressed with the two strides). */
int csd_mlp_predict(const csd_mlp *model, const float *q, int64_t n,
                    int64_t stride_n, int64_t stride_j, float *out,
                    void *stream);
/* Mean squared error over one batch (batch a multiple of 128): q (device)
 * [batch][in_dim], target (device) [batch]. _gradients writes dLoss/dparams
 * (device, num_params floats, may be NULL) and the loss (device float, may be
 * NULL) without updating; _train_step also applies one Adam step. */
int csd_mlp_gradients(csd_mlp *model, const float *q, const float *target,
                      int32_t batch, float *grads, float *loss, void *stream);
int csd_mlp_train_step(csd_mlp *model, const float *q, const float *target,
                       int32_t batch, float learning_rate, float *loss,
                       void *stream);
/* Labels: the smallest signed distance (metres, FP32, negative in
 * penetration) over the plant's collision pairs, for plants whose moving
 * collision geometry is all spheres (CSD_ERR_UNSUPPORTED otherwise);
 * voxels are not included. out (device) [n]. */
int csd_signed_distance(const csd_plant *plant, const float *q, int64_t n,
                        int64_t stride_n, int64_t stride_j, float *out,
                        void *stream);

#ifdef __cplusplus
}
#endif
#endif
