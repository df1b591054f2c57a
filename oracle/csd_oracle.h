/*
 * csd_oracle.h -- CPU restatement of csdecomp's hot path, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Nothing in the product (csdecomp_amd/) links, loads or calls this code.
 *
 * Two flavours of every check:
 *   - an FP32 "mirror" that follows the reference op order and its
 *     comparison operators (variant ORC_RULE_CPU = collision_checker.cpp,
 *     variant ORC_RULE_GPU = cuda_geometry_utilities.h), and
 *   - an FP64 oracle that returns signed distances, used to define the
 *     1e-5 m contact band in which flags are allowed to differ.
 *
 * Plant layout mirrors MinimalPlant (plant/minimal_plant.h:16-39):
 * 4x4 matrices are column-major (Eigen default), pairs are (A,B) int32.
 */
#ifndef CSD_ORACLE_H
#define CSD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FIXED = 0, ORC_REVOLUTE = 1, ORC_PRISMATIC = 2 };
enum { ORC_BOX = 0, ORC_SPHERE = 1, ORC_CYLINDER = 2, ORC_CAPSULE = 3 };
enum { ORC_RULE_CPU = 0, ORC_RULE_GPU = 1 };

typedef struct {
  int32_t num_links, num_joints, num_config_vars;
  const int32_t *joint_type, *joint_parent_link, *joint_child_link;
  const int32_t *joint_config_idx, *traversal;
  const float *joint_X_PL_J; /* [J][16] column-major */
  const float *joint_axis;   /* [J][3] */
  int32_t num_geoms;
  const int32_t *geom_type, *geom_link, *geom_group;
  const float *geom_dims;  /* [G][3]: box = full edge lengths, sphere = (r,-1,-1),
                              capsule = (r, length, r) */
  const float *geom_X_L_B; /* [G][16] column-major */
  int32_t num_pairs;
  const int32_t *pairs; /* [P][2] */
} orc_plant;

/* ---- forward kinematics (minimal_kinematic_tree.cpp:10-58) ---- */
void orc_fk_f32(const orc_plant *p, const float *q, float *X /*[L][16]*/);
void orc_fk_f64(const orc_plant *p, const double *q, double *X /*[L][16]*/);

/* ---- pair primitives ---- */
int orc_pair_free_f32(const orc_plant *p, int gA, const float *X_W_LA, int gB,
                      const float *X_W_LB, int rule);
double orc_pair_sd_f64(const orc_plant *p, int gA, const double *X_W_LA, int gB,
                       const double *X_W_LB);
int orc_geom_voxel_free_f32(const orc_plant *p, int g, const float *X_W_L,
                            const float *vox, float voxel_radius, int rule);
double orc_geom_voxel_sd_f64(const orc_plant *p, int g, const double *X_W_L,
                             const double *vox, double voxel_radius);

/* ---- batched drivers (configs AoS [N][dof], the (dof,N) column-major
 * layout the reference's Eigen::MatrixXf uses). nthreads<=0 -> all cores. */
void orc_check_batch(const orc_plant *p, const float *q, int64_t n, int rule,
                     uint8_t *out, int nthreads);
void orc_check_voxels_batch(const orc_plant *p, const float *q, int64_t n,
                            const float *voxels, int32_t num_voxels,
                            float voxel_radius, const int32_t *robot_geoms,
                            int32_t num_robot_geoms, int rule, uint8_t *out,
                            int nthreads);
void orc_voxel_flags_batch(const orc_plant *p, const float *q, int64_t n,
                           const float *voxels, int32_t num_voxels,
                           float voxel_radius, const int32_t *robot_geoms,
                           int32_t num_robot_geoms, int rule,
                           uint8_t *out /*[N][V]*/, int nthreads);
/* FP64 minimum signed distance over all pairs (+ robot geoms x voxels). */
void orc_min_sd_batch(const orc_plant *p, const float *q, int64_t n,
                      const float *voxels, int32_t num_voxels,
                      float voxel_radius, const int32_t *robot_geoms,
                      int32_t num_robot_geoms, double *out, int nthreads);
/* per-link FK error |X_f32 - X_f64| (max abs over the 3x4 block), [N][L] */
void orc_fk_error_batch(const orc_plant *p, const float *q, int64_t n,
                        double *out, int nthreads);

/* checkEdgeCollisionFree (collision_checker.cpp:41-53) */
int orc_check_edge(const orc_plant *p, const float *q1, const float *q2,
                   float step, int rule);
/* number of configurations the rule above checks for an edge */
int32_t orc_edge_num_steps(const float *q1, const float *q2, int32_t dof,
                           float step);

/* ---- counter-based RNG (Philox4x32-10, Salmon et al. SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

/* ---- H-polytope + hit-and-run (cuda_hit_and_run_sampling.cu:74-112) ----
 * A row-major [F][dim] (hpolyhedron.h:15-21), b [F].
 * chain i of a call draws its randomness from counters
 * (chain_offset + i, step, stream, round*16 + call). */
void orc_hit_and_run(const float *A, const float *b, int32_t num_faces,
                     int32_t dim, const float *start, int64_t num_chains,
                     int64_t chain_offset, int32_t mixing_steps, uint64_t seed,
                     uint32_t stream, uint32_t round, float *out /*[n][dim]*/);

/* chain c = chain_offset + i of `total` starts at fraction (c + 1/2) / total
 * of the segment ls -> le, with counters (c, step, stream, round*16 + call)
 * (the batched EI-ZO round's sampler; the reference starts every chain at
 * line_start, cuda_edge_inflation_zero_order.cu:122-131) */
void orc_hit_and_run_segment(const float *A, const float *b, int32_t num_faces,
                             int32_t dim, const float *ls, const float *le,
                             int64_t num_chains, int64_t chain_offset,
                             int64_t total, int32_t mixing_steps, uint64_t seed,
                             uint32_t stream, uint32_t round,
                             float *out /*[n][dim]*/);

/* ---- EI-ZO building blocks (cuda_set_builder_utils.*) ---- */
uint32_t orc_num_samples_unadaptive(float epsilon, float delta, float tau);
float orc_delta_k(float delta, int32_t iteration);
void orc_project_onto_segment(const float *p, const float *q1, const float *q2,
                              int32_t dim, float *proj, float *dist);
/* fused bisection of one particle; returns the optimised collision point */
void orc_bisect(const orc_plant *p, const float *particle, const float *proj,
                int32_t dim, int32_t steps, const float *voxels,
                int32_t num_voxels, float voxel_radius,
                const int32_t *robot_geoms, int32_t num_robot_geoms, int rule,
                float *scol);
/* greedy face placement over u particles (cuda_edge_inflation_zero_order.cu
 * :268-306, FP64 placement). Appends faces to A/b (capacity cap). Returns
 * number of faces added, or -1 on capacity overflow. */
int32_t orc_place_faces(float *A, float *b, int32_t *num_faces, int32_t cap,
                        int32_t dim, const float *scol, const float *proj,
                        int32_t u, const float *ls, const float *le,
                        float margin, int32_t max_faces);

/* the same with an ellipsoid metric M (SPD, dim x dim row-major, NULL =
 * identity): particles in ascending sqrt(e' M e), e = scol - proj, face
 * normal M e / |M e| */
int32_t orc_place_faces_metric(float *A, float *b, int32_t *num_faces,
                               int32_t cap, int32_t dim, const float *scol,
                               const float *proj, int32_t u, const float *ls,
                               const float *le, float margin,
                               int32_t max_faces, const double *M);

/* region editing for one region (cuda_edit_regions.cu:236-293): u bisected
 * collisions scol with projections proj; appends faces to A/b (capacity
 * cap). Returns faces added, or -1 on capacity overflow. */
int32_t orc_edit_region(float *A, float *b, int32_t *num_faces, int32_t cap,
                        int32_t dim, const float *scol, const float *proj,
                        int32_t u, const float *ls, const float *le,
                        float margin);

/* distance field (geometry/cuda_distance_field.*): voxelization of points
 * [m][3] into a G^3 uint8 grid ([z][y][x]) and the squared distance in
 * voxels to the nearest occupied voxel by definition (0x3fffffff if none). */
void orc_voxelize(const float *points, int64_t m, const float *origin,
                  float voxel, int32_t G, uint8_t *occ);
void orc_edt_brute(const uint8_t *occ, int32_t G, int32_t *d2, int threads);

typedef struct {
  float configuration_margin, tau, delta, epsilon;
  uint32_t bisection_steps, num_particles, max_hyperplanes_per_iteration;
  uint32_t max_iterations, mixing_steps;
  uint64_t seed;
} orc_eizo_options;

typedef struct {
  int32_t iterations;
  int32_t converged;
  int64_t samples_drawn;
  int32_t faces_added;
} orc_eizo_info;

/* full EI-ZO for one edge; A/b hold the domain on entry (num_faces rows)
 * and the region on exit. Returns 0 ok, -1 capacity overflow.
 * spread_starts = 0: every chain starts at ls (the reference);
 * 1: the batched round's spread starts (orc_hit_and_run_segment). */
int32_t orc_eizo(const orc_plant *p, const int32_t *robot_geoms,
                 int32_t num_robot_geoms, const float *voxels,
                 int32_t num_voxels, float voxel_radius, float *A, float *b,
                 int32_t *num_faces, int32_t cap, const float *ls,
                 const float *le, const orc_eizo_options *opt, uint32_t stream,
                 int rule, int spread_starts, orc_eizo_info *info);

/* orc_eizo over R regions, one per OpenMP thread (ls/le rows [R][dim],
 * outputs [R][cap][dim], [R][cap], [R]; region r uses stream r). */
void orc_eizo_batch(const orc_plant *p, const int32_t *robot_geoms,
                    int32_t num_robot_geoms, const float *A_dom, const float *b_dom,
                    int32_t F, int32_t cap, int32_t R, const float *ls, const float *le,
                    const orc_eizo_options *opt, int rule, int spread_starts,
                    float *A_out, float *b_out, int32_t *nf_out, int32_t *status,
                    orc_eizo_info *infos, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
