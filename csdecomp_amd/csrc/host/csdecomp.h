// csdecomp.h -- C++ host API mirroring csdecomp's plant, collision-checker,
// sampler and EI-ZO entry points (same class / function names and argument
// meaning), Eigen-free. The batched names (the *Cuda functions, EI-ZO, the
// sampler, the visibility graph) go through the C-ABI of
// include/csdecomp_amd.h to the gfx950 kernels. The reference's CPU names for
// one configuration (checkCollisionFree, checkCollisionFreeVoxels,
// checkEdgeCollisionFree, pairCollisionFree, geomToVoxelCollisionFree,
// computeLinkFrameToWorldTransformsMinimal) are CPU functions, as in the
// reference (cpu_checker.cpp).
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../../include/csdecomp_amd.h"

namespace csdecomp {

using Matrix4f = std::array<float, 16>;  // column-major, like Eigen::Matrix4f
using Vector3f = std::array<float, 3>;
using GeometryIndex = int32_t;
Matrix4f identity4();

// Dense float matrix, column-major (the memory layout of Eigen::MatrixXf).
struct MatrixXf {
  int rows = 0, cols = 0;
  std::vector<float> data;
  MatrixXf() = default;
  MatrixXf(int r, int c) : rows(r), cols(c), data((size_t)r * c, 0.f) {}
  float &operator()(int r, int c) { return data[(size_t)c * rows + r]; }
  float operator()(int r, int c) const { return data[(size_t)c * rows + r]; }
};

// Throws std::runtime_error carrying the library's message if rc != CSD_OK.
void check_status(int rc, const char *what);
// Selects device 0 once per process; throws if no gfx950 device is visible.
void ensure_device();

// ---------------------------------------------------------------------------
// kinematics (plant/kinematic_tree.h, plant/minimal_kinematic_tree.h)
// ---------------------------------------------------------------------------
enum JointType { FIXED = 0, REVOLUTE = 1, PRISMATIC = 2 };

struct Joint {
  std::string name;
  JointType type = FIXED;
  Matrix4f X_PL_J = identity4();
  int parent_link = -1;
  int child_link = -1;
  Vector3f axis{0.f, 0.f, 0.f};
  float position_lower_limit = 0.f;
  float position_upper_limit = 0.f;
};

struct Link {
  std::string name;
  int parent_joint = -1;
  std::vector<int> child_joints;
};

struct MinimalJoint {
  JointType type = FIXED;
  Matrix4f X_PL_J = identity4();
  int index = 0;
  int parent_link = -1;
  int child_link = -1;
  Vector3f axis{0.f, 0.f, 0.f};
};

struct MinimalLink {
  int index = 0;
  int parent_joint_index = -1;
  std::vector<int> child_joint_index_array;
  int num_child_joints = 0;
};

struct MinimalKinematicTree {
  std::vector<MinimalJoint> joints;
  std::vector<MinimalLink> links;
  std::vector<int> joint_traversal_sequence;
  std::vector<int> joint_idx_to_config_idx;
  int num_configuration_variables = 0;
  int num_joints = 0;
  int num_links = 0;
};

class KinematicTree {
 public:
  // kinematic_tree.cpp:14-52
  void addLink(const Link &link);
  void addJoint(const Joint &joint);
  // kinematic_tree.cpp:82-146: DFS joint order (stack, children pushed in
  // order), config indices in joint order, limits of non-fixed joints
  void finalize();
  int getLinkIndexByName(const std::string &name) const;
  std::string getLinkNameByIndex(int index) const;
  const Link &getLink(int index) const;
  int getJointIndex(const std::string &name) const;
  const Joint &getJoint(int index) const;
  bool isFinalized() const { return is_finalized_; }
  const std::vector<int> &getLeafLinkIndices() const { return leaf_link_indices_; }
  const std::vector<Link> &getLinks() const { return links_; }
  const std::vector<Joint> &getJoints() const { return joints_; }
  std::vector<float> getPositionLowerLimits() const { return lower_; }
  std::vector<float> getPositionUpperLimits() const { return upper_; }
  // kinematic_tree.cpp:195-252, evaluated on the CPU (cpu_checker.cpp)
  std::vector<Matrix4f> computeLinkFrameToWorldTransforms(const std::vector<float> &q) const;
  MinimalKinematicTree getMyMinimalKinematicTreeStruct() const;
  int numLinks() const { return (int)links_.size(); }
  int numJoints() const { return (int)joints_.size(); }
  int numConfigurationVariables() const { return num_configuration_variables_; }

 private:
  std::vector<Link> links_;
  std::vector<Joint> joints_;
  std::unordered_map<std::string, int> link_name_to_index_, joint_name_to_index_;
  std::vector<int> traversal_, joint_idx_to_config_idx_, leaf_link_indices_;
  std::vector<float> lower_, upper_;
  int num_configuration_variables_ = 0;
  bool is_finalized_ = false;
};

// ---------------------------------------------------------------------------
// collision geometry (plant/collision_geometry.h)
// ---------------------------------------------------------------------------
enum ShapePrimitive { BOX = 0, SPHERE = 1, CYLINDER = 2, CAPSULE = 3 };
enum CollisionGroup { ROBOT = 0, WORLD = 1, VOXELS = 2 };

struct CollisionGeometry {
  ShapePrimitive type = BOX;
  Vector3f dimensions{0.f, 0.f, 0.f};  // box: full lengths; sphere: (r,-1,-1)
  int link_index = 0;
  CollisionGroup group = ROBOT;
  Matrix4f X_L_B = identity4();
};

struct SceneInspector {
  std::vector<std::vector<int>> link_index_to_scene_collision_geometry_indices;
  std::unordered_map<std::string, int> scene_collision_geometry_name_to_geometry_index;
  std::vector<std::pair<int, int>> robot_to_world_collision_pairs;
  std::vector<GeometryIndex> robot_geometry_ids;
  // 2 x P, column-major
  std::vector<int32_t> getCollisionPairMatrix() const;
};

// urdf_parser.cpp:485-616: assigns WORLD/ROBOT groups and the collision pairs
SceneInspector finalizeScene(const KinematicTree &tree, std::vector<CollisionGeometry> &geoms,
                             const std::unordered_map<std::string, int> &names);

// Plant data (plant/minimal_plant.h:16-39) without compile-time capacities.
// The compiled device image is cached and rebuilt whenever the content
// changes.
struct MinimalPlant {
  MinimalKinematicTree kin_tree;
  std::vector<CollisionGeometry> scene_collision_geometries;
  std::vector<GeometryIndex> collision_pairs_flat;  // [2P]
  int num_scene_geometries = 0;
  int num_collision_pairs = 0;
  // device image for the current content (compiled on first use)
  csd_plant *device() const;

 private:
  mutable std::shared_ptr<csd_plant> dev_;
  mutable size_t dev_hash_ = 0;
};

class Plant {
 public:
  Plant(const KinematicTree &kt, const SceneInspector &insp,
        const std::vector<CollisionGeometry> &geoms)
      : kin_tree_(kt), inspector_(insp), geoms_(geoms) {}
  int numLinks() const { return kin_tree_.numLinks(); }
  int numJoints() const { return kin_tree_.numJoints(); }
  std::vector<float> getPositionLowerLimits() const { return kin_tree_.getPositionLowerLimits(); }
  std::vector<float> getPositionUpperLimits() const { return kin_tree_.getPositionUpperLimits(); }
  int numConfigurationVariables() const { return kin_tree_.numConfigurationVariables(); }
  MinimalPlant getMinimalPlant() const;
  const KinematicTree &getKinematicTree() const { return kin_tree_; }
  const SceneInspector &getSceneInspector() const { return inspector_; }
  const std::vector<GeometryIndex> &getRobotGeometryIds() const {
    return inspector_.robot_geometry_ids;
  }
  const std::vector<CollisionGeometry> &getSceneCollisionGeometries() const { return geoms_; }
  std::vector<int32_t> getCollisionPairMatrix() const {
    return inspector_.getCollisionPairMatrix();
  }
  std::vector<std::string> getSceneCollisionGeometryNames() const;
  int getSceneCollisionGeometryIndexByName(const std::string &name) const;

 private:
  KinematicTree kin_tree_;
  SceneInspector inspector_;
  std::vector<CollisionGeometry> geoms_;
};

// ---------------------------------------------------------------------------
// H-polytopes (geometry/hpolyhedron.h)
// ---------------------------------------------------------------------------
struct MinimalHPolyhedron {
  std::vector<float> A;  // row-major [num_faces][dim]
  std::vector<float> b;
  int32_t num_faces = 0;
  int32_t dim = 0;
};

class HPolyhedron {
 public:
  HPolyhedron() = default;
  explicit HPolyhedron(const MinimalHPolyhedron &m);
  HPolyhedron(const MatrixXf &A, const std::vector<float> &b);
  const MatrixXf &A() const { return A_; }
  const std::vector<float> &b() const { return b_; }
  void MakeBox(const std::vector<float> &lower, const std::vector<float> &upper);
  MinimalHPolyhedron GetMyMinimalHPolyhedron() const;
  bool PointInSet(const std::vector<float> &x, float tol = 1e-9f) const;
  // hit-and-run from previous_sample on the GPU sampler (one chain)
  std::vector<float> UniformSample(const std::vector<float> &previous_sample,
                                   int mixing_steps) const;
  // maximises the inscribed-ball radius (dense simplex on the host)
  std::vector<float> ChebyshevCenter() const;
  std::vector<float> GetFeasiblePoint() const;
  int ambient_dimension() const { return dim_; }

 private:
  MatrixXf A_;
  std::vector<float> b_;
  int dim_ = 0;
};

// ---------------------------------------------------------------------------
// GPU entry points (plant/cuda_collision_checker.h, cuda_forward_kinematics.h,
// geometry/cuda_hit_and_run_sampling.h). configurations are (dim, N).
// ---------------------------------------------------------------------------
std::vector<uint8_t> checkCollisionFreeCuda(const MatrixXf &configurations,
                                            const MinimalPlant &plant);
std::vector<uint8_t> checkCollisionFreeVoxelsCuda(const MatrixXf &configurations,
                                                  const MatrixXf &voxels, float voxel_radius,
                                                  const MinimalPlant &plant,
                                                  const std::vector<GeometryIndex> &robot_geoms);
// [N][V] flags, 1 = voxel free at config n
std::vector<uint8_t> checkCollisionFreeVoxelsWithoutSelfCollisionCuda(
    const MatrixXf &configurations, const MatrixXf &voxels, float voxel_radius,
    const MinimalPlant &plant, const std::vector<GeometryIndex> &robot_geoms);
// (4, 4*L*N): per configuration, L column-major 4x4 link transforms
MatrixXf computeForwardKinematicsCuda(const MatrixXf &configurations,
                                      const MinimalKinematicTree &tree);
std::vector<MatrixXf> UniformSampleInHPolyhedronCuda(const std::vector<HPolyhedron> &polyhedra,
                                                     const MatrixXf &starting_points,
                                                     int num_samples_per_hpolyhedron,
                                                     int mixing_steps, uint64_t seed = 1337);
// single-configuration names of the reference's CPU API
// (plant/collision_checker.h:15-75, minimal_kinematic_tree.h), evaluated on
// the calling CPU thread with the reference's CPU comparison rules
bool checkCollisionFree(const std::vector<float> &q, const MinimalPlant &plant);
bool checkCollisionFreeVoxels(const std::vector<float> &q, const MatrixXf &voxels,
                              float voxel_radius, const MinimalPlant &plant,
                              const std::vector<GeometryIndex> &robot_geoms);
bool checkEdgeCollisionFree(const std::vector<float> &q1, const std::vector<float> &q2,
                            const MinimalPlant &plant, float step_size = 0.01f);
std::vector<Matrix4f> computeLinkFrameToWorldTransformsMinimal(const std::vector<float> &q,
                                                                const MinimalKinematicTree &tree);
bool pairCollisionFree(const CollisionGeometry &a, const Matrix4f &X_W_LA,
                       const CollisionGeometry &b, const Matrix4f &X_W_LB);
// collision_checker.cpp:205-253: robot geometry vs a voxel sphere
bool geomToVoxelCollisionFree(const CollisionGeometry &geom, const Matrix4f &X_W_L,
                              const Vector3f &p_W_voxel, float voxel_radius);

// Visibility graph (planning/cuda_visibility_graph.h): upper-triangular edge
// list (i < j) of mutually visible configurations, step rule of
// checkEdgeCollisionFree.
std::vector<std::pair<int, int>> VisibilityGraph(const MatrixXf &configurations,
                                                 const MinimalPlant &plant, float step_size,
                                                 int64_t max_configs_per_batch = 10000000);

// ---------------------------------------------------------------------------
// EI-ZO (planning/cuda_edge_inflation_zero_order.h, cuda_polytope_builder.h)
// ---------------------------------------------------------------------------
struct EizoOptions {
  float configuration_margin{0.01f};
  float tau{0.5f};
  float delta{0.05f};
  float epsilon{0.01f};
  uint32_t bisection_steps{9};
  uint32_t num_particles{1000};
  uint32_t max_hyperplanes_per_iteration{10};
  uint32_t max_iterations{500};
  uint32_t mixing_steps{40};
  uint64_t seed{1337};
  bool verbose{false};
  bool track_iteration_information{true};
};

// One round (cuda_edge_inflation_zero_order.h:37-47). The first seven
// fields are the reference's; num_collisions_unadaptive is -1 in a guided
// round whose pilot skipped the test. points_post_bisection is not kept.
struct EizoIterationInfo {
  int iteration_index = 0;
  int num_samples_drawn = 0;
  int collisions_threshold = 0;
  int num_collisions_unadaptive = 0;
  int num_collisions_updated = 0;
  int num_faces_placed = 0;
  float iteration_time_ms = 0.f;
  // batched and guided rounds
  int num_collisions_found = 0;  // among all samples of the test
  int num_faces_at_start = 0;
  int num_pilot_samples = 0;
  int num_pilot_collisions = 0;
  int num_pilot_checked = 0;  // pilot samples checked exactly (fewer when pruned)
};

struct EizoInfo {
  std::vector<EizoIterationInfo> iteration_info;
  // per-region summary of the batched driver
  int iterations = 0;
  bool converged = false;
  bool face_capacity_reached = false;
  int64_t samples_drawn = 0;
  int faces_added = 0;
  // guided rounds left the region uncertified and plain EI-ZO inflated it
  // again (setGuidance plain_fallback); iterations and samples_drawn count
  // both runs
  bool plain_fallback = false;
};

uint32_t numSamplesUnadaptive(float epsilon, float delta, float tau);

// Round-by-round table of an EizoInfo (printEizoSummary,
// cuda_edge_inflation_zero_order.h:168-213).
void printEizoSummary(const EizoInfo &info);

// The particles of one region in its last round: colliding samples, their
// bisected boundary points and their projections, each dim x count.
struct EizoParticles {
  MatrixXf particles, scol, proj;
};

class CudaEdgeInflator {
 public:
  CudaEdgeInflator(const MinimalPlant &plant, const std::vector<GeometryIndex> &robot_geometry_ids,
                   const EizoOptions &options, const HPolyhedron &domain,
                   int max_regions = 1, int max_voxels = 10000, int face_capacity = 0);
  ~CudaEdgeInflator();
  CudaEdgeInflator(const CudaEdgeInflator &) = delete;
  CudaEdgeInflator &operator=(const CudaEdgeInflator &) = delete;
  HPolyhedron inflateEdge(const std::vector<float> &line_start, const std::vector<float> &line_end,
                          const MatrixXf &voxels, float voxel_radius, bool verbose = true);
  // many edges at once; budgets (optional) override num_particles per edge
  std::vector<HPolyhedron> inflateEdges(const MatrixXf &line_starts, const MatrixXf &line_ends,
                                        const MatrixXf &voxels, float voxel_radius,
                                        const std::vector<uint32_t> &particle_budgets,
                                        std::vector<EizoInfo> *infos = nullptr);
  // EI-ZO guidance (csd_eizo_set_guidance): an exact pilot that lets clearly
  // colliding regions skip a round's test, and surrogate ranking of the
  // particles to bisect. surrogate: a csd_mlp* (or null).
  // plain_fallback: regions that guided rounds leave uncertified are
  // inflated again by plain EI-ZO, whose result replaces the guided one.
  // prune_below (finite): the pilot's samples are predicted by the surrogate
  // and only those predicted below it are checked exactly
  // (csd_eizo_set_pruning); NaN checks every pilot sample.
  void setGuidance(uint32_t pilot_samples, float pilot_factor, const void *surrogate,
                   int keep_particles = 256, int line_points = 4, bool plain_fallback = false,
                   float prune_below = NAN);
  // Device-resident, graph-captured rounds (csd_eizo_set_graph): the host
  // reads the region status every rounds_per_sync rounds; 0 for host-driven
  // rounds. Same polytopes either way.
  void setGraph(int rounds_per_sync);
  // Ellipsoid metric of the face placement (csd_eizo_set_metric): a
  // symmetric positive definite dim x dim matrix, or an empty one for the
  // Euclidean metric of the reference.
  void setMetric(const MatrixXf &metric);
  // Sampler tile (csd_eizo_set_sampler_tile / csd_eizo_tune_sampler):
  // 1024-thread workgroups from this many faces on; tuneSampler() times both
  // tiles and returns the count it picked. Same polytopes either way.
  void setSamplerTile(int wide_from);
  int tuneSampler();
  const EizoInfo &lastInfo() const { return last_info_; }
  // region r of the last inflateEdge(s) call, in its last round
  EizoParticles lastParticles(int region) const;
  int faceCapacity() const { return face_capacity_; }

 private:
  MinimalPlant plant_;
  std::vector<GeometryIndex> robot_ids_;
  EizoOptions options_;
  HPolyhedron domain_;
  int max_regions_, max_voxels_, face_capacity_;
  // guidance as last set (restored after a plain fallback run)
  uint32_t pilot_samples_ = 0;
  float pilot_factor_ = 10.f;
  const void *surrogate_ = nullptr;
  int keep_particles_ = 256, line_points_ = 4;
  bool plain_fallback_ = false;
  csd_eizo *ws_ = nullptr;
  EizoInfo last_info_;
};

MinimalHPolyhedron EizoCuda(const std::vector<float> &line_start, const std::vector<float> &line_end,
                            const MatrixXf &voxels, float voxel_radius, const HPolyhedron &domain,
                            const MinimalPlant &plant,
                            const std::vector<GeometryIndex> &robot_geometry_ids,
                            const EizoOptions &options, EizoInfo &eizo_info);

// ---------------------------------------------------------------------------
// Distance field (geometry/cuda_distance_field.h:133-172): exact distance
// from each voxel of a gridSize^3 grid to the nearest occupied voxel.
// Points are (3, N) matrices; a point maps to voxel round((p - origin) /
// voxelSize) per axis.
// ---------------------------------------------------------------------------
class DistanceField {
 public:
  DistanceField(const Vector3f &origin, size_t gridSize, float voxelSize);
  ~DistanceField();
  DistanceField(const DistanceField &) = delete;
  DistanceField &operator=(const DistanceField &) = delete;
  // the PBA band parameters of the reference are accepted and unused
  void computeEDT(int phase1Band = 1, int phase2Band = 1, int phase3Band = 2);
  void setOccupancy(const MatrixXf &P_VG);
  void resetOccupancy();
  // 0 outside the grid, +inf if no voxel is occupied
  std::vector<float> getDistance(const MatrixXf &P_VG);
  const int gridSize;
  const size_t numVoxels;

 private:
  csd_distance_field *field_ = nullptr;
};

// ---------------------------------------------------------------------------
// Region editing (planning/cuda_edit_regions.h)
// ---------------------------------------------------------------------------
struct EditRegionsOptions {
  float configuration_margin{0.01f};
  uint32_t bisection_steps{9};
  bool verbose{false};
};

// Cuts trajectory collisions out of regions while keeping each region's
// segment inside it (planning/cuda_edit_regions.cu:44-300). collisions is
// (dim, C); line_segment_idxs[c] names the segment / region of collision c.
// Returns the edited regions and (projections, optimised collisions), both
// (dim, C) with the collisions grouped by segment in order of appearance, as
// the reference returns them.
std::pair<std::vector<HPolyhedron>, std::pair<MatrixXf, MatrixXf>> EditRegionsCuda(
    const MatrixXf &collisions, const std::vector<uint32_t> &line_segment_idxs,
    const MatrixXf &line_start_points, const MatrixXf &line_end_points,
    const std::vector<HPolyhedron> &regions, const MinimalPlant &plant,
    const std::vector<GeometryIndex> &robot_geometry_ids, const MatrixXf &voxels,
    float voxel_radius, const EditRegionsOptions &options);

}  // namespace csdecomp
