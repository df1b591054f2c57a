// csdecomp.cpp -- host side of the csdecomp-compatible API. Plant assembly,
// polytope bookkeeping and argument checking live here; every numerical
// evaluation of the hot path is a call through include/csdecomp_amd.h.
#include "csdecomp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <queue>
#include <stack>

namespace csdecomp {

Matrix4f identity4() {
  Matrix4f m{};
  m[0] = m[5] = m[10] = m[15] = 1.f;
  return m;
}

void check_status(int rc, const char *what) {
  if (rc == CSD_OK) return;
  std::string msg = std::string("[csdecomp_amd] ") + what + ": " + csd_status_string(rc);
  const char *detail = csd_last_error();
  if (detail && *detail) msg += std::string(" (") + detail + ")";
  if (rc == CSD_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

void ensure_device() {
  static std::once_flag once;
  static int rc = CSD_OK;
  static std::string err;
  std::call_once(once, [] {
    rc = csd_init(0);
    if (rc != CSD_OK) err = csd_last_error();
  });
  if (rc != CSD_OK)
    throw std::runtime_error("[csdecomp_amd] no usable gfx950 GPU: " + err +
                             " (this library has no CPU fallback)");
}

namespace {

// ---- device scratch: grows, never shrinks, one set per host thread ----
struct DevBuf {
  void *ptr = nullptr;
  size_t cap = 0;
  ~DevBuf() {
    if (ptr) csd_free(ptr);
  }
  void *get(size_t bytes) {
    if (bytes > cap) {
      if (ptr) csd_free(ptr);
      ptr = nullptr;
      cap = 0;
      check_status(csd_malloc(&ptr, bytes), "device allocation");
      cap = bytes;
    }
    return ptr;
  }
};
struct Scratch {
  DevBuf q, out, aux, aux2, aux3;
};
Scratch &scratch() {
  static thread_local Scratch s;
  return s;
}

void require(bool ok, const std::string &msg) {
  if (!ok) throw std::invalid_argument("[csdecomp_amd] " + msg);
}

size_t fnv(size_t h, const void *p, size_t n) {
  const unsigned char *c = (const unsigned char *)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

}  // namespace

// ===========================================================================
// KinematicTree
// ===========================================================================
void KinematicTree::addLink(const Link &link) {
  if (is_finalized_) throw std::runtime_error("Cannot add link to finalized tree.");
  if (link_name_to_index_.count(link.name))
    throw std::runtime_error("Duplicate link name detected in URDF. Name : " + link.name);
  link_name_to_index_[link.name] = (int)links_.size();
  links_.push_back(link);
}

void KinematicTree::addJoint(const Joint &joint) {
  if (is_finalized_) throw std::runtime_error("Cannot add joint to finalized tree.");
  if (joint_name_to_index_.count(joint.name))
    throw std::runtime_error("Duplicate joint name detected in URDF.");
  if (joint.parent_link < 0 || joint.child_link < 0 || joint.parent_link >= (int)links_.size() ||
      joint.child_link >= (int)links_.size())
    throw std::runtime_error("Parent or child link not found for joint: " + joint.name);
  int idx = (int)joints_.size();
  joint_name_to_index_[joint.name] = idx;
  joints_.push_back(joint);
  links_[joint.parent_link].child_joints.push_back(idx);
  links_[joint.child_link].parent_joint = idx;
}

void KinematicTree::finalize() {
  is_finalized_ = true;
  leaf_link_indices_.clear();
  joint_idx_to_config_idx_.assign(joints_.size(), 0);
  traversal_.clear();
  lower_.clear();
  upper_.clear();
  num_configuration_variables_ = 0;
  for (int i = 0; i < (int)links_.size(); ++i)
    if (links_[i].child_joints.empty()) leaf_link_indices_.push_back(i);
  for (size_t j = 0; j < joints_.size(); ++j) {
    if (joints_[j].type != FIXED) {
      joint_idx_to_config_idx_[j] = num_configuration_variables_++;
      lower_.push_back(joints_[j].position_lower_limit);
      upper_.push_back(joints_[j].position_upper_limit);
    } else {
      joint_idx_to_config_idx_[j] = -1;
    }
  }
  if (links_.empty()) return;
  // depth first from the root with an explicit stack: children are pushed in
  // order, so the last child is visited first (kinematic_tree.cpp:117-132)
  std::stack<int> todo;
  for (int cj : links_[0].child_joints) todo.push(cj);
  while (!todo.empty()) {
    int j = todo.top();
    todo.pop();
    traversal_.push_back(j);
    for (int cj : links_[joints_[j].child_link].child_joints) todo.push(cj);
  }
}

int KinematicTree::getLinkIndexByName(const std::string &name) const {
  auto it = link_name_to_index_.find(name);
  if (it == link_name_to_index_.end())
    throw std::runtime_error("[KinematicTree] Invalid link name " + name);
  return it->second;
}

std::string KinematicTree::getLinkNameByIndex(int index) const {
  if (index < 0 || index >= (int)links_.size()) throw std::runtime_error("link index not found");
  return links_[index].name;
}

const Link &KinematicTree::getLink(int index) const { return links_.at(index); }

int KinematicTree::getJointIndex(const std::string &name) const {
  auto it = joint_name_to_index_.find(name);
  return it == joint_name_to_index_.end() ? -1 : it->second;
}

const Joint &KinematicTree::getJoint(int index) const { return joints_.at(index); }

MinimalKinematicTree KinematicTree::getMyMinimalKinematicTreeStruct() const {
  MinimalKinematicTree t;
  t.num_configuration_variables = num_configuration_variables_;
  t.num_joints = (int)joints_.size();
  t.num_links = (int)links_.size();
  t.joint_traversal_sequence = traversal_;
  t.joint_idx_to_config_idx = joint_idx_to_config_idx_;
  for (size_t l = 0; l < links_.size(); ++l) {
    MinimalLink ml;
    ml.index = (int)l;
    ml.parent_joint_index = links_[l].parent_joint;
    ml.child_joint_index_array = links_[l].child_joints;
    ml.num_child_joints = (int)links_[l].child_joints.size();
    t.links.push_back(ml);
  }
  for (size_t j = 0; j < joints_.size(); ++j) {
    MinimalJoint mj;
    mj.index = (int)j;
    mj.type = joints_[j].type;
    mj.X_PL_J = joints_[j].X_PL_J;
    mj.parent_link = joints_[j].parent_link;
    mj.child_link = joints_[j].child_link;
    mj.axis = joints_[j].axis;
    t.joints.push_back(mj);
  }
  return t;
}

std::vector<Matrix4f> KinematicTree::computeLinkFrameToWorldTransforms(
    const std::vector<float> &q) const {
  if (!is_finalized_) throw std::runtime_error("KinematicTree is not finalized");
  if ((int)q.size() != num_configuration_variables_)
    throw std::invalid_argument("Incorrect number of configuration variables");
  return computeLinkFrameToWorldTransformsMinimal(q, getMyMinimalKinematicTreeStruct());
}

// ===========================================================================
// scene
// ===========================================================================
std::vector<int32_t> SceneInspector::getCollisionPairMatrix() const {
  std::vector<int32_t> m;
  m.reserve(2 * robot_to_world_collision_pairs.size());
  for (auto &p : robot_to_world_collision_pairs) {
    m.push_back(p.first);
    m.push_back(p.second);
  }
  return m;
}

namespace {
// breadth-first path between two links over the undirected joint graph
// (urdf_parser.cpp:38-71); returns link indices, empty if disconnected
std::vector<int> link_path(const KinematicTree &t, int a, int b) {
  int n = t.numLinks();
  std::vector<int> parent(n, -1);
  std::vector<char> seen(n, 0);
  std::vector<std::vector<int>> adj(n);
  for (const Joint &j : t.getJoints()) {
    adj[j.parent_link].push_back(j.child_link);
    adj[j.child_link].push_back(j.parent_link);
  }
  for (auto &v : adj) std::sort(v.begin(), v.end());
  std::queue<int> q;
  q.push(a);
  seen[a] = 1;
  while (!q.empty()) {
    int c = q.front();
    q.pop();
    if (c == b) {
      std::vector<int> path;
      for (int v = b; v != -1; v = parent[v]) path.push_back(v);
      std::reverse(path.begin(), path.end());
      return path;
    }
    for (int nb : adj[c])
      if (!seen[nb]) {
        seen[nb] = 1;
        parent[nb] = c;
        q.push(nb);
      }
  }
  return {};
}

int joint_between(const KinematicTree &t, int la, int lb) {
  const auto &js = t.getJoints();
  for (size_t j = 0; j < js.size(); ++j)
    if ((js[j].parent_link == la && js[j].child_link == lb) ||
        (js[j].parent_link == lb && js[j].child_link == la))
      return (int)j;
  return -1;
}
}  // namespace

SceneInspector finalizeScene(const KinematicTree &tree, std::vector<CollisionGeometry> &geoms,
                             const std::unordered_map<std::string, int> &names) {
  SceneInspector insp;
  insp.link_index_to_scene_collision_geometry_indices.assign(tree.numLinks(), {});
  for (size_t g = 0; g < geoms.size(); ++g) {
    if (geoms[g].link_index < 0 || geoms[g].link_index >= tree.numLinks())
      throw std::runtime_error("collision geometry attached to an unknown link");
    insp.link_index_to_scene_collision_geometry_indices[geoms[g].link_index].push_back((int)g);
  }
  // WORLD iff the link reaches the root through fixed joints only
  for (auto &co : geoms) {
    const Link &lk = tree.getLink(co.link_index);
    if (lk.parent_joint < 0) {
      co.group = WORLD;  // geometry on the root link itself
      continue;
    }
    const Joint *prev = &tree.getJoint(lk.parent_joint);
    while (true) {
      if (prev->type != FIXED) {
        co.group = ROBOT;
        break;
      }
      if (prev->parent_link == 0) {
        co.group = WORLD;
        break;
      }
      int pj = tree.getLink(prev->parent_link).parent_joint;
      if (pj == -1)
        throw std::runtime_error("The kinematic tree appears to be disconnected.");
      prev = &tree.getJoint(pj);
    }
  }
  for (size_t g = 0; g < geoms.size(); ++g)
    if (geoms[g].group == ROBOT) insp.robot_geometry_ids.push_back((GeometryIndex)g);

  auto parent_link_of = [&](int link) {
    int pj = tree.getLink(link).parent_joint;
    return pj < 0 ? -1 : tree.getJoint(pj).parent_link;
  };
  for (size_t a = 0; a + 1 < geoms.size(); ++a) {
    for (size_t b = a + 1; b < geoms.size(); ++b) {
      if (geoms[a].group == WORLD && geoms[b].group == WORLD) continue;
      int la = geoms[a].link_index, lb = geoms[b].link_index;
      bool not_successive = parent_link_of(la) != lb && parent_link_of(lb) != la;
      bool not_same = la != lb;
      std::vector<int> path = link_path(tree, la, lb);
      bool fixed_only = true;
      for (size_t s = 0; s + 1 < path.size(); ++s) {
        int j = joint_between(tree, path[s], path[s + 1]);
        if (tree.getJoint(j).type != FIXED) {
          fixed_only = false;
          break;
        }
      }
      if (not_successive && not_same && !fixed_only)
        insp.robot_to_world_collision_pairs.emplace_back((int)a, (int)b);
    }
  }
  insp.scene_collision_geometry_name_to_geometry_index = names;
  return insp;
}

csd_plant *MinimalPlant::device() const {
  ensure_device();
  const MinimalKinematicTree &t = kin_tree;
  require(t.num_links >= 1 && (int)t.links.size() >= t.num_links, "plant has no links");
  require((int)t.joints.size() >= t.num_joints && (int)t.joint_traversal_sequence.size() >= t.num_joints &&
              (int)t.joint_idx_to_config_idx.size() >= t.num_joints,
          "kinematic tree arrays are shorter than num_joints");
  require((int)scene_collision_geometries.size() >= num_scene_geometries,
          "fewer geometries than num_scene_geometries");
  require((int)collision_pairs_flat.size() >= 2 * num_collision_pairs,
          "collision_pairs_flat shorter than 2*num_collision_pairs");
  const int J = t.num_joints, G = num_scene_geometries, P = num_collision_pairs;
  std::vector<int32_t> jt(J), jp(J), jc(J), jq(J), tr(J), gt(G), gl(G), gg(G);
  std::vector<float> jx(16 * (size_t)J), ja(3 * (size_t)J), gd(3 * (size_t)G), gx(16 * (size_t)G);
  for (int j = 0; j < J; ++j) {
    const MinimalJoint &m = t.joints[j];
    jt[j] = m.type;
    jp[j] = m.parent_link;
    jc[j] = m.child_link;
    jq[j] = t.joint_idx_to_config_idx[j];
    tr[j] = t.joint_traversal_sequence[j];
    std::memcpy(&jx[16 * j], m.X_PL_J.data(), 64);
    std::memcpy(&ja[3 * j], m.axis.data(), 12);
  }
  for (int g = 0; g < G; ++g) {
    const CollisionGeometry &c = scene_collision_geometries[g];
    gt[g] = c.type;
    gl[g] = c.link_index;
    gg[g] = c.group;
    std::memcpy(&gd[3 * g], c.dimensions.data(), 12);
    std::memcpy(&gx[16 * g], c.X_L_B.data(), 64);
  }
  size_t h = 1469598103934665603ull;
  int hdr[5] = {t.num_links, J, t.num_configuration_variables, G, P};
  h = fnv(h, hdr, sizeof(hdr));
  for (auto *v : {&jt, &jp, &jc, &jq, &tr, &gt, &gl, &gg}) h = fnv(h, v->data(), v->size() * 4);
  for (auto *v : {&jx, &ja, &gd, &gx}) h = fnv(h, v->data(), v->size() * 4);
  h = fnv(h, collision_pairs_flat.data(), 8 * (size_t)P);
  if (dev_ && h == dev_hash_) return dev_.get();
  csd_plant_desc d;
  d.num_links = t.num_links;
  d.num_joints = J;
  d.num_configuration_variables = t.num_configuration_variables;
  d.joint_type = jt.data();
  d.joint_parent_link = jp.data();
  d.joint_child_link = jc.data();
  d.joint_config_idx = jq.data();
  d.traversal = tr.data();
  d.joint_X_PL_J = jx.data();
  d.joint_axis = ja.data();
  d.num_geometries = G;
  d.geom_type = gt.data();
  d.geom_link = gl.data();
  d.geom_group = gg.data();
  d.geom_dims = gd.data();
  d.geom_X_L_B = gx.data();
  d.num_collision_pairs = P;
  d.collision_pairs = collision_pairs_flat.data();
  csd_plant *p = nullptr;
  check_status(csd_plant_create(&d, &p), "plant upload");
  dev_.reset(p, [](csd_plant *x) { csd_plant_destroy(x); });
  dev_hash_ = h;
  return p;
}

MinimalPlant Plant::getMinimalPlant() const {
  MinimalPlant mp;
  mp.kin_tree = kin_tree_.getMyMinimalKinematicTreeStruct();
  mp.scene_collision_geometries = geoms_;
  mp.collision_pairs_flat = inspector_.getCollisionPairMatrix();
  mp.num_scene_geometries = (int)geoms_.size();
  mp.num_collision_pairs = (int)inspector_.robot_to_world_collision_pairs.size();
  return mp;
}

std::vector<std::string> Plant::getSceneCollisionGeometryNames() const {
  std::vector<std::string> keys(inspector_.scene_collision_geometry_name_to_geometry_index.size());
  for (auto &kv : inspector_.scene_collision_geometry_name_to_geometry_index)
    if (kv.second >= 0 && kv.second < (int)keys.size()) keys[kv.second] = kv.first;
  return keys;
}

int Plant::getSceneCollisionGeometryIndexByName(const std::string &name) const {
  auto it = inspector_.scene_collision_geometry_name_to_geometry_index.find(name);
  if (it == inspector_.scene_collision_geometry_name_to_geometry_index.end())
    throw std::out_of_range("Geometry name not found: " + name);
  return it->second;
}

// ===========================================================================
// HPolyhedron (hpolyhedron.cpp)
// ===========================================================================
HPolyhedron::HPolyhedron(const MinimalHPolyhedron &m) : A_(m.num_faces, m.dim), dim_(m.dim) {
  for (int r = 0; r < m.num_faces; ++r)
    for (int c = 0; c < m.dim; ++c) A_(r, c) = m.A[(size_t)r * m.dim + c];
  b_.assign(m.b.begin(), m.b.begin() + m.num_faces);
}

HPolyhedron::HPolyhedron(const MatrixXf &A, const std::vector<float> &b)
    : A_(A), b_(b), dim_(A.cols) {
  if ((int)b.size() != A.rows) throw std::invalid_argument("A and b dimensions do not match");
}

void HPolyhedron::MakeBox(const std::vector<float> &lo, const std::vector<float> &hi) {
  if (lo.size() != hi.size())
    throw std::invalid_argument("Lower and upper limits must have the same dimension");
  int d = (int)lo.size();
  dim_ = d;
  A_ = MatrixXf(2 * d, d);
  b_.assign(2 * d, 0.f);
  for (int i = 0; i < d; ++i) {  // rows interleaved (-e_i, +e_i), hpolyhedron.cpp:104-122
    A_(2 * i, i) = -1.f;
    A_(2 * i + 1, i) = 1.f;
    b_[2 * i] = -lo[i];
    b_[2 * i + 1] = hi[i];
  }
}

MinimalHPolyhedron HPolyhedron::GetMyMinimalHPolyhedron() const {
  MinimalHPolyhedron m;
  m.num_faces = A_.rows;
  m.dim = A_.cols;
  m.A.resize((size_t)m.num_faces * m.dim);
  for (int r = 0; r < m.num_faces; ++r)
    for (int c = 0; c < m.dim; ++c) m.A[(size_t)r * m.dim + c] = A_(r, c);
  m.b = b_;
  return m;
}

bool HPolyhedron::PointInSet(const std::vector<float> &x, float tol) const {
  if ((int)x.size() != dim_)
    throw std::invalid_argument("Point dimension does not match polyhedron dimension");
  for (int r = 0; r < A_.rows; ++r) {
    float s = 0.f;
    for (int c = 0; c < dim_; ++c) s += A_(r, c) * x[c];
    if (!(s <= b_[r] + tol)) return false;
  }
  return true;
}

namespace {
// Dense two-phase simplex (Chvatal's auxiliary-variable start, Bland's rule):
// maximise c.z s.t. M z <= h, z >= 0. Returns false if infeasible/unbounded.
bool simplex_max(const std::vector<std::vector<double>> &M, const std::vector<double> &h,
                 const std::vector<double> &c, std::vector<double> &z) {
  const int m = (int)M.size(), n = (int)c.size();
  // tableau columns: z (n), w (1), slack (m), rhs
  const int W = n, S = n + 1, R = n + 1 + m, C = R + 1;
  std::vector<std::vector<double>> T(m + 1, std::vector<double>(C, 0.0));
  std::vector<int> basis(m);
  for (int i = 0; i < m; ++i) {
    for (int j = 0; j < n; ++j) T[i][j] = M[i][j];
    T[i][W] = -1.0;
    T[i][S + i] = 1.0;
    T[i][R] = h[i];
    basis[i] = S + i;
  }
  const double eps = 1e-10;
  auto pivot = [&](int r, int col) {
    double p = T[r][col];
    for (double &v : T[r]) v /= p;
    for (int i = 0; i <= m; ++i)
      if (i != r && T[i][col] != 0.0) {
        double f = T[i][col];
        for (int k = 0; k < C; ++k) T[i][k] -= f * T[r][k];
      }
    basis[r] = col;
  };
  auto run = [&](int forbid) {
    for (int it = 0; it < 50000; ++it) {
      int col = -1;
      for (int k = 0; k < R; ++k)
        if (k != forbid && T[m][k] < -eps) {
          col = k;
          break;
        }
      if (col < 0) return true;
      int r = -1;
      double best = 0;
      for (int i = 0; i < m; ++i)
        if (T[i][col] > eps) {
          double ratio = T[i][R] / T[i][col];
          if (r < 0 || ratio < best - 1e-12 || (std::fabs(ratio - best) <= 1e-12 && basis[i] < basis[r])) {
            r = i;
            best = ratio;
          }
        }
      if (r < 0) return false;  // unbounded
      pivot(r, col);
    }
    return false;
  };
  // phase 1: minimise w (objective row stores -(max -w) coefficients)
  int worst = 0;
  for (int i = 1; i < m; ++i)
    if (T[i][R] < T[worst][R]) worst = i;
  if (m > 0 && T[worst][R] < 0) {
    std::fill(T[m].begin(), T[m].end(), 0.0);
    T[m][W] = 1.0;
    pivot(worst, W);
    if (!run(-1)) return false;
    if (T[m][R] < -1e-7) return false;  // infeasible
    for (int i = 0; i < m; ++i)
      if (basis[i] == W) {
        for (int k = 0; k < R; ++k)
          if (k != W && std::fabs(T[i][k]) > eps) {
            pivot(i, k);
            break;
          }
      }
  }
  // phase 2
  std::fill(T[m].begin(), T[m].end(), 0.0);
  for (int j = 0; j < n; ++j) T[m][j] = -c[j];
  for (int i = 0; i < m; ++i) {
    int bcol = basis[i];
    if (bcol < n && T[m][bcol] != 0.0) {
      double f = T[m][bcol];
      for (int k = 0; k < C; ++k) T[m][k] -= f * T[i][k];
    }
  }
  for (int i = 0; i <= m; ++i) T[i][W] = 0.0;  // w stays at zero
  if (!run(W)) return false;
  z.assign(n, 0.0);
  for (int i = 0; i < m; ++i)
    if (basis[i] < n) z[basis[i]] = T[i][R];
  return true;
}
}  // namespace

std::vector<float> HPolyhedron::ChebyshevCenter() const {
  // max r s.t. a_i x + |a_i| r <= b_i, r >= 0, x = xp - xn
  const int m = A_.rows, d = dim_;
  std::vector<std::vector<double>> M(m, std::vector<double>(2 * d + 1, 0.0));
  std::vector<double> h(m), c(2 * d + 1, 0.0), z;
  for (int i = 0; i < m; ++i) {
    double nrm = 0;
    for (int j = 0; j < d; ++j) {
      M[i][j] = A_(i, j);
      M[i][d + j] = -A_(i, j);
      nrm += (double)A_(i, j) * A_(i, j);
    }
    M[i][2 * d] = std::sqrt(nrm);
    h[i] = b_[i];
  }
  c[2 * d] = 1.0;
  // bound r so the LP stays bounded for unbounded polyhedra
  M.push_back(std::vector<double>(2 * d + 1, 0.0));
  M.back()[2 * d] = 1.0;
  h.push_back(1e6);
  if (!simplex_max(M, h, c, z)) throw std::runtime_error("Failed to solve LP for Chebyshev center");
  std::vector<float> x(d);
  for (int j = 0; j < d; ++j) x[j] = (float)(z[j] - z[d + j]);
  return x;
}

std::vector<float> HPolyhedron::GetFeasiblePoint() const {
  std::vector<float> x = ChebyshevCenter();
  if (!PointInSet(x, 1e-5f)) throw std::runtime_error("The polyhedron is empty.");
  return x;
}

// ===========================================================================
// GPU entry points
// ===========================================================================
namespace {
const float *upload_configs(const MatrixXf &q, int dof) {
  require(q.rows == dof, "configurations have " + std::to_string(q.rows) +
                             " rows but the plant has " + std::to_string(dof) +
                             " configuration variables");
  size_t bytes = q.data.size() * sizeof(float);
  void *d = scratch().q.get(std::max<size_t>(bytes, 4));
  check_status(csd_memcpy_h2d(d, q.data.data(), bytes, nullptr), "configuration upload");
  return (const float *)d;
}
}  // namespace

std::vector<uint8_t> checkCollisionFreeCuda(const MatrixXf &configurations,
                                            const MinimalPlant &plant) {
  csd_plant *p = plant.device();
  const int dof = plant.kin_tree.num_configuration_variables;
  const int64_t n = configurations.cols;
  std::vector<uint8_t> out(n);
  if (n == 0) return out;
  const float *dq = upload_configs(configurations, dof);
  uint8_t *dout = (uint8_t *)scratch().out.get(n);
  // (dim, N) column-major: config n at q[n*dim], joint stride 1
  check_status(csd_check_collision_free(p, dq, n, dof, 1, dout, nullptr), "collision check");
  check_status(csd_memcpy_d2h(out.data(), dout, n, nullptr), "result download");
  check_status(csd_stream_synchronize(nullptr), "synchronize");
  return out;
}

std::vector<uint8_t> checkCollisionFreeVoxelsCuda(const MatrixXf &configurations,
                                                  const MatrixXf &voxels, float voxel_radius,
                                                  const MinimalPlant &plant,
                                                  const std::vector<GeometryIndex> &robot_geoms) {
  csd_plant *p = plant.device();
  const int dof = plant.kin_tree.num_configuration_variables;
  const int64_t n = configurations.cols;
  std::vector<uint8_t> out(n);
  if (n == 0) return out;
  require(voxels.cols == 0 || voxels.rows == 3, "voxels must be a 3 x V matrix");
  for (auto g : robot_geoms)
    require(g >= 0 && g < plant.num_scene_geometries, "invalid robot geometry id");
  const float *dq = upload_configs(configurations, dof);
  float *dv = nullptr;
  if (voxels.cols > 0) {
    dv = (float *)scratch().aux.get(voxels.data.size() * 4);
    check_status(csd_memcpy_h2d(dv, voxels.data.data(), voxels.data.size() * 4, nullptr),
                 "voxel upload");
  }
  uint8_t *dout = (uint8_t *)scratch().out.get(n);
  check_status(csd_check_collision_free_voxels(p, dq, n, dof, 1, dv, voxels.cols, voxel_radius,
                                               robot_geoms.data(), (int)robot_geoms.size(), dout,
                                               nullptr),
               "voxel collision check");
  check_status(csd_memcpy_d2h(out.data(), dout, n, nullptr), "result download");
  check_status(csd_stream_synchronize(nullptr), "synchronize");
  return out;
}

std::vector<uint8_t> checkCollisionFreeVoxelsWithoutSelfCollisionCuda(
    const MatrixXf &configurations, const MatrixXf &voxels, float voxel_radius,
    const MinimalPlant &plant, const std::vector<GeometryIndex> &robot_geoms) {
  csd_plant *p = plant.device();
  const int dof = plant.kin_tree.num_configuration_variables;
  const int64_t n = configurations.cols, V = voxels.cols;
  require(V > 0, "at least one voxel is required");
  require(voxels.rows == 3, "voxels must be a 3 x V matrix");
  std::vector<uint8_t> out((size_t)n * V);
  if (n == 0) return out;
  const float *dq = upload_configs(configurations, dof);
  float *dv = (float *)scratch().aux.get(voxels.data.size() * 4);
  check_status(csd_memcpy_h2d(dv, voxels.data.data(), voxels.data.size() * 4, nullptr),
               "voxel upload");
  uint8_t *dout = (uint8_t *)scratch().out.get(out.size());
  check_status(csd_voxel_flags(p, dq, n, dof, 1, dv, (int)V, voxel_radius, robot_geoms.data(),
                               (int)robot_geoms.size(), dout, nullptr),
               "voxel flags");
  check_status(csd_memcpy_d2h(out.data(), dout, out.size(), nullptr), "result download");
  check_status(csd_stream_synchronize(nullptr), "synchronize");
  return out;
}


MatrixXf computeForwardKinematicsCuda(const MatrixXf &configurations,
                                      const MinimalKinematicTree &tree) {
  static thread_local MinimalPlant cache;
  cache.kin_tree = tree;
  csd_plant *p = cache.device();
  const int L = tree.num_links;
  const int64_t n = configurations.cols;
  MatrixXf out(4, (int)(4 * L * n));
  if (n == 0) return out;
  const float *dq = upload_configs(configurations, tree.num_configuration_variables);
  size_t bytes = (size_t)n * L * 16 * sizeof(float);
  float *dt = (float *)scratch().out.get(bytes);
  check_status(csd_forward_kinematics(p, dq, n, tree.num_configuration_variables, 1, dt, nullptr),
               "forward kinematics");
  check_status(csd_memcpy_d2h(out.data.data(), dt, bytes, nullptr), "result download");
  check_status(csd_stream_synchronize(nullptr), "synchronize");
  return out;
}

std::vector<MatrixXf> UniformSampleInHPolyhedronCuda(const std::vector<HPolyhedron> &polys,
                                                     const MatrixXf &starts, int n,
                                                     int mixing_steps, uint64_t seed) {
  ensure_device();
  std::vector<MatrixXf> out;
  if (polys.empty()) return out;
  const int dim = polys[0].ambient_dimension();
  require(n >= 0 && mixing_steps >= 0, "negative sample or mixing count");
  require(starts.rows == dim && starts.cols == (int)polys.size(),
          "starting_points must be dim x num_polyhedra");
  int fmax = 0;
  for (auto &p : polys) {
    require(p.ambient_dimension() == dim, "polyhedra of different dimensions");
    fmax = std::max(fmax, p.A().rows);
  }
  for (size_t k = 0; k < polys.size(); ++k) {
    std::vector<float> x(dim);
    for (int d = 0; d < dim; ++d) x[d] = starts(d, (int)k);
    if (!polys[k].PointInSet(x, 1e-4f))
      throw std::runtime_error("[UniformSampleInHPolyhedronCuda] starting point " +
                               std::to_string(k) + " is not contained in its polyhedron");
  }
  const int P = (int)polys.size();
  const int F = std::max(fmax, 1);
  std::vector<float> A((size_t)P * F * dim, 0.f), b((size_t)P * F, 0.f);
  std::vector<int32_t> nf(P);
  for (int k = 0; k < P; ++k) {
    MinimalHPolyhedron m = polys[k].GetMyMinimalHPolyhedron();
    nf[k] = m.num_faces;
    std::copy(m.A.begin(), m.A.end(), A.begin() + (size_t)k * F * dim);
    std::copy(m.b.begin(), m.b.end(), b.begin() + (size_t)k * F);
  }
  float *dA = (float *)scratch().aux.get(A.size() * 4);
  float *db = (float *)scratch().aux2.get(b.size() * 4 + P * 4 + (size_t)P * dim * 4);
  int32_t *dnf = (int32_t *)(db + b.size());
  float *dst = (float *)(dnf + P);
  check_status(csd_memcpy_h2d(dA, A.data(), A.size() * 4, nullptr), "upload");
  check_status(csd_memcpy_h2d(db, b.data(), b.size() * 4, nullptr), "upload");
  check_status(csd_memcpy_h2d(dnf, nf.data(), P * 4, nullptr), "upload");
  check_status(csd_memcpy_h2d(dst, starts.data.data(), (size_t)P * dim * 4, nullptr), "upload");
  size_t obytes = (size_t)P * n * dim * 4;
  float *dout = (float *)scratch().out.get(std::max<size_t>(obytes, 4));
  check_status(csd_hit_and_run(dA, db, dnf, F, dim, P, dst, n, mixing_steps, seed, 0, dout,
                               nullptr),
               "hit-and-run");
  std::vector<float> host(obytes / 4);
  check_status(csd_memcpy_d2h(host.data(), dout, obytes, nullptr), "download");
  check_status(csd_stream_synchronize(nullptr), "synchronize");
  for (int k = 0; k < P; ++k) {
    MatrixXf m(dim, n);
    std::copy(host.begin() + (size_t)k * n * dim, host.begin() + (size_t)(k + 1) * n * dim,
              m.data.begin());
    out.push_back(std::move(m));
  }
  return out;
}

std::vector<float> HPolyhedron::UniformSample(const std::vector<float> &prev, int mixing) const {
  MatrixXf s(dim_, 1);
  require((int)prev.size() == dim_, "wrong dimension of previous sample!");
  s.data = prev;
  static thread_local uint64_t calls = 0;
  auto r = UniformSampleInHPolyhedronCuda({*this}, s, 1, mixing, 1337 + calls++);
  return r[0].data;
}

std::vector<std::pair<int, int>> VisibilityGraph(const MatrixXf &configurations,
                                                 const MinimalPlant &plant, float step_size,
                                                 int64_t max_configs_per_batch) {
  csd_plant *p = plant.device();
  const int dof = plant.kin_tree.num_configuration_variables;
  const int n = configurations.cols;
  require(configurations.rows == dof, "configurations have the wrong number of rows");
  require(n >= 2, "at least two configurations are required");
  require(step_size >= 1e-12f, "step size must be positive");
  require(max_configs_per_batch > 1, "max_configs_per_batch must exceed 1");
  const int64_t E = (int64_t)n * (n - 1) / 2;
  float *dn = (float *)scratch().aux.get(configurations.data.size() * 4);
  check_status(csd_memcpy_h2d(dn, configurations.data.data(), configurations.data.size() * 4,
                              nullptr),
               "upload");
  // edges are checked in fixed-size batches of the upper triangle
  const int64_t batch = std::min<int64_t>(E, 1 << 20);
  std::vector<int32_t> edges(2 * batch);
  std::vector<uint8_t> flags(batch);
  int32_t *de = (int32_t *)scratch().aux2.get(edges.size() * 4);
  uint8_t *dout = (uint8_t *)scratch().out.get(batch);
  std::vector<std::pair<int, int>> result;
  int i = 0, j = 1;
  for (int64_t done = 0; done < E; done += batch) {
    int64_t cnt = std::min(batch, E - done);
    for (int64_t k = 0; k < cnt; ++k) {
      edges[2 * k] = i;
      edges[2 * k + 1] = j;
      if (++j == n) {
        ++i;
        j = i + 1;
      }
    }
    check_status(csd_memcpy_h2d(de, edges.data(), cnt * 8, nullptr), "upload");
    check_status(csd_validate_edges(p, dn, n, de, cnt, step_size, 1, dout, nullptr),
                 "edge validation");
    check_status(csd_memcpy_d2h(flags.data(), dout, cnt, nullptr), "download");
    check_status(csd_stream_synchronize(nullptr), "synchronize");
    for (int64_t k = 0; k < cnt; ++k)
      if (flags[k]) result.emplace_back(edges[2 * k], edges[2 * k + 1]);
  }
  return result;
}

// ===========================================================================
// EI-ZO
// ===========================================================================
uint32_t numSamplesUnadaptive(float epsilon, float delta, float tau) {
  // cuda_set_builder_utils.h:198-203, float arithmetic
  return (uint32_t)std::ceil(2 * std::log(1 / delta) / (epsilon * tau * tau));
}

namespace {
csd_eizo_options to_c(const EizoOptions &o) {
  csd_eizo_options c;
  c.configuration_margin = o.configuration_margin;
  c.tau = o.tau;
  c.delta = o.delta;
  c.epsilon = o.epsilon;
  c.bisection_steps = o.bisection_steps;
  c.num_particles = o.num_particles;
  c.max_hyperplanes_per_iteration = o.max_hyperplanes_per_iteration;
  c.max_iterations = o.max_iterations;
  c.mixing_steps = o.mixing_steps;
  c.seed = o.seed;
  return c;
}
}  // namespace

CudaEdgeInflator::CudaEdgeInflator(const MinimalPlant &plant,
                                   const std::vector<GeometryIndex> &robot_geometry_ids,
                                   const EizoOptions &options, const HPolyhedron &domain,
                                   int max_regions, int max_voxels, int face_capacity)
    : plant_(plant),
      robot_ids_(robot_geometry_ids),
      options_(options),
      domain_(domain),
      max_regions_(max_regions),
      max_voxels_(max_voxels) {
  const int dim = domain.ambient_dimension();
  require(dim == plant.kin_tree.num_configuration_variables,
          "domain dimension differs from the plant's configuration variables");
  for (auto g : robot_geometry_ids)
    require(g >= 0 && g < plant.num_scene_geometries, "invalid robot geometry id");
  require(max_regions >= 1, "max_regions must be positive");
  // face capacity (domain faces included): as given, else the domain plus
  // every face the iterations could add, rounded up, at most 4096. A region
  // whose faces reach the capacity stops inflating (status "capacity").
  if (face_capacity > 0) {
    face_capacity_ = face_capacity;
  } else {
    int64_t cap = (int64_t)domain.A().rows +
                  (int64_t)options.max_iterations * options.max_hyperplanes_per_iteration;
    cap = std::min<int64_t>(cap, 4096);
    face_capacity_ = (int)((cap + 15) / 16 * 16);
  }
  require(domain.A().rows < face_capacity_, "domain uses the whole face capacity");
  csd_eizo_options o = to_c(options);
  check_status(csd_eizo_create(plant_.device(), robot_ids_.data(), (int)robot_ids_.size(), &o,
                               max_regions_, face_capacity_, (int)options.num_particles,
                               max_voxels_, &ws_),
               "EI-ZO workspace");
}

void CudaEdgeInflator::setGuidance(uint32_t pilot_samples, float pilot_factor,
                                   const void *surrogate, int keep_particles, int line_points,
                                   bool plain_fallback, float prune_below) {
  check_status(csd_eizo_set_guidance(ws_, pilot_samples, pilot_factor,
                                     static_cast<const csd_mlp *>(surrogate), keep_particles,
                                     line_points),
               "EI-ZO guidance");
  const bool prune = std::isfinite(prune_below);
  check_status(csd_eizo_set_pruning(ws_, prune ? 1 : 0, prune ? prune_below : 0.f),
               "EI-ZO pruned pilot");
  pilot_samples_ = pilot_samples;
  pilot_factor_ = pilot_factor;
  surrogate_ = surrogate;
  keep_particles_ = keep_particles;
  line_points_ = line_points;
  plain_fallback_ = plain_fallback;
}

void CudaEdgeInflator::setMetric(const MatrixXf &metric) {
  if (metric.rows == 0 && metric.cols == 0) {
    check_status(csd_eizo_set_metric(ws_, nullptr), "EI-ZO metric");
    return;
  }
  const int dim = domain_.ambient_dimension();
  require(metric.rows == dim && metric.cols == dim, "metric must be dim x dim");
  std::vector<float> rm((size_t)dim * dim);
  for (int r = 0; r < dim; ++r)
    for (int c = 0; c < dim; ++c) rm[(size_t)r * dim + c] = metric(r, c);
  check_status(csd_eizo_set_metric(ws_, rm.data()), "EI-ZO metric");
}

void CudaEdgeInflator::setSamplerTile(int wide_from) {
  check_status(csd_eizo_set_sampler_tile(ws_, wide_from), "EI-ZO sampler tile");
}

int CudaEdgeInflator::tuneSampler() {
  int32_t from = 0;
  check_status(csd_eizo_tune_sampler(ws_, &from), "EI-ZO sampler tuning");
  return from;
}

void CudaEdgeInflator::setGraph(int rounds_per_sync) {
  check_status(csd_eizo_set_graph(ws_, rounds_per_sync), "EI-ZO device-resident rounds");
}

CudaEdgeInflator::~CudaEdgeInflator() {
  if (ws_) csd_eizo_destroy(ws_);
}

std::vector<HPolyhedron> CudaEdgeInflator::inflateEdges(const MatrixXf &ls, const MatrixXf &le,
                                                        const MatrixXf &voxels, float voxel_radius,
                                                        const std::vector<uint32_t> &budgets,
                                                        std::vector<EizoInfo> *infos) {
  const int dim = domain_.ambient_dimension();
  const int R = ls.cols;
  require(ls.rows == dim && le.rows == dim && le.cols == R, "line segments must be dim x R");
  require(R <= max_regions_, "more edges than the inflator was created for");
  require(voxels.cols <= max_voxels_, "more voxels than the inflator was created for");
  require(voxels.cols == 0 || voxels.rows == 3, "voxels must be a 3 x V matrix");
  require(budgets.empty() || (int)budgets.size() == R, "one particle budget per edge");
  for (int r = 0; r < R; ++r) {
    std::vector<float> a(dim), b(dim);
    for (int d = 0; d < dim; ++d) {
      a[d] = ls(d, r);
      b[d] = le(d, r);
    }
    require(domain_.PointInSet(a, 1e-5f) && domain_.PointInSet(b, 1e-5f),
            "line segment endpoints must lie in the domain");
  }
  MinimalHPolyhedron dom = domain_.GetMyMinimalHPolyhedron();
  std::vector<float> A((size_t)R * face_capacity_ * dim), b((size_t)R * face_capacity_);
  std::vector<csd_eizo_region_info> inf(R);
  check_status(csd_eizo_run(ws_, dom.A.data(), dom.b.data(), dom.num_faces, ls.data.data(),
                            le.data.data(), R, budgets.empty() ? nullptr : budgets.data(),
                            voxels.data.data(), voxels.cols, voxel_radius, A.data(), b.data(),
                            inf.data()),
               "EI-ZO");
  std::vector<std::vector<EizoIterationInfo>> rounds(R);
  const bool track = options_.track_iteration_information && infos;
  auto collect = [&](int k, int r) {
    int32_t n = 0;
    check_status(csd_eizo_get_rounds(ws_, k, nullptr, 0, &n), "EI-ZO rounds");
    std::vector<csd_eizo_round_info> v(n);
    if (n) check_status(csd_eizo_get_rounds(ws_, k, v.data(), n, &n), "EI-ZO rounds");
    const int base = (int)rounds[r].size();
    for (const csd_eizo_round_info &e : v) {
      EizoIterationInfo it;
      it.iteration_index = base + e.iteration_index;
      it.num_samples_drawn = e.num_samples_drawn;
      it.collisions_threshold = (int)e.collisions_threshold;
      it.num_collisions_unadaptive = e.num_collisions_unadaptive;
      it.num_collisions_updated = e.num_collisions_updated;
      it.num_faces_placed = e.num_faces_placed;
      it.iteration_time_ms = e.iteration_time_ms;
      it.num_collisions_found = e.num_collisions_found;
      it.num_faces_at_start = e.num_faces_at_start;
      it.num_pilot_samples = e.pilot_samples;
      it.num_pilot_collisions = e.pilot_collisions;
      it.num_pilot_checked = e.pilot_checked;
      rounds[r].push_back(it);
    }
  };
  if (track)
    for (int r = 0; r < R; ++r) collect(r, r);
  std::vector<char> fell_back(R, 0);
  if (plain_fallback_ && (pilot_samples_ || surrogate_)) {
    // plain EI-ZO for the regions the guided rounds left uncertified, in the
    // same workspace with the guidance off (restored afterwards)
    std::vector<int> redo;
    for (int r = 0; r < R; ++r)
      if (!inf[r].converged) redo.push_back(r);
    if (!redo.empty()) {
      const int n = (int)redo.size();
      MatrixXf ls2(dim, n), le2(dim, n);
      std::vector<uint32_t> bud2;
      for (int k = 0; k < n; ++k) {
        for (int d = 0; d < dim; ++d) {
          ls2(d, k) = ls(d, redo[k]);
          le2(d, k) = le(d, redo[k]);
        }
        if (!budgets.empty()) bud2.push_back(budgets[redo[k]]);
      }
      std::vector<float> A2((size_t)n * face_capacity_ * dim), b2((size_t)n * face_capacity_);
      std::vector<csd_eizo_region_info> inf2(n);
      int rc = csd_eizo_set_guidance(ws_, 0, 1.f, nullptr, 1, 1);
      if (rc == CSD_OK)
        rc = csd_eizo_run(ws_, dom.A.data(), dom.b.data(), dom.num_faces, ls2.data.data(),
                          le2.data.data(), n, bud2.empty() ? nullptr : bud2.data(),
                          voxels.data.data(), voxels.cols, voxel_radius, A2.data(), b2.data(),
                          inf2.data());
      const std::string err = rc == CSD_OK ? std::string() : std::string(csd_last_error());
      check_status(csd_eizo_set_guidance(ws_, pilot_samples_, pilot_factor_,
                                         static_cast<const csd_mlp *>(surrogate_),
                                         keep_particles_, line_points_),
                   "EI-ZO guidance");
      if (rc != CSD_OK) throw std::runtime_error("[csdecomp_amd] EI-ZO plain fallback: " + err);
      for (int k = 0; k < n; ++k) {
        const int r = redo[k];
        std::copy(A2.begin() + (size_t)k * face_capacity_ * dim,
                  A2.begin() + (size_t)(k + 1) * face_capacity_ * dim,
                  A.begin() + (size_t)r * face_capacity_ * dim);
        std::copy(b2.begin() + (size_t)k * face_capacity_, b2.begin() + (size_t)(k + 1) * face_capacity_,
                  b.begin() + (size_t)r * face_capacity_);
        csd_eizo_region_info g = inf[r];
        inf[r] = inf2[k];
        inf[r].iterations += g.iterations;
        inf[r].samples_drawn += g.samples_drawn;
        inf[r].collisions_found += g.collisions_found;
        fell_back[r] = 1;
        if (track) collect(k, r);  // the plain rounds follow the guided ones
      }
    }
  }
  std::vector<HPolyhedron> out;
  if (infos) infos->assign(R, EizoInfo());
  for (int r = 0; r < R; ++r) {
    MinimalHPolyhedron m;
    m.dim = dim;
    m.num_faces = inf[r].num_faces;
    m.A.assign(A.begin() + (size_t)r * face_capacity_ * dim,
               A.begin() + ((size_t)r * face_capacity_ + m.num_faces) * dim);
    m.b.assign(b.begin() + (size_t)r * face_capacity_,
               b.begin() + (size_t)r * face_capacity_ + m.num_faces);
    out.emplace_back(m);
    if (infos) {
      EizoInfo &e = (*infos)[r];
      e.iterations = inf[r].iterations;
      e.converged = inf[r].converged != 0;
      e.face_capacity_reached = inf[r].capacity_reached != 0;
      e.samples_drawn = inf[r].samples_drawn;
      e.faces_added = inf[r].faces_added;
      e.plain_fallback = fell_back[r] != 0;
      e.iteration_info = std::move(rounds[r]);
    }
  }
  return out;
}

EizoParticles CudaEdgeInflator::lastParticles(int region) const {
  int32_t n = 0;
  check_status(csd_eizo_last_particles(ws_, region, nullptr, nullptr, nullptr, 0, &n),
               "EI-ZO particles");
  const int dim = domain_.ambient_dimension();
  EizoParticles p;
  p.particles = MatrixXf(dim, n);
  p.scol = MatrixXf(dim, n);
  p.proj = MatrixXf(dim, n);
  if (n)
    check_status(csd_eizo_last_particles(ws_, region, p.particles.data.data(), p.scol.data.data(),
                                         p.proj.data.data(), n, &n),
                 "EI-ZO particles");
  return p;
}

void printEizoSummary(const EizoInfo &info) {
  printf("EI-ZO summary\n=============================\nTotal iterations: %zu\n\n",
         info.iteration_info.size());
  printf("%-8s%-10s%-12s%-10s%-12s%-8s%-12s\n", "Iter.", "Samples", "Col. (UT)", "Thresh.",
         "Col. (OPT)", "Faces", "Time (ms)");
  printf("%s\n", std::string(72, '-').c_str());
  long long samples = 0, col_ut = 0, col_opt = 0, faces = 0;
  double ms = 0.0;
  for (const EizoIterationInfo &it : info.iteration_info) {
    printf("%-8d%-10d%-12d%-10d%-12d%-8d%-12.2f\n", it.iteration_index, it.num_samples_drawn,
           it.num_collisions_unadaptive, it.collisions_threshold, it.num_collisions_updated,
           it.num_faces_placed, it.iteration_time_ms);
    samples += it.num_samples_drawn;
    col_ut += std::max(it.num_collisions_unadaptive, 0);
    col_opt += it.num_collisions_updated;
    faces += it.num_faces_placed;
    ms += it.iteration_time_ms;
  }
  printf("%s\nTotals:\n  Samples drawn: %lld\n  Collisions (UT): %lld\n  Collisions (OPT): %lld\n"
         "  Faces placed: %lld\n  Total time: %.2f ms\n",
         std::string(72, '-').c_str(), samples, col_ut, col_opt, faces, ms);
  fflush(stdout);
}

HPolyhedron CudaEdgeInflator::inflateEdge(const std::vector<float> &line_start,
                                          const std::vector<float> &line_end,
                                          const MatrixXf &voxels, float voxel_radius,
                                          bool verbose) {
  const int dim = domain_.ambient_dimension();
  require((int)line_start.size() == dim && (int)line_end.size() == dim,
          "line segment endpoints have the wrong dimension");
  MatrixXf ls(dim, 1), le(dim, 1);
  ls.data = line_start;
  le.data = line_end;
  std::vector<EizoInfo> infos;
  std::vector<HPolyhedron> r = inflateEdges(ls, le, voxels, voxel_radius, {}, &infos);
  last_info_ = infos[0];
  // the reference stops with this error when MinimalHPolyhedron's static
  // face storage overflows (cuda_edge_inflation_zero_order.cu:299-304)
  if (last_info_.face_capacity_reached)
    throw std::runtime_error(
        "Max face count exceeded: the region filled the inflator's face capacity (" +
        std::to_string(face_capacity_) + " faces); create it with a larger face_capacity.");
  if (verbose)
    printf("EI-ZO: %d iterations, %s, %lld samples, %d faces added\n", last_info_.iterations,
           last_info_.converged ? "certified" : "iteration limit reached",
           (long long)last_info_.samples_drawn, last_info_.faces_added);
  return r[0];
}

MinimalHPolyhedron EizoCuda(const std::vector<float> &line_start, const std::vector<float> &line_end,
                            const MatrixXf &voxels, float voxel_radius, const HPolyhedron &domain,
                            const MinimalPlant &plant,
                            const std::vector<GeometryIndex> &robot_geometry_ids,
                            const EizoOptions &options, EizoInfo &eizo_info) {
  // argument checks of cuda_edge_inflation_zero_order.cu:355-386 (without the
  // reference's dim < 14 restriction)
  if (domain.ambient_dimension() > 32)
    throw std::runtime_error("EI-ZO supports at most 32 configuration variables");
  if (domain.A().rows > 0.9 * 4000)
    throw std::runtime_error(
        "ERROR: The domain polyhedron is already reserving almost all of the faces.");
  float delta_min = (float)(6 * options.delta /
                            (M_PI * M_PI * options.max_iterations * options.max_iterations));
  uint32_t max_cfg = numSamplesUnadaptive(options.epsilon, delta_min, options.tau);
  if (options.num_particles > max_cfg)
    throw std::runtime_error(
        "Current required particle count exceeds buffer size. There were " +
        std::to_string(options.num_particles) + " num_particles requested, and " +
        std::to_string(max_cfg) + " max_num_configurations space provided by the unadaptive test.");
  CudaEdgeInflator inf(plant, robot_geometry_ids, options, domain, 1,
                       std::max(voxels.cols, 1));
  HPolyhedron P = inf.inflateEdge(line_start, line_end, voxels, voxel_radius, options.verbose);
  eizo_info = inf.lastInfo();
  return P.GetMyMinimalHPolyhedron();
}

// ---------------------------------------------------------------------------
// distance field
// ---------------------------------------------------------------------------
DistanceField::DistanceField(const Vector3f &origin, size_t gridSize, float voxelSize)
    : gridSize((int)gridSize), numVoxels(gridSize * gridSize * gridSize) {
  ensure_device();
  check_status(csd_distance_field_create(origin.data(), (int32_t)gridSize, voxelSize, &field_),
               "distance field");
}

DistanceField::~DistanceField() { csd_distance_field_destroy(field_); }

void DistanceField::computeEDT(int, int, int) {
  check_status(csd_distance_field_compute(field_, nullptr), "distance transform");
}

void DistanceField::resetOccupancy() {
  check_status(csd_distance_field_reset_occupancy(field_, nullptr), "reset occupancy");
}

void DistanceField::setOccupancy(const MatrixXf &P) {
  require(P.cols == 0 || P.rows == 3, "points must be a 3 x N matrix");
  if (P.cols == 0) return;
  DevBuf d;
  void *dp = d.get(P.data.size() * 4);
  check_status(csd_memcpy_h2d(dp, P.data.data(), P.data.size() * 4, nullptr), "upload");
  check_status(csd_distance_field_set_occupancy(field_, (const float *)dp, P.cols, nullptr),
               "set occupancy");
  check_status(csd_stream_synchronize(nullptr), "synchronize");
}

std::vector<float> DistanceField::getDistance(const MatrixXf &P) {
  require(P.cols == 0 || P.rows == 3, "points must be a 3 x N matrix");
  std::vector<float> out(P.cols);
  if (P.cols == 0) return out;
  DevBuf dpts, dout;
  void *dp = dpts.get(P.data.size() * 4);
  void *dd = dout.get(out.size() * 4);
  check_status(csd_memcpy_h2d(dp, P.data.data(), P.data.size() * 4, nullptr), "upload");
  check_status(csd_distance_field_get_distance(field_, (const float *)dp, P.cols, (float *)dd,
                                               nullptr),
               "distance lookup");
  check_status(csd_memcpy_d2h(out.data(), dd, out.size() * 4, nullptr), "download");
  check_status(csd_stream_synchronize(nullptr), "synchronize");
  return out;
}

// ---------------------------------------------------------------------------
// region editing
// ---------------------------------------------------------------------------
std::pair<std::vector<HPolyhedron>, std::pair<MatrixXf, MatrixXf>> EditRegionsCuda(
    const MatrixXf &collisions, const std::vector<uint32_t> &line_segment_idxs,
    const MatrixXf &line_start_points, const MatrixXf &line_end_points,
    const std::vector<HPolyhedron> &regions, const MinimalPlant &plant,
    const std::vector<GeometryIndex> &robot_geometry_ids, const MatrixXf &voxels,
    float voxel_radius, const EditRegionsOptions &options) {
  // argument checks of cuda_edit_regions.cu:55-86
  require(!regions.empty(), "Regions vector cannot be empty");
  const int R = (int)regions.size();
  require(line_start_points.cols == R && line_end_points.cols == R,
          "Number of line segments must match number of regions");
  require(line_start_points.rows == line_end_points.rows,
          "Line start and end points must have same dimension");
  const int dim = regions[0].ambient_dimension();
  for (const auto &r : regions)
    require(r.ambient_dimension() == dim, "all regions must have the same dimension");
  require(collisions.rows == dim, "collisions must live in the same space as the regions");
  require(line_start_points.rows == dim,
          "the linesegments must live in the same dimension as the regions");
  require(dim == plant.kin_tree.num_configuration_variables,
          "the regions must live in the plant's configuration space");
  require(!robot_geometry_ids.empty(), "Robot geometry IDs cannot be empty");
  for (auto g : robot_geometry_ids)
    require(g >= 0 && g < plant.num_scene_geometries, "invalid robot geometry id");
  require(voxels.cols == 0 || voxels.rows == 3, "voxels must be a 3 x V matrix");
  require(voxel_radius >= 0.f, "Voxel radius must not be negative");
  auto column = [dim](const MatrixXf &m, int c) {
    return std::vector<float>(m.data.begin() + (size_t)c * dim,
                              m.data.begin() + (size_t)(c + 1) * dim);
  };
  for (int r = 0; r < R; ++r) {
    if (!regions[r].PointInSet(column(line_start_points, r)))
      throw std::runtime_error("line start is not contained in region " + std::to_string(r));
    if (!regions[r].PointInSet(column(line_end_points, r)))
      throw std::runtime_error("line end is not contained in region " + std::to_string(r));
  }
  // group the collisions by segment, keeping their order (cuda_edit_regions.cu:88-113)
  const int C = collisions.cols;
  if ((int)line_segment_idxs.size() < C)
    throw std::out_of_range("[EditRegionsCuda] fewer line segment indices than collisions");
  std::vector<uint32_t> count(R, 0), start(R, 0);
  for (int c = 0; c < C; ++c) {
    if (line_segment_idxs[c] >= (uint32_t)R)
      throw std::out_of_range("[EditRegionsCuda] line segment index " +
                              std::to_string(line_segment_idxs[c]) + " has no region");
    ++count[line_segment_idxs[c]];
  }
  for (int r = 1; r < R; ++r) start[r] = start[r - 1] + count[r - 1];
  if (options.verbose) printf("[EditRegionsCuda] Num collisions to be optimized %d\n", C);
  if (C == 0)
    throw std::runtime_error(
        "[EditRegionsCuda] No collisions are contained in the regions, this is not supported.\n");
  for (int r = 0; r < R; ++r)
    if (count[r] > 4096)
      throw std::invalid_argument("[EditRegionsCuda] region " + std::to_string(r) + " has " +
                                  std::to_string(count[r]) +
                                  " collisions; at most 4096 per region are supported");
  MatrixXf grouped(dim, C);
  std::vector<uint32_t> seg(C);
  {
    std::vector<uint32_t> next(start);
    for (int c = 0; c < C; ++c) {
      const uint32_t s = line_segment_idxs[c], k = next[s]++;
      std::memcpy(&grouped.data[(size_t)k * dim], &collisions.data[(size_t)c * dim],
                  sizeof(float) * dim);
      seg[k] = s;
    }
  }
  // regions packed [R][cap][dim] with room for one face per collision
  int cap = 1;
  for (int r = 0; r < R; ++r) cap = std::max(cap, regions[r].A().rows + (int)count[r]);
  std::vector<float> A((size_t)R * cap * dim, 0.f), b((size_t)R * cap, 0.f);
  std::vector<int32_t> nf(R);
  for (int r = 0; r < R; ++r) {
    const MatrixXf &Ar = regions[r].A();
    nf[r] = Ar.rows;
    for (int f = 0; f < Ar.rows; ++f) {
      for (int d = 0; d < dim; ++d) A[((size_t)r * cap + f) * dim + d] = Ar(f, d);
      b[(size_t)r * cap + f] = regions[r].b()[f];
    }
  }
  csd_plant *p = plant.device();
  std::vector<int32_t> offsets(R + 1);
  for (int r = 0; r < R; ++r) offsets[r] = (int32_t)start[r];
  offsets[R] = C;
  DevBuf d_col, d_seg, d_off, d_ls, d_le, d_A, d_b, d_nf, d_vox, d_sc, d_pr, d_st;
  auto up = [](DevBuf &buf, const void *src, size_t bytes) {
    void *dst = buf.get(std::max<size_t>(bytes, 4));
    if (bytes) check_status(csd_memcpy_h2d(dst, src, bytes, nullptr), "upload");
    return dst;
  };
  const size_t cbytes = sizeof(float) * (size_t)C * dim;
  up(d_col, grouped.data.data(), cbytes);
  up(d_seg, seg.data(), sizeof(uint32_t) * C);
  up(d_off, offsets.data(), sizeof(int32_t) * (R + 1));
  up(d_ls, line_start_points.data.data(), sizeof(float) * (size_t)R * dim);
  up(d_le, line_end_points.data.data(), sizeof(float) * (size_t)R * dim);
  up(d_A, A.data(), sizeof(float) * A.size());
  up(d_b, b.data(), sizeof(float) * b.size());
  up(d_nf, nf.data(), sizeof(int32_t) * R);
  if (voxels.cols) up(d_vox, voxels.data.data(), sizeof(float) * voxels.data.size());
  d_sc.get(cbytes);
  d_pr.get(cbytes);
  d_st.get(sizeof(int32_t) * R);
  check_status(csd_edit_regions(p, (const float *)d_col.ptr, (const uint32_t *)d_seg.ptr,
                                (const int32_t *)d_off.ptr, C,
                                (const float *)d_ls.ptr, (const float *)d_le.ptr, R,
                                (float *)d_A.ptr, (float *)d_b.ptr, (int32_t *)d_nf.ptr, cap,
                                voxels.cols ? (const float *)d_vox.ptr : nullptr, voxels.cols,
                                voxel_radius, robot_geometry_ids.data(),
                                (int)robot_geometry_ids.size(), options.configuration_margin,
                                (int32_t)options.bisection_steps, (float *)d_sc.ptr,
                                (float *)d_pr.ptr, (int32_t *)d_st.ptr, nullptr),
               "region editing");
  std::vector<int32_t> status(R);
  MatrixXf projections(dim, C), optimized(dim, C);
  check_status(csd_memcpy_d2h(A.data(), d_A.ptr, sizeof(float) * A.size(), nullptr), "download");
  check_status(csd_memcpy_d2h(b.data(), d_b.ptr, sizeof(float) * b.size(), nullptr), "download");
  check_status(csd_memcpy_d2h(nf.data(), d_nf.ptr, sizeof(int32_t) * R, nullptr), "download");
  check_status(csd_memcpy_d2h(status.data(), d_st.ptr, sizeof(int32_t) * R, nullptr), "download");
  check_status(csd_memcpy_d2h(projections.data.data(), d_pr.ptr, cbytes, nullptr), "download");
  check_status(csd_memcpy_d2h(optimized.data.data(), d_sc.ptr, cbytes, nullptr), "download");
  check_status(csd_stream_synchronize(nullptr), "synchronize");
  std::vector<HPolyhedron> edited;
  edited.reserve(R);
  for (int r = 0; r < R; ++r) {
    if (status[r] < 0)
      throw std::runtime_error("[EditRegionsCuda] region " + std::to_string(r) +
                               " could not be edited (status " + std::to_string(status[r]) + ")");
    if (count[r] == 0) {
      if (options.verbose) printf("[EditRegionsCuda] region %d remains unchanged\n", r);
      edited.push_back(regions[r]);
      continue;
    }
    if (options.verbose) printf("[EditRegionsCuda] region %d added %d faces\n", r, status[r]);
    MatrixXf Ar(nf[r], dim);
    std::vector<float> br(nf[r]);
    for (int f = 0; f < nf[r]; ++f) {
      for (int d = 0; d < dim; ++d) Ar(f, d) = A[((size_t)r * cap + f) * dim + d];
      br[f] = b[(size_t)r * cap + f];
    }
    edited.emplace_back(Ar, br);
  }
  return {std::move(edited), {std::move(projections), std::move(optimized)}};
}

}  // namespace csdecomp
