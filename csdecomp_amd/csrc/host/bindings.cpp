// bindings.cpp -- pybind11 module with the Python surface of csdecomp's
// pycsdecomp_bindings (src/pybind/csdecomp/{plant,hpolyhedron,
// collision_checker,cuda}_bindings.h): same class, method and argument names.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>

#include "csdecomp.h"

namespace py = pybind11;
using namespace csdecomp;

namespace {
using FArr = py::array_t<float, py::array::f_style | py::array::forcecast>;

MatrixXf to_mat(const py::object &a_in) {
  FArr a = FArr::ensure(a_in);
  if (!a) throw std::invalid_argument("expected a float array");
  MatrixXf m;
  if (a.ndim() == 1) {
    m = MatrixXf((int)a.shape(0), 1);
  } else if (a.ndim() == 2) {
    m = MatrixXf((int)a.shape(0), (int)a.shape(1));
  } else {
    throw std::invalid_argument("expected a 1-D or 2-D array");
  }
  if (!m.data.empty()) std::memcpy(m.data.data(), a.data(), m.data.size() * sizeof(float));
  return m;
}

py::array_t<float> from_mat(const MatrixXf &m) {
  py::array_t<float, py::array::f_style> a({m.rows, m.cols});
  if (!m.data.empty()) std::memcpy(a.mutable_data(), m.data.data(), m.data.size() * sizeof(float));
  return a;
}

std::vector<float> to_vec(const py::object &a_in) {
  py::array_t<float, py::array::c_style | py::array::forcecast> a =
      py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(a_in);
  if (!a) throw std::invalid_argument("expected a float vector");
  return std::vector<float>(a.data(), a.data() + a.size());
}

py::array_t<float> from_vec(const std::vector<float> &v) {
  py::array_t<float> a((py::ssize_t)v.size());
  if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(float));
  return a;
}

Matrix4f to_m4(const py::object &a_in) {
  MatrixXf m = to_mat(a_in);
  if (m.rows != 4 || m.cols != 4) throw std::invalid_argument("expected a 4x4 matrix");
  Matrix4f r;
  std::memcpy(r.data(), m.data.data(), 64);
  return r;
}

py::array_t<float> from_m4(const Matrix4f &x) {
  MatrixXf m(4, 4);
  std::memcpy(m.data.data(), x.data(), 64);
  return from_mat(m);
}

Vector3f to_v3(const py::object &a) {
  std::vector<float> v = to_vec(a);
  if (v.size() != 3) throw std::invalid_argument("expected 3 values");
  return {v[0], v[1], v[2]};
}

struct Voxels {
  MatrixXf matrix{3, 0};
};

MatrixXf voxels_of(const py::object &o) {
  if (o.is_none()) return MatrixXf(3, 0);
  if (py::isinstance<Voxels>(o)) return o.cast<const Voxels &>().matrix;
  return to_mat(o);
}
}  // namespace

PYBIND11_MODULE(_pycsdecomp, m) {
  m.doc() = "MI355X (gfx950) implementation of csdecomp's FK / collision / EI-ZO path";

  // ---------------------------------------------------------------- plant
  py::enum_<JointType>(m, "JointType")
      .value("FIXED", FIXED)
      .value("REVOLUTE", REVOLUTE)
      .value("PRISMATIC", PRISMATIC)
      .export_values();

  py::class_<Joint>(m, "Joint")
      .def(py::init<>())
      .def_readwrite("name", &Joint::name)
      .def_readwrite("type", &Joint::type)
      .def_property(
          "X_PL_J", [](const Joint &j) { return from_m4(j.X_PL_J); },
          [](Joint &j, const py::object &a) { j.X_PL_J = to_m4(a); })
      .def_readwrite("parent_link", &Joint::parent_link)
      .def_readwrite("child_link", &Joint::child_link)
      .def_property(
          "axis", [](const Joint &j) { return from_vec({j.axis[0], j.axis[1], j.axis[2]}); },
          [](Joint &j, const py::object &a) { j.axis = to_v3(a); })
      .def_readwrite("position_lower_limit", &Joint::position_lower_limit)
      .def_readwrite("position_upper_limit", &Joint::position_upper_limit);

  py::class_<Link>(m, "Link")
      .def(py::init<>())
      .def_readwrite("name", &Link::name)
      .def_readwrite("parent_joint", &Link::parent_joint)
      .def_readwrite("child_joints", &Link::child_joints);

  py::class_<KinematicTree>(m, "KinematicTree")
      .def(py::init<>())
      .def("add_link", &KinematicTree::addLink)
      .def("add_joint", &KinematicTree::addJoint)
      .def("finalize", &KinematicTree::finalize)
      .def("get_link_index", &KinematicTree::getLinkIndexByName)
      .def("get_link_name", &KinematicTree::getLinkNameByIndex)
      .def("get_link", &KinematicTree::getLink)
      .def("get_joint_index", &KinematicTree::getJointIndex)
      .def("get_joint", &KinematicTree::getJoint)
      .def("is_finalized", &KinematicTree::isFinalized)
      .def("get_leaf_link_indices", &KinematicTree::getLeafLinkIndices)
      .def("get_links", &KinematicTree::getLinks)
      .def("get_joints", &KinematicTree::getJoints)
      .def("get_position_lower_limits",
           [](const KinematicTree &t) { return from_vec(t.getPositionLowerLimits()); })
      .def("get_position_upper_limits",
           [](const KinematicTree &t) { return from_vec(t.getPositionUpperLimits()); })
      .def("compute_link_frame_to_world_transforms",
           [](const KinematicTree &t, const py::object &q) {
             std::vector<py::array_t<float>> out;
             for (auto &x : t.computeLinkFrameToWorldTransforms(to_vec(q))) out.push_back(from_m4(x));
             return out;
           })
      .def("get_minimal_kinematic_tree", &KinematicTree::getMyMinimalKinematicTreeStruct)
      .def("num_links", &KinematicTree::numLinks)
      .def("num_joints", &KinematicTree::numJoints)
      .def("num_configuration_variables", &KinematicTree::numConfigurationVariables);

  py::class_<MinimalJoint>(m, "MinimalJoint")
      .def(py::init<>())
      .def_readwrite("type", &MinimalJoint::type)
      .def_property(
          "X_PL_J", [](const MinimalJoint &j) { return from_m4(j.X_PL_J); },
          [](MinimalJoint &j, const py::object &a) { j.X_PL_J = to_m4(a); })
      .def_readwrite("index", &MinimalJoint::index)
      .def_readwrite("parent_link", &MinimalJoint::parent_link)
      .def_readwrite("child_link", &MinimalJoint::child_link)
      .def_property(
          "axis", [](const MinimalJoint &j) { return from_vec({j.axis[0], j.axis[1], j.axis[2]}); },
          [](MinimalJoint &j, const py::object &a) { j.axis = to_v3(a); });

  py::class_<MinimalLink>(m, "MinimalLink")
      .def(py::init<>())
      .def_readwrite("index", &MinimalLink::index)
      .def_readwrite("parent_joint_index", &MinimalLink::parent_joint_index)
      .def_readwrite("child_joint_index_array", &MinimalLink::child_joint_index_array)
      .def_readwrite("num_child_joints", &MinimalLink::num_child_joints);

  py::class_<MinimalKinematicTree>(m, "MinimalKinematicTree")
      .def(py::init<>())
      .def_readwrite("joints", &MinimalKinematicTree::joints)
      .def_readwrite("links", &MinimalKinematicTree::links)
      .def_readwrite("joint_traversal_sequence", &MinimalKinematicTree::joint_traversal_sequence)
      .def_readwrite("joint_idx_to_config_idx", &MinimalKinematicTree::joint_idx_to_config_idx)
      .def_readwrite("num_configuration_variables",
                     &MinimalKinematicTree::num_configuration_variables)
      .def_readwrite("num_joints", &MinimalKinematicTree::num_joints)
      .def_readwrite("num_links", &MinimalKinematicTree::num_links);

  m.def(
      "computeLinkFrameToWorldTransformsMinimal",
      [](const py::object &q, const MinimalKinematicTree &tree) {
        auto tfs = computeLinkFrameToWorldTransformsMinimal(to_vec(q), tree);
        MatrixXf out(4 * tree.num_links, 4);  // link l at rows 4l..4l+3
        for (int l = 0; l < tree.num_links; ++l)
          for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) out(4 * l + r, c) = tfs[l][c * 4 + r];
        return from_mat(out);
      },
      "Compute link frame to world transforms for a minimal kinematic tree",
      py::arg("joint_values"), py::arg("tree"));

  py::enum_<ShapePrimitive>(m, "ShapePrimitive")
      .value("BOX", BOX)
      .value("SPHERE", SPHERE)
      .value("CYLINDER", CYLINDER)
      .value("CAPSULE", CAPSULE);
  py::enum_<CollisionGroup>(m, "CollisionGroup")
      .value("ROBOT", ROBOT)
      .value("WORLD", WORLD)
      .value("VOXELS", VOXELS);

  py::class_<CollisionGeometry>(m, "CollisionGeometry")
      .def(py::init<>())
      .def_readwrite("type", &CollisionGeometry::type)
      .def_property(
          "dimensions",
          [](const CollisionGeometry &g) {
            return from_vec({g.dimensions[0], g.dimensions[1], g.dimensions[2]});
          },
          [](CollisionGeometry &g, const py::object &a) { g.dimensions = to_v3(a); })
      .def_readwrite("link_index", &CollisionGeometry::link_index)
      .def_readwrite("group", &CollisionGeometry::group)
      .def_property(
          "X_L_B", [](const CollisionGeometry &g) { return from_m4(g.X_L_B); },
          [](CollisionGeometry &g, const py::object &a) { g.X_L_B = to_m4(a); });

  py::class_<SceneInspector>(m, "SceneInspector")
      .def(py::init<>())
      .def_readwrite("link_index_to_scene_collision_geometry_indices",
                     &SceneInspector::link_index_to_scene_collision_geometry_indices)
      .def_readwrite("scene_collision_geometry_name_to_geometry_index",
                     &SceneInspector::scene_collision_geometry_name_to_geometry_index)
      .def_readwrite("robot_to_world_collision_pairs",
                     &SceneInspector::robot_to_world_collision_pairs)
      .def_readwrite("robot_geometry_ids", &SceneInspector::robot_geometry_ids)
      .def("get_collision_pair_matrix", [](const SceneInspector &s) {
        std::vector<int32_t> v = s.getCollisionPairMatrix();
        py::array_t<int32_t, py::array::f_style> a({(py::ssize_t)2, (py::ssize_t)(v.size() / 2)});
        if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * 4);
        return a;
      });

  m.def(
      "finalize_scene",
      [](const KinematicTree &tree, std::vector<CollisionGeometry> geoms,
         const std::unordered_map<std::string, int> &names) {
        SceneInspector s = finalizeScene(tree, geoms, names);
        return py::make_tuple(geoms, s);
      },
      py::arg("tree"), py::arg("geometries"), py::arg("names"),
      "Collision groups and pairs of a scene (URDFParser::finalizeScene rule)");

  py::class_<Voxels>(m, "Voxels")
      .def(py::init<>())
      .def(py::init([](const py::object &a) {
        Voxels v;
        v.matrix = to_mat(a);
        if (v.matrix.cols > 0 && v.matrix.rows != 3)
          throw std::invalid_argument("voxels must be a 3 x V matrix");
        return v;
      }))
      .def("set_matrix",
           [](Voxels &v, const py::object &a) {
             v.matrix = to_mat(a);
             if (v.matrix.cols > 0 && v.matrix.rows != 3)
               throw std::invalid_argument("voxels must be a 3 x V matrix");
           })
      .def("get_matrix", [](const Voxels &v) { return from_mat(v.matrix); });

  py::class_<MinimalPlant>(m, "MinimalPlant")
      .def(py::init<>())
      .def_readwrite("kin_tree", &MinimalPlant::kin_tree)
      .def_property_readonly("scene_collision_geometries",
                             [](const MinimalPlant &p) {
                               return std::vector<CollisionGeometry>(
                                   p.scene_collision_geometries.begin(),
                                   p.scene_collision_geometries.begin() + p.num_scene_geometries);
                             })
      .def("set_scene_collision_geometries",
           [](MinimalPlant &p, const std::vector<CollisionGeometry> &g) {
             p.scene_collision_geometries = g;
             p.num_scene_geometries = (int)g.size();
           })
      .def("set_collision_pairs_from_matrix",
           [](MinimalPlant &p, const py::object &mat) {
             auto a = py::array_t<int32_t, py::array::forcecast>::ensure(mat);
             if (!a || a.ndim() != 2 || a.shape(0) != 2)
               throw std::runtime_error("Matrix must be shape (2, N)");
             auto u = a.unchecked<2>();
             const py::ssize_t n = a.shape(1);
             p.collision_pairs_flat.resize(2 * n);
             for (py::ssize_t i = 0; i < n; ++i) {
               p.collision_pairs_flat[2 * i] = u(0, i);
               p.collision_pairs_flat[2 * i + 1] = u(1, i);
             }
             p.num_collision_pairs = (int)n;
           })
      .def_property_readonly("collision_pairs_flat",
                             [](const MinimalPlant &p) {
                               py::array_t<int32_t> a(2 * (py::ssize_t)p.num_collision_pairs);
                               if (p.num_collision_pairs)
                                 std::memcpy(a.mutable_data(), p.collision_pairs_flat.data(),
                                             8 * (size_t)p.num_collision_pairs);
                               return a;
                             })
      .def_readwrite("num_scene_geometries", &MinimalPlant::num_scene_geometries)
      .def_readwrite("num_collision_pairs", &MinimalPlant::num_collision_pairs)
      .def("device_info", [](const MinimalPlant &p) {
        int32_t L, dof, P, fast;
        check_status(csd_plant_info(p.device(), &L, &dof, &P, &fast), "plant info");
        py::dict d;
        d["num_links"] = L;
        d["num_configuration_variables"] = dof;
        d["num_device_pairs"] = P;
        d["fast_path"] = fast != 0;
        return d;
      });

  py::class_<Plant>(m, "Plant")
      .def(py::init<const KinematicTree &, const SceneInspector &,
                    const std::vector<CollisionGeometry> &>())
      .def("numLinks", &Plant::numLinks)
      .def("numJoints", &Plant::numJoints)
      .def("getPositionLowerLimits", [](const Plant &p) { return from_vec(p.getPositionLowerLimits()); })
      .def("getPositionUpperLimits", [](const Plant &p) { return from_vec(p.getPositionUpperLimits()); })
      .def("numConfigurationVariables", &Plant::numConfigurationVariables)
      .def("getMinimalPlant", &Plant::getMinimalPlant)
      .def("getKinematicTree", &Plant::getKinematicTree, py::return_value_policy::reference_internal)
      .def("getSceneInspector", &Plant::getSceneInspector,
           py::return_value_policy::reference_internal)
      .def("getRobotGeometryIds", &Plant::getRobotGeometryIds)
      .def("getSceneCollisionGeometries", &Plant::getSceneCollisionGeometries)
      .def("getCollisionPairMatrix",
           [](const Plant &p) {
             std::vector<int32_t> v = p.getCollisionPairMatrix();
             py::array_t<int32_t, py::array::f_style> a({(py::ssize_t)2, (py::ssize_t)(v.size() / 2)});
             if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * 4);
             return a;
           })
      .def("getSceneCollisionGeometryNames", &Plant::getSceneCollisionGeometryNames)
      .def("getSceneCollisionGeometryIndexByName", &Plant::getSceneCollisionGeometryIndexByName);

  // ---------------------------------------------------------------- polytopes
  py::class_<MinimalHPolyhedron>(m, "MinimalHPolyhedron")
      .def(py::init<>())
      .def("A",
           [](const MinimalHPolyhedron &s) {
             MatrixXf A(s.num_faces, s.dim);
             for (int r = 0; r < s.num_faces; ++r)
               for (int c = 0; c < s.dim; ++c) A(r, c) = s.A[(size_t)r * s.dim + c];
             return from_mat(A);
           })
      .def("b", [](const MinimalHPolyhedron &s) {
        return from_vec(std::vector<float>(s.b.begin(), s.b.begin() + s.num_faces));
      })
      .def_property_readonly("num_faces", [](const MinimalHPolyhedron &s) { return s.num_faces; })
      .def_property_readonly("dim", [](const MinimalHPolyhedron &s) { return s.dim; });

  py::class_<HPolyhedron>(m, "HPolyhedron")
      .def(py::init<>())
      .def(py::init<const MinimalHPolyhedron &>())
      .def(py::init([](const py::object &A, const py::object &b) {
             return HPolyhedron(to_mat(A), to_vec(b));
           }),
           py::arg("A"), py::arg("b"))
      .def("A", [](const HPolyhedron &h) { return from_mat(h.A()); })
      .def("b", [](const HPolyhedron &h) { return from_vec(h.b()); })
      .def("MakeBox", [](HPolyhedron &h, const py::object &lo,
                         const py::object &hi) { h.MakeBox(to_vec(lo), to_vec(hi)); })
      .def("GetMyMinimalHPolyhedron", &HPolyhedron::GetMyMinimalHPolyhedron)
      .def(
          "PointInSet",
          [](const HPolyhedron &h, const py::object &x, float tol) {
            return h.PointInSet(to_vec(x), tol);
          },
          py::arg("previous_sample"), py::arg("tol") = 1e-9f)
      .def(
          "UniformSample",
          [](const HPolyhedron &h, const py::object &x, int mixing) {
            return from_vec(h.UniformSample(to_vec(x), mixing));
          },
          py::arg("previous_sample"), py::arg("mixing_steps"))
      .def("GetFeasiblePoint", [](const HPolyhedron &h) { return from_vec(h.GetFeasiblePoint()); })
      .def("ChebyshevCenter", [](const HPolyhedron &h) { return from_vec(h.ChebyshevCenter()); })
      .def("ambient_dimension", &HPolyhedron::ambient_dimension);

  // ---------------------------------------------------------------- checker
  m.def(
      "PairCollisionFree",
      [](const CollisionGeometry &a, const py::object &xa, const CollisionGeometry &b,
         const py::object &xb) { return pairCollisionFree(a, to_m4(xa), b, to_m4(xb)); },
      py::arg("geom_a"), py::arg("x_w_la"), py::arg("geom_b"), py::arg("x_w_lb"));
  m.def(
      "CheckCollisionFree",
      [](const py::object &q, const MinimalPlant &p) { return checkCollisionFree(to_vec(q), p); },
      py::arg("configuration"), py::arg("plant"));
  m.def(
      "CheckCollisionFreeVoxels",
      [](const py::object &q, const py::object &vox, float r, const MinimalPlant &p,
         const std::vector<int> &ids) {
        for (int id : ids)
          if (id < 0 || id >= p.num_scene_geometries)
            throw std::runtime_error(
                "Invalid robot geometry ID, check the SceneInspector for vaild ids.");
        return checkCollisionFreeVoxels(to_vec(q), voxels_of(vox), r, p,
                                        std::vector<GeometryIndex>(ids.begin(), ids.end()));
      },
      py::arg("configuration"), py::arg("voxels"), py::arg("voxel_radius"), py::arg("plant"),
      py::arg("robot_geometry_index"));
  m.def(
      "CheckEdgeCollisionFree",
      [](const py::object &a, const py::object &b, const MinimalPlant &p, float step) {
        return checkEdgeCollisionFree(to_vec(a), to_vec(b), p, step);
      },
      py::arg("configuration_1"), py::arg("configuration_2"), py::arg("plant"),
      py::arg("step_size") = 0.01f);

  // ---------------------------------------------------------------- cuda names
  py::class_<EizoOptions>(m, "EizoOptions")
      .def(py::init<>())
      .def_readwrite("configuration_margin", &EizoOptions::configuration_margin)
      .def_readwrite("tau", &EizoOptions::tau)
      .def_readwrite("delta", &EizoOptions::delta)
      .def_readwrite("epsilon", &EizoOptions::epsilon)
      .def_readwrite("bisection_steps", &EizoOptions::bisection_steps)
      .def_readwrite("num_particles", &EizoOptions::num_particles)
      .def_readwrite("max_hyperplanes_per_iteration", &EizoOptions::max_hyperplanes_per_iteration)
      .def_readwrite("max_iterations", &EizoOptions::max_iterations)
      .def_readwrite("mixing_steps", &EizoOptions::mixing_steps)
      .def_readwrite("seed", &EizoOptions::seed)
      .def_readwrite("verbose", &EizoOptions::verbose)
      .def_readwrite("track_iteration_information", &EizoOptions::track_iteration_information);

  py::class_<EizoIterationInfo>(m, "EizoIterationInfo")
      .def_readonly("iteration_index", &EizoIterationInfo::iteration_index)
      .def_readonly("num_samples_drawn", &EizoIterationInfo::num_samples_drawn)
      .def_readonly("collisions_threshold", &EizoIterationInfo::collisions_threshold)
      .def_readonly("num_collisions_unadaptive", &EizoIterationInfo::num_collisions_unadaptive)
      .def_readonly("num_collisions_updated", &EizoIterationInfo::num_collisions_updated)
      .def_readonly("num_faces_placed", &EizoIterationInfo::num_faces_placed)
      .def_readonly("iteration_time_ms", &EizoIterationInfo::iteration_time_ms)
      .def_readonly("num_collisions_found", &EizoIterationInfo::num_collisions_found)
      .def_readonly("num_faces_at_start", &EizoIterationInfo::num_faces_at_start)
      .def_readonly("num_pilot_samples", &EizoIterationInfo::num_pilot_samples)
      .def_readonly("num_pilot_collisions", &EizoIterationInfo::num_pilot_collisions)
      .def_readonly("num_pilot_checked", &EizoIterationInfo::num_pilot_checked);
  m.def("printEizoSummary", &printEizoSummary, py::arg("info"));

  py::class_<EizoInfo>(m, "EizoInfo")
      .def(py::init<>())
      .def_readonly("iterations", &EizoInfo::iterations)
      .def_readonly("converged", &EizoInfo::converged)
      .def_readonly("face_capacity_reached", &EizoInfo::face_capacity_reached)
      .def_readonly("plain_fallback", &EizoInfo::plain_fallback)
      .def_readonly("samples_drawn", &EizoInfo::samples_drawn)
      .def_readonly("faces_added", &EizoInfo::faces_added)
      .def_readonly("iteration_info", &EizoInfo::iteration_info);

  py::class_<CudaEdgeInflator>(m, "CudaEdgeInflator")
      .def(py::init<const MinimalPlant &, const std::vector<GeometryIndex> &, const EizoOptions &,
                    const HPolyhedron &, int, int, int>(),
           py::arg("plant"), py::arg("robot_geometry_ids"), py::arg("options"), py::arg("domain"),
           py::arg("max_regions") = 1, py::arg("max_voxels") = 10000,
           py::arg("face_capacity") = 0)
      .def(
          "inflateEdge",
          [](CudaEdgeInflator &self, const py::object &ls, const py::object &le,
             const py::object &vox, float r, bool verbose) {
            return self.inflateEdge(to_vec(ls), to_vec(le), voxels_of(vox), r, verbose);
          },
          py::arg("line_start"), py::arg("line_end"), py::arg("voxels"), py::arg("voxel_radius"),
          py::arg("verbose") = true)
      .def(
          "inflateEdges",
          [](CudaEdgeInflator &self, const py::object &ls, const py::object &le,
             const py::object &vox, float r, const std::vector<uint32_t> &budgets) {
            std::vector<EizoInfo> infos;
            auto regions = self.inflateEdges(to_mat(ls), to_mat(le), voxels_of(vox), r, budgets,
                                             &infos);
            return py::make_tuple(regions, infos);
          },
          py::arg("line_starts"), py::arg("line_ends"), py::arg("voxels"),
          py::arg("voxel_radius"), py::arg("particle_budgets") = std::vector<uint32_t>())
      .def(
          "set_guidance",
          [](CudaEdgeInflator &self, uint32_t pilot_samples, float pilot_factor,
             const py::object &surrogate, int keep_particles, int line_points,
             bool plain_fallback, const py::object &prune_below) {
            uintptr_t addr = 0;
            if (py::hasattr(surrogate, "address")) addr = surrogate.attr("address").cast<uintptr_t>();
            else if (!surrogate.is_none()) addr = surrogate.cast<uintptr_t>();
            self.setGuidance(pilot_samples, pilot_factor, reinterpret_cast<const void *>(addr),
                             keep_particles, line_points, plain_fallback,
                             prune_below.is_none() ? NAN : prune_below.cast<float>());
          },
          py::arg("pilot_samples") = 0, py::arg("pilot_factor") = 10.0f,
          py::arg("surrogate") = py::none(), py::arg("keep_particles") = 256,
          py::arg("line_points") = 4, py::arg("plain_fallback") = false,
          py::arg("prune_below") = py::none(),
          // the inflator keeps the surrogate object alive: it reads the model
          // in every guided round
          py::keep_alive<1, 4>(),
          "surrogate: a csdecomp_amd.capi.Mlp (kept alive by the inflator), the address of a "
          "csd_mlp, or None; prune_below: with a pilot and a surrogate, only the pilot samples "
          "the surrogate predicts below this value get the exact check (None: all)")
      .def(
          "set_metric",
          [](CudaEdgeInflator &self, const py::object &m) {
            self.setMetric(m.is_none() ? MatrixXf() : to_mat(m));
          },
          py::arg("metric") = py::none(),
          "ellipsoid metric of the face placement: a symmetric positive definite dim x dim "
          "matrix (faces tangent to M-ellipsoids around the segment), None for the Euclidean "
          "metric of the reference")
      .def("set_sampler_tile", &CudaEdgeInflator::setSamplerTile, py::arg("wide_from"),
           "hit-and-run in 1024-thread workgroups (16 waves sharing the polytope in LDS) from "
           "wide_from faces on when 256-thread workgroups also fit (0: always; 2**31 - 1: only "
           "above their 64 KB); the samples are the same either way")
      .def("tune_sampler", &CudaEdgeInflator::tuneSampler,
           "time both sampler tiles on synthetic polytopes and keep the faster per face count; "
           "returns the face count from which 1024-thread workgroups are used")
      .def("set_graph", &CudaEdgeInflator::setGraph, py::arg("rounds_per_sync") = 1,
           "replay one graph-captured round per launch; the host reads the region status "
           "every rounds_per_sync rounds (0: host-driven rounds)")
      .def("last_info", &CudaEdgeInflator::lastInfo)
      .def(
          "last_particles",
          [](const CudaEdgeInflator &self, int region) {
            EizoParticles p = self.lastParticles(region);
            return py::make_tuple(from_mat(p.particles), from_mat(p.scol), from_mat(p.proj));
          },
          py::arg("region") = 0,
          "(particles, scol, proj), each dim x count: region's particles in its last round of "
          "the last inflateEdge(s) call, their bisected boundary points and projections")
      .def("face_capacity", &CudaEdgeInflator::faceCapacity);

  m.def(
      "EizoCuda",
      [](const py::object &ls, const py::object &le, const py::object &vox, float r,
         const HPolyhedron &domain, const MinimalPlant &plant, const std::vector<GeometryIndex> &ids,
         const EizoOptions &opt) {
        EizoInfo info;
        MinimalHPolyhedron P =
            EizoCuda(to_vec(ls), to_vec(le), voxels_of(vox), r, domain, plant, ids, opt, info);
        return py::make_tuple(P, info);
      },
      py::arg("line_start"), py::arg("line_end"), py::arg("voxels"), py::arg("voxel_radius"),
      py::arg("domain"), py::arg("plant"), py::arg("robot_geometry_ids"), py::arg("options"));

  m.def(
      "UniformSampleInHPolyhedraCuda",
      [](const std::vector<HPolyhedron> &polys, const py::object &starts, int n, int mixing,
         uint64_t seed) {
        std::vector<py::array_t<float>> out;
        for (auto &s : UniformSampleInHPolyhedronCuda(polys, to_mat(starts), n, mixing, seed))
          out.push_back(from_mat(s));
        return out;
      },
      py::arg("polyhedra"), py::arg("starting_points"), py::arg("num_samples_per_hpolyhedron"),
      py::arg("mixing_steps"), py::arg("seed") = 1337);

  py::class_<DistanceField>(m, "DistanceField")
      .def(py::init([](const py::object &origin, size_t grid_size, float voxel_size) {
             return new DistanceField(to_v3(origin), grid_size, voxel_size);
           }),
           py::arg("origin"), py::arg("grid_size"), py::arg("voxel_size"))
      .def("computeEDT", &DistanceField::computeEDT, py::arg("phase1Band") = 1,
           py::arg("phase2Band") = 1, py::arg("phase3Band") = 2)
      .def("setOccupancy",
           [](DistanceField &f, const py::object &P) { f.setOccupancy(to_mat(P)); },
           py::arg("P_VG"))
      .def("resetOccupancy", &DistanceField::resetOccupancy)
      .def("getDistance",
           [](DistanceField &f, const py::object &P) { return from_vec(f.getDistance(to_mat(P))); },
           py::arg("P_VG"))
      .def_readonly("gridSize", &DistanceField::gridSize)
      .def_readonly("numVoxels", &DistanceField::numVoxels);

  py::class_<EditRegionsOptions>(m, "EditRegionsOptions")
      .def(py::init<>())
      .def_readwrite("configuration_margin", &EditRegionsOptions::configuration_margin)
      .def_readwrite("bisection_steps", &EditRegionsOptions::bisection_steps)
      .def_readwrite("verbose", &EditRegionsOptions::verbose);

  m.def(
      "EditRegionsCuda",
      [](const py::object &collisions, const std::vector<uint32_t> &line_segment_idxs,
         const py::object &line_start_points, const py::object &line_end_points,
         const std::vector<HPolyhedron> &regions, const MinimalPlant &plant,
         const std::vector<GeometryIndex> &robot_geometry_ids, const py::object &voxels,
         float voxel_radius, const EditRegionsOptions &options) {
        auto res = EditRegionsCuda(to_mat(collisions), line_segment_idxs,
                                   to_mat(line_start_points), to_mat(line_end_points), regions,
                                   plant, robot_geometry_ids, voxels_of(voxels), voxel_radius,
                                   options);
        return py::make_tuple(res.first, py::make_tuple(from_mat(res.second.first),
                                                        from_mat(res.second.second)));
      },
      py::arg("collisions"), py::arg("line_segment_idxs"), py::arg("line_start_points"),
      py::arg("line_end_points"), py::arg("regions"), py::arg("plant"),
      py::arg("robot_geometry_ids"), py::arg("voxels"), py::arg("voxel_radius"),
      py::arg("options"),
      "Cuts colliding configurations out of regions while keeping each region's line segment "
      "inside it. Returns (edited regions, (projections, optimized collisions)), the last two "
      "(dim, C) with the collisions grouped by line segment.");

  m.def(
      "CheckCollisionFreeVoxelsCuda",
      [](const py::object &q, const py::object &vox, float r, const MinimalPlant &p,
         const std::vector<GeometryIndex> &ids) {
        return checkCollisionFreeVoxelsCuda(to_mat(q), voxels_of(vox), r, p, ids);
      },
      py::arg("configurations"), py::arg("voxels"), py::arg("voxel_radius"), py::arg("plant"),
      py::arg("robot_geometries"));

  m.def(
      "CheckCollisionFreeCuda",
      [](const py::object &q, const MinimalPlant &p) { return checkCollisionFreeCuda(to_mat(q), p); },
      py::arg("configurations"), py::arg("plant"));

  m.def(
      "CheckCollisionFreeVoxelsWithoutSelfCollisionCuda",
      [](const py::object &q, const py::object &vox, float r, const MinimalPlant &p,
         const std::vector<GeometryIndex> &ids) {
        MatrixXf qm = to_mat(q), vm = voxels_of(vox);
        std::vector<uint8_t> f = checkCollisionFreeVoxelsWithoutSelfCollisionCuda(qm, vm, r, p, ids);
        py::array_t<uint8_t> a({(py::ssize_t)qm.cols, (py::ssize_t)vm.cols});
        if (!f.empty()) std::memcpy(a.mutable_data(), f.data(), f.size());
        return a;
      },
      py::arg("configurations"), py::arg("voxels"), py::arg("voxel_radius"), py::arg("plant"),
      py::arg("robot_geometries"));

  m.def(
      "computeForwardKinematicsCuda",
      [](const py::object &q, const MinimalKinematicTree &t) {
        return from_mat(computeForwardKinematicsCuda(to_mat(q), t));
      },
      py::arg("configurations"), py::arg("tree"));

  m.def(
      "_visibility_graph_edges",
      [](const py::object &q, const MinimalPlant &p, float step, int64_t batch) {
        auto e = VisibilityGraph(to_mat(q), p, step, batch);
        py::array_t<int32_t> rows((py::ssize_t)e.size()), cols((py::ssize_t)e.size());
        for (size_t k = 0; k < e.size(); ++k) {
          rows.mutable_data()[k] = e[k].first;
          cols.mutable_data()[k] = e[k].second;
        }
        return py::make_tuple(rows, cols);
      },
      py::arg("configurations"), py::arg("plant"), py::arg("step_size"),
      py::arg("max_configs_per_batch") = 10000000);

  m.def("numSamplesUnadaptive", &numSamplesUnadaptive, py::arg("epsilon"), py::arg("delta"),
        py::arg("tau"));
}
