"""CPU-side checks of the drop-in boundary: the C-ABI library loads and
exports every entry point the public header declares; host logic runs
without a device; no compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, parse

HEADER = os.path.join(ROOT, "include", "csdecomp_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(csd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    import csdecomp_amd
    lib = ctypes.CDLL(csdecomp_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_without_device():
    from csdecomp_amd import capi
    l = capi.lib()
    assert l.csd_status_string(0) == b"ok"
    assert l.csd_status_string(2) == b"capacity exceeded"


def test_pybind_surface_names():
    import csdecomp_amd as csd
    for name in ["EizoOptions", "CudaEdgeInflator", "UniformSampleInHPolyhedraCuda",
                 "CheckCollisionFreeVoxelsCuda", "CheckCollisionFreeCuda", "VisibilityGraph",
                 "computeLinkFrameToWorldTransformsMinimal", "MinimalPlant", "CollisionGeometry",
                 "ShapePrimitive", "CollisionGroup", "SceneInspector", "Voxels", "Plant",
                 "HPolyhedron", "MinimalHPolyhedron", "CheckCollisionFree", "CheckCollisionFreeVoxels",
                 "PairCollisionFree", "KinematicTree", "Joint", "Link", "MinimalKinematicTree",
                 "MinimalJoint", "MinimalLink", "JointType", "URDFParser"]:
        assert hasattr(csd, name), name
    o = csd.EizoOptions()
    assert (o.configuration_margin, o.tau, o.delta, o.epsilon, o.bisection_steps, o.num_particles,
            o.max_hyperplanes_per_iteration, o.max_iterations, o.mixing_steps, o.seed) == \
        (np.float32(0.01), 0.5, np.float32(0.05), np.float32(0.01), 9, 1000, 10, 500, 40, 1337)


def test_kinematic_tree_traversal_and_limits():
    # kinematic_tree.cpp:82-146: DFS with an explicit stack; config indices in
    # joint order; limits of non-fixed joints
    p = parse("2_branch_5dof.urdf")
    t = p.get_kinematic_tree()
    m = t.get_minimal_kinematic_tree()
    # root children pushed [joint1(0), jointwb1(6)] -> jointwb1 popped first
    assert list(m.joint_traversal_sequence)[:2] == [6, 7]
    assert sorted(m.joint_traversal_sequence) == list(range(8))
    assert list(m.joint_idx_to_config_idx) == [0, 1, 2, 3, 4, -1, -1, -1]
    np.testing.assert_allclose(t.get_position_lower_limits(), [-3.14] * 5, atol=1e-6)
    assert sorted(t.get_leaf_link_indices()) == [3, 6, 8]
    assert t.num_links() == 9


def test_minimal_plant_pairs_roundtrip():
    p = parse("2_branch_5dof.urdf")
    mp = p.get_minimal_plant()
    flat = np.asarray(mp.collision_pairs_flat)
    mp.set_collision_pairs_from_matrix(flat.reshape(-1, 2).T.copy())
    assert np.array_equal(np.asarray(mp.collision_pairs_flat), flat)
    assert mp.num_collision_pairs == 27


def test_hpolyhedron_host_logic():
    import csdecomp_amd as csd
    P = csd.HPolyhedron()
    P.MakeBox(np.array([-1, -2], np.float32), np.array([1, 2], np.float32))
    A, b = P.A(), P.b()
    np.testing.assert_array_equal(A, [[-1, 0], [1, 0], [0, -1], [0, 1]])
    np.testing.assert_array_equal(b, [1, 1, 2, 2])
    assert P.PointInSet(np.array([0.5, 1.9], np.float32))
    assert not P.PointInSet(np.array([1.5, 0], np.float32))
    c = P.ChebyshevCenter()
    np.testing.assert_allclose(c[0], 0, atol=1e-6)
    assert abs(c[1]) <= 1.0 + 1e-6
    m = P.GetMyMinimalHPolyhedron()
    Q = csd.HPolyhedron(m)
    np.testing.assert_array_equal(Q.A(), A)
    # triangle: Chebyshev centre is the incentre
    T = csd.HPolyhedron(np.array([[-1, 0], [0, -1], [1, 1]], np.float32), np.array([0, 0, 1], np.float32))
    r = 1 / (2 + np.sqrt(2))
    np.testing.assert_allclose(T.ChebyshevCenter(), [r, r], atol=1e-5)


def test_device_pointer_checks_refuse_host_memory():
    """Entry points that take device memory refuse host arrays and CPU
    tensors before any kernel sees them; strided reads must stay inside the
    tensor. Host logic only (no GPU call is made)."""
    import torch
    from csdecomp_amd import capi
    with pytest.raises(TypeError):
        capi._dptr(np.zeros(4, np.float32), "q")
    with pytest.raises(TypeError):
        capi._dptr(torch.zeros(4), "q")
    assert capi._dptr(None) is None
    assert capi._dptr(0x1000).value == 0x1000  # a raw address the caller vouches for
    with pytest.raises(TypeError):
        capi._ptr(torch.zeros(4))  # host entry points take numpy arrays
    q = torch.zeros(7 * 100)
    capi._span(q, 100, 1, 100, 7)  # SoA, in range
    capi._span(q, 100, 7, 1, 7)  # (dim, N) column-major, in range
    with pytest.raises(ValueError):
        capi._span(q, 101, 1, 100, 7)
    with pytest.raises(ValueError):
        capi._span(q, 100, 1, 101, 7)
    with pytest.raises(TypeError):
        capi._dtensor(torch.zeros(8), "q", torch.float32, 8)  # on the CPU
