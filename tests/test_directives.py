"""Model-directive loading (URDFParser::parseDirectives,
plant/urdf_parser.cpp:618-778) checked against the reference's own tests
(plant/tests/urdf_parser_test.cpp:436-541) on its fixtures."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

ASSETS = os.path.join(GOLDEN, "urdf")
DIRECTIVES = os.path.join(ASSETS, "directives")


def _parser(csd):
    p = csd.URDFParser()
    assert p.register_package("test_assets", ASSETS)
    return p


def test_register_package(csd):
    p = csd.URDFParser()
    assert not p.register_package("test_assets", os.path.join(ASSETS, "missing"))
    assert not p.register_package("", ASSETS)
    assert p.register_package("test_assets", ASSETS)


def test_simple_directives(csd):
    p = _parser(csd)
    assert p.parse_directives(os.path.join(DIRECTIVES, "simple.yml"))
    kt = p.get_kinematic_tree()
    assert len(kt.get_links()) == 18 and len(kt.get_joints()) == 17
    assert len(p.get_scene_collision_geometries()) == 10
    links, joints = kt.get_links(), kt.get_joints()
    t = links[kt.get_link_index("table_origin")]
    X = np.asarray(joints[t.parent_joint].X_PL_J)
    want = np.eye(4, dtype=np.float32)
    want[0, 3] = 0.4
    np.testing.assert_allclose(X, want, atol=1e-7)


def test_floating_model_raises(csd):
    p = _parser(csd)
    with pytest.raises(RuntimeError):
        p.parse_directives(os.path.join(DIRECTIVES, "simple_no_anchor_test.yml"))


def test_kinova_table_robotiq(csd):
    p = _parser(csd)
    assert p.parse_directives(os.path.join(DIRECTIVES, "kinova_sens_on_table.yml"))
    kt = p.get_kinematic_tree()
    assert len(kt.get_links()) == 33 and len(kt.get_joints()) == 32
    assert len(p.get_scene_collision_geometries()) == 20
    lower = [-3.1, -2.24, -3.1, -2.57, -3.1, -2.09, -3.1]
    np.testing.assert_allclose(kt.get_position_lower_limits(), lower, rtol=1e-6)
    np.testing.assert_allclose(kt.get_position_upper_limits(), np.negative(lower), rtol=1e-6)
    cpm = p.get_scene_inspector().get_collision_pair_matrix()
    assert cpm.shape[0] == 2 and cpm.shape[1] > 0
    assert not any(cpm[0, c] == 7 and cpm[1, c] == 9 for c in range(cpm.shape[1]))


def test_kinova_plant_names(csd):
    p = _parser(csd)
    with pytest.raises(RuntimeError):
        p.build_plant()  # nothing parsed yet
    assert p.parse_directives(os.path.join(DIRECTIVES, "kinova_sens_on_table.yml"))
    plant = p.build_plant()
    name = "kinova::half_arm_2_link_collision::1"
    names = plant.getSceneCollisionGeometryNames()
    assert name in names
    idx = plant.getSceneCollisionGeometryIndexByName(name)
    assert p.get_scene_inspector().scene_collision_geometry_name_to_geometry_index[name] == idx
    assert plant.getSceneCollisionGeometries()[idx].type == csd.ShapePrimitive.BOX
    assert plant.numConfigurationVariables() == 7


def test_combined_urdf_text(csd):
    p = _parser(csd)
    text = p.parse_directives_to_urdf_string(os.path.join(DIRECTIVES, "simple.yml"))
    assert text.startswith('<robot name="combined_robot">') and text.endswith("</robot>")
    assert 'name="weld_world_to_kinova::base_link" type="fixed"' in text
    assert '<link name="table_origin"/>' in text
    assert 'name="world_to_table_origin" type="fixed"' in text
    assert 'origin xyz="0.4 0 0" rpy="0 0 0"' in text
    assert 'link name="kinova::base_link"' in text
    # a model that names `world` as a parent is refused (urdf_parser.cpp:748-762)
    bad = os.path.join(DIRECTIVES, "..", "3dofrobot.urdf")
    q = csd.URDFParser()
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".yml", delete=False) as f:
        f.write(f"directives:\n- add_model:\n    name: arm\n    file: {bad}\n")
    try:
        with pytest.raises(RuntimeError, match="world link as parent"):
            q.parse_directives(f.name)
    finally:
        os.unlink(f.name)
    assert q.parse_directives_to_urdf_string(os.path.join(ASSETS, "table.urdf")) == ""
