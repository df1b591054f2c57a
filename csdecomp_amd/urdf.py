"""URDF and model-directive loading into a csdecomp plant (the reference's
URDFParser, plant/urdf_parser.cpp:78-158, :209-616 and :618-778, on the
standard-library XML reader and PyYAML's safe loader).

Link 0 is the implicit ``world`` link; links and joints are indexed in
document order. Collision primitives: box (full edge lengths), sphere
(radius), capsule (radius, length; axis along z). Cylinders raise, as in the
reference.
"""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import _pycsdecomp as _c


def yaml_safe_loader():
    """PyYAML's SafeLoader (constructs plain data only); the directives'
    `!Rpy {deg: [...]}` tag reads as its mapping."""
    import yaml

    class Loader(yaml.SafeLoader):
        pass

    Loader.add_constructor("!Rpy", lambda loader, node: loader.construct_mapping(node, deep=True))
    return Loader


def yaml_load(text, loader):
    import yaml
    return yaml.load(text, Loader=loader)  # loader derives from yaml.SafeLoader

FLT_MAX = np.finfo(np.float32).max


def _rpy_to_matrix(r, p, y):
    """R = Rz(yaw) Ry(pitch) Rx(roll) (urdf_parser.cpp:356-362)."""
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    return np.array([
        [cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr],
        [sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr],
        [-sp, cp * sr, cp * cr],
    ])


def _origin(elem):
    """4x4 float32 transform of an <origin> element (urdf_parser.cpp:321-365)."""
    X = np.eye(4)
    if elem is not None:
        xyz = elem.get("xyz")
        if xyz is None:
            raise RuntimeError("[CSDECOMP:URDFPARSER] no position specified for joint origin")
        X[:3, 3] = [float(v) for v in xyz.split()]
        rpy = elem.get("rpy")
        if rpy is not None:
            X[:3, :3] = _rpy_to_matrix(*[float(v) for v in rpy.split()])
    return X.astype(np.float32)


def _rename(elem, prefix):
    """Prefix every name of a model with `prefix::`, and the parent / child
    links of its joints (urdf_parser.cpp:732-778). A joint whose parent is
    `world` is refused: models are attached with add_weld."""
    name = elem.get("name")
    if name is not None:
        elem.set("name", f"{prefix}::{name}")
    if elem.tag == "joint":
        parent, child = elem.find("parent"), elem.find("child")
        if parent is not None and parent.get("link") is not None:
            if parent.get("link") == "world":
                raise RuntimeError(
                    f"[CSDECOMP:URDFParser] Directly specifying the world link as parent in one of "
                    f"the URDFS (found in {name}) is not supported due to potential ambiguity. To "
                    f"weld the robot to the world specify this in a directive using the -add_weld "
                    f"directive.")
            parent.set("link", f"{prefix}::{parent.get('link')}")
        if child is not None and child.get("link") is not None:
            child.set("link", f"{prefix}::{child.get('link')}")
    for c in elem:
        _rename(c, prefix)


def _num(v):
    """A number as an iostream with default precision writes it (6
    significant digits), so combined URDFs match the reference's text."""
    return f"{float(v):g}"


class _DirectiveLoader(yaml_safe_loader()):
    pass


class URDFParser:
    """Python counterpart of csdecomp's URDFParser with the same method names."""

    def __init__(self):
        self._reset()
        self._packages = {}

    def _reset(self):
        self._tree = _c.KinematicTree()
        self._geoms = []
        self._names = {}
        self._inspector = _c.SceneInspector()
        self._parsed = False

    # -- packages / directives -------------------------------------------
    def register_package(self, name, path):
        """Maps package://name/... to path (urdf_parser.cpp:780-815); False
        for an empty name or a path that does not exist."""
        if not name or not path or not os.path.exists(path):
            return False
        self._packages[name] = path
        return True

    def parse_directives_to_urdf_string(self, directives_file):
        """One URDF combining the models of a directives file
        (urdf_parser.cpp:618-701): add_model renames the model's links,
        joints and geometries to `model::name`; add_weld adds a fixed joint;
        add_frame adds a link on a fixed joint with the given translation and
        rotation (Rpy degrees). Returns "" if the file has no directives or a
        model cannot be read."""
        with open(directives_file) as f:
            cfg = yaml_load(f.read(), _DirectiveLoader)
        if not isinstance(cfg, dict) or not cfg.get("directives"):
            return ""
        out = ['<robot name="combined_robot">']
        for d in cfg["directives"]:
            if "add_model" in d:
                m = d["add_model"]
                try:
                    root = ET.parse(self._resolve_lenient(m["file"])).getroot()
                except (OSError, ET.ParseError):
                    return ""
                _rename(root, m["name"])
                out += [ET.tostring(c, encoding="unicode") for c in root]
            elif "add_weld" in d:
                w = d["add_weld"]
                out.append(f'  <joint name="weld_{w["parent"]}_to_{w["child"]}" type="fixed">\n'
                           f'    <parent link="{w["parent"]}"/>\n'
                           f'    <child link="{w["child"]}"/>\n  </joint>')
            elif "add_frame" in d:
                fr = d["add_frame"]
                X = fr["X_PF"]
                base, t = X["base_frame"], X["translation"]
                rot = [float(v) * math.pi / 180.0 for v in X["rotation"]["deg"]]
                out.append(f'  <link name="{fr["name"]}"/>')
                out.append(f'  <joint name="{base}_to_{fr["name"]}" type="fixed">\n'
                           f'    <parent link="{base}"/>\n    <child link="{fr["name"]}"/>\n'
                           f'    <origin xyz="{_num(t[0])} {_num(t[1])} {_num(t[2])}" '
                           f'rpy="{_num(rot[0])} {_num(rot[1])} {_num(rot[2])}"/>\n  </joint>')
        out.append("</robot>")
        return "\n".join(out)

    def parse_directives(self, directives_file):
        """Parses the combined URDF of a directives file (urdf_parser.cpp:703-713)."""
        text = self.parse_directives_to_urdf_string(directives_file)
        if not text:
            return False
        return self.parse_urdf_string(text)

    def override_kinematic_tree(self, kinematic_tree):
        """Replaces the parsed kinematic tree (urdf_parser.h:127-129)."""
        self._tree = kinematic_tree

    def override_scene_collision_geometries(self, geometries):
        """Replaces the parsed collision geometries (urdf_parser.h:116-122)."""
        self._geoms = list(geometries)

    def _resolve_lenient(self, uri):
        """package:// paths of registered packages resolve; anything else is
        used as written (urdf_parser.cpp:716-729)."""
        if uri.startswith("package://"):
            pkg, _, rel = uri[len("package://"):].partition("/")
            if pkg in self._packages:
                return os.path.join(self._packages[pkg], rel)
        return uri

    def _resolve(self, uri):
        if uri.startswith("package://"):
            rest = uri[len("package://"):]
            pkg, _, rel = rest.partition("/")
            if pkg not in self._packages:
                raise RuntimeError(f"unknown package {pkg}")
            return os.path.join(self._packages[pkg], rel)
        return uri

    # -- URDF ----------------------------------------------------------------
    def parse_urdf(self, urdf_file):
        if not os.path.exists(urdf_file):
            raise RuntimeError(f"Failed to load URDF file: {urdf_file}")
        with open(urdf_file) as f:
            return self.parse_urdf_string(f.read())

    def parse_urdf_string(self, text):
        try:
            robot = ET.fromstring(text)
        except ET.ParseError as e:
            raise RuntimeError(f"Failed to load URDF file: {e}") from e
        if robot.tag != "robot":
            return False
        self._reset()
        world = _c.Link()
        world.name = "world"
        world.parent_joint = -1
        self._tree.add_link(world)
        for link in robot.findall("link"):
            self._parse_link(link)
        for joint in robot.findall("joint"):
            if not self._parse_joint(joint):
                return False
        joints = self._tree.get_joints()
        if not any(j.parent_link == 0 for j in joints):
            raise RuntimeError(
                "[CSDECOMP:URDFPARSER] The provided urdf is not connected to the 'world' link via a joint.")
        for l in self._tree.get_links():
            if l.name != "world" and not (0 <= l.parent_joint < len(joints)):
                raise RuntimeError(
                    f"[CSDECOMP:URDFPARSER] The link {l.name} has no valid parent joint.")
        self._tree.finalize()
        self.finalize_scene()
        self._parsed = True
        return True

    def _parse_link(self, elem):
        name = elem.get("name")
        if elem.get("origin") is not None:
            raise RuntimeError("[CSDECOMP:URDFParser:ParseLink] Global origin tags are not supported in links.")
        link = _c.Link()
        link.name = name
        link.parent_joint = -1
        self._tree.add_link(link)
        idx = self._tree.get_link_index(name)
        for n, col in enumerate(elem.findall("collision")):
            cname = col.get("name")
            cname = f"{cname}::{n + 1}" if cname is not None else f"{name}::col_geom::{n + 1}"
            g = self._parse_collision(col)
            g.link_index = idx
            self._names[cname] = len(self._geoms)
            self._geoms.append(g)

    def _parse_collision(self, col):
        g = _c.CollisionGeometry()
        g.type = _c.ShapePrimitive.BOX
        g.dimensions = np.zeros(3, np.float32)
        g.group = _c.CollisionGroup.ROBOT
        g.X_L_B = _origin(col.find("origin")) if col.find("origin") is not None else np.eye(4, dtype=np.float32)
        geo = col.find("geometry")
        if geo is not None:
            shapes = list(geo)
            if not shapes:
                raise RuntimeError("No valid geometry type specified. Spheres, Boxes and Capsules are supported.")
            s = shapes[0]
            if s.tag == "box":
                g.type = _c.ShapePrimitive.BOX
                g.dimensions = np.array([float(v) for v in s.get("size").split()], np.float32)
            elif s.tag == "sphere":
                g.type = _c.ShapePrimitive.SPHERE
                g.dimensions = np.array([float(s.get("radius")), -1, -1], np.float32)
            elif s.tag == "capsule":
                g.type = _c.ShapePrimitive.CAPSULE
                r, ln = float(s.get("radius")), float(s.get("length"))
                g.dimensions = np.array([r, ln, r], np.float32)
            elif s.tag == "cylinder":
                raise RuntimeError("cylider not yet implemented")
        return g

    def _parse_joint(self, elem):
        kinds = {"fixed": _c.JointType.FIXED, "revolute": _c.JointType.REVOLUTE,
                 "prismatic": _c.JointType.PRISMATIC}
        t = elem.get("type")
        if t not in kinds:
            return False
        j = _c.Joint()
        j.name = elem.get("name")
        j.type = kinds[t]
        j.X_PL_J = _origin(elem.find("origin"))
        if j.type != _c.JointType.FIXED:
            ax = elem.find("axis")
            if ax is None:
                raise RuntimeError(
                    "[CSDECOMP:URDFPARSER] Error: One or more revolute or prismatic joints have no axis specified!")
            a = np.array([float(v) for v in ax.get("xyz").split()], np.float32)
            j.axis = (a / np.linalg.norm(a)).astype(np.float32)
        else:
            j.axis = np.zeros(3, np.float32)
        parent, child = elem.find("parent"), elem.find("child")
        if parent is None or child is None:
            return False
        j.parent_link = self._tree.get_link_index(parent.get("link"))
        j.child_link = self._tree.get_link_index(child.get("link"))
        lim = elem.find("limit")
        if lim is None:
            j.position_lower_limit, j.position_upper_limit = -FLT_MAX, FLT_MAX
        else:
            j.position_lower_limit = float(lim.get("lower", 0.0))
            j.position_upper_limit = float(lim.get("upper", 0.0))
        self._tree.add_joint(j)
        return True

    # -- scene ----------------------------------------------------------------
    def finalize_scene(self):
        geoms, insp = _c.finalize_scene(self._tree, self._geoms, self._names)
        self._geoms = list(geoms)
        self._inspector = insp

    def get_kinematic_tree(self):
        return self._tree

    def get_minimal_kinematic_tree(self):
        return self._tree.get_minimal_kinematic_tree()

    def get_scene_collision_geometries(self):
        return list(self._geoms)

    def get_scene_inspector(self):
        return self._inspector

    def build_plant(self):
        if not self._tree.is_finalized():
            raise RuntimeError("No URDF has been parsed yet, cannot build plant!")
        return _c.Plant(self._tree, self._inspector, self._geoms)

    def get_minimal_plant(self):
        return self.build_plant().getMinimalPlant()
