"""ctypes wrapper of the CPU restatement (csd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; nothing in csdecomp_amd/ imports
it. Build with ``make -C oracle``.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libcsd_oracle.so")
# the same source built for speed (FMA contraction): CPU baseline timing only
LIB_FAST = os.path.join(HERE, "build", "libcsd_oracle_fast.so")

RULE_CPU, RULE_GPU = 0, 1

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class OrcPlant(C.Structure):
    _fields_ = [
        ("num_links", C.c_int32), ("num_joints", C.c_int32), ("num_config_vars", C.c_int32),
        ("joint_type", C.c_void_p), ("joint_parent_link", C.c_void_p), ("joint_child_link", C.c_void_p),
        ("joint_config_idx", C.c_void_p), ("traversal", C.c_void_p),
        ("joint_X_PL_J", C.c_void_p), ("joint_axis", C.c_void_p),
        ("num_geoms", C.c_int32),
        ("geom_type", C.c_void_p), ("geom_link", C.c_void_p), ("geom_group", C.c_void_p),
        ("geom_dims", C.c_void_p), ("geom_X_L_B", C.c_void_p),
        ("num_pairs", C.c_int32), ("pairs", C.c_void_p),
    ]


class EizoOptions(C.Structure):
    _fields_ = [
        ("configuration_margin", C.c_float), ("tau", C.c_float), ("delta", C.c_float),
        ("epsilon", C.c_float), ("bisection_steps", C.c_uint32), ("num_particles", C.c_uint32),
        ("max_hyperplanes_per_iteration", C.c_uint32), ("max_iterations", C.c_uint32),
        ("mixing_steps", C.c_uint32), ("seed", C.c_uint64),
    ]


class EizoInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("samples_drawn", C.c_int64), ("faces_added", C.c_int32)]


def _load(path=LIB):
    if not os.path.exists(path):
        raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
    lib = C.CDLL(path)
    P = C.POINTER(OrcPlant)
    sig = {
        "orc_fk_f32": (None, [P, _f32p, _f32p]),
        "orc_fk_f64": (None, [P, _f64p, _f64p]),
        "orc_check_batch": (None, [P, _f32p, C.c_int64, C.c_int, _u8p, C.c_int]),
        "orc_check_voxels_batch": (None, [P, _f32p, C.c_int64, _f32p, C.c_int32, C.c_float,
                                          _i32p, C.c_int32, C.c_int, _u8p, C.c_int]),
        "orc_voxel_flags_batch": (None, [P, _f32p, C.c_int64, _f32p, C.c_int32, C.c_float,
                                         _i32p, C.c_int32, C.c_int, _u8p, C.c_int]),
        "orc_min_sd_batch": (None, [P, _f32p, C.c_int64, _f32p, C.c_int32, C.c_float, _i32p,
                                    C.c_int32, _f64p, C.c_int]),
        "orc_fk_error_batch": (None, [P, _f32p, C.c_int64, _f64p, C.c_int]),
        "orc_check_edge": (C.c_int, [P, _f32p, _f32p, C.c_float, C.c_int]),
        "orc_edge_num_steps": (C.c_int32, [_f32p, _f32p, C.c_int32, C.c_float]),
        "orc_philox4x32_10": (None, [np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")]),
        "orc_hit_and_run": (None, [_f32p, _f32p, C.c_int32, C.c_int32, _f32p, C.c_int64,
                                   C.c_int64, C.c_int32, C.c_uint64, C.c_uint32, C.c_uint32, _f32p]),
        "orc_hit_and_run_segment": (None, [_f32p, _f32p, C.c_int32, C.c_int32, _f32p, _f32p,
                                           C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_uint64,
                                           C.c_uint32, C.c_uint32, _f32p]),
        "orc_num_samples_unadaptive": (C.c_uint32, [C.c_float, C.c_float, C.c_float]),
        "orc_delta_k": (C.c_float, [C.c_float, C.c_int32]),
        "orc_project_onto_segment": (None, [_f32p, _f32p, _f32p, C.c_int32, _f32p, C.c_void_p]),
        "orc_bisect": (None, [P, _f32p, _f32p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                              C.c_float, C.c_void_p, C.c_int32, C.c_int, _f32p]),
        "orc_place_faces": (C.c_int32, [_f32p, _f32p, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                        _f32p, _f32p, C.c_int32, _f32p, _f32p, C.c_float, C.c_int32]),
        "orc_place_faces_metric": (C.c_int32, [_f32p, _f32p, C.POINTER(C.c_int32), C.c_int32,
                                               C.c_int32, _f32p, _f32p, C.c_int32, _f32p, _f32p,
                                               C.c_float, C.c_int32, C.c_void_p]),
        "orc_voxelize": (None, [_f32p, C.c_int64, _f32p, C.c_float, C.c_int32, _u8p]),
        "orc_edt_brute": (None, [_u8p, C.c_int32, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                 C.c_int]),
        "orc_edit_region": (C.c_int32, [_f32p, _f32p, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                        _f32p, _f32p, C.c_int32, _f32p, _f32p, C.c_float]),
        "orc_eizo": (C.c_int32, [P, _i32p, C.c_int32, C.c_void_p, C.c_int32, C.c_float, _f32p,
                                 _f32p, C.POINTER(C.c_int32), C.c_int32, _f32p, _f32p,
                                 C.POINTER(EizoOptions), C.c_uint32, C.c_int, C.c_int,
                                 C.POINTER(EizoInfo)]),
        "orc_eizo_batch": (None, [P, _i32p, C.c_int32, _f32p, _f32p, C.c_int32, C.c_int32,
                                  C.c_int32, _f32p, _f32p, C.POINTER(EizoOptions), C.c_int,
                                  C.c_int, _f32p, _f32p, _i32p, _i32p, C.c_void_p, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = None
_LIB_FAST = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def lib_fast():
    """The speed build (CPU baseline timing; results are not checked)."""
    global _LIB_FAST
    if _LIB_FAST is None:
        _LIB_FAST = _load(LIB_FAST)
    return _LIB_FAST


class Plant:
    """Flat arrays of a plant (the content of csdecomp's MinimalPlant)."""

    def __init__(self, minimal_plant):
        mp = minimal_plant
        t = mp.kin_tree
        J = t.num_joints
        self.num_links = t.num_links
        self.dof = t.num_configuration_variables
        joints = t.joints
        self.joint_type = np.array([int(joints[j].type) for j in range(J)], np.int32)
        self.joint_parent = np.array([joints[j].parent_link for j in range(J)], np.int32)
        self.joint_child = np.array([joints[j].child_link for j in range(J)], np.int32)
        self.joint_cfg = np.array(list(t.joint_idx_to_config_idx)[:J], np.int32)
        self.traversal = np.array(list(t.joint_traversal_sequence)[:J], np.int32)
        self.joint_X = np.ascontiguousarray(
            np.stack([np.asarray(joints[j].X_PL_J, np.float32).T.reshape(16) for j in range(J)])
            if J else np.zeros((0, 16), np.float32))
        self.joint_axis = np.ascontiguousarray(
            np.stack([np.asarray(joints[j].axis, np.float32) for j in range(J)])
            if J else np.zeros((0, 3), np.float32))
        geoms = mp.scene_collision_geometries
        G = len(geoms)
        self.geom_type = np.array([int(g.type) for g in geoms], np.int32)
        self.geom_link = np.array([g.link_index for g in geoms], np.int32)
        self.geom_group = np.array([int(g.group) for g in geoms], np.int32)
        self.geom_dims = np.ascontiguousarray(
            np.stack([np.asarray(g.dimensions, np.float32) for g in geoms])
            if G else np.zeros((0, 3), np.float32))
        self.geom_X = np.ascontiguousarray(
            np.stack([np.asarray(g.X_L_B, np.float32).T.reshape(16) for g in geoms])
            if G else np.zeros((0, 16), np.float32))
        self.pairs = np.ascontiguousarray(np.asarray(mp.collision_pairs_flat, np.int32))
        self.num_pairs = mp.num_collision_pairs
        self._s = OrcPlant(
            self.num_links, J, self.dof,
            self.joint_type.ctypes.data, self.joint_parent.ctypes.data, self.joint_child.ctypes.data,
            self.joint_cfg.ctypes.data, self.traversal.ctypes.data,
            self.joint_X.ctypes.data, self.joint_axis.ctypes.data, G,
            self.geom_type.ctypes.data, self.geom_link.ctypes.data, self.geom_group.ctypes.data,
            self.geom_dims.ctypes.data, self.geom_X.ctypes.data, self.num_pairs,
            self.pairs.ctypes.data)

    @property
    def ref(self):
        return C.byref(self._s)


def _aos(q, dof):
    """(dof, N) configurations -> contiguous [N][dof] float32."""
    q = np.asarray(q, np.float32)
    if q.ndim == 1:
        q = q[:, None]
    assert q.shape[0] == dof
    return np.ascontiguousarray(q.T)


def fk(plant, q, double=False):
    """Link transforms of one configuration: [L][4][4] (row-major view)."""
    L = plant.num_links
    if double:
        out = np.zeros(L * 16, np.float64)
        lib().orc_fk_f64(plant.ref, np.ascontiguousarray(q, np.float64), out)
    else:
        out = np.zeros(L * 16, np.float32)
        lib().orc_fk_f32(plant.ref, np.ascontiguousarray(q, np.float32), out)
    return out.reshape(L, 4, 4).transpose(0, 2, 1)


def check(plant, q, rule=RULE_GPU, threads=0, fast=False):
    """Flags (1 = free); fast=True runs the speed build (baseline timing)."""
    qa = _aos(q, plant.dof)
    out = np.zeros(qa.shape[0], np.uint8)
    (lib_fast() if fast else lib()).orc_check_batch(plant.ref, qa, qa.shape[0], rule, out, threads)
    return out


def _vox(voxels):
    if voxels is None:
        return np.zeros((0, 3), np.float32), 0
    v = np.asarray(voxels, np.float32)
    v = np.ascontiguousarray(v.T) if v.shape[0] == 3 else np.ascontiguousarray(v)
    return v, v.shape[0]


def check_voxels(plant, q, voxels, radius, robot_geoms, rule=RULE_GPU, threads=0):
    qa = _aos(q, plant.dof)
    v, nv = _vox(voxels)
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    out = np.zeros(qa.shape[0], np.uint8)
    lib().orc_check_voxels_batch(plant.ref, qa, qa.shape[0], v, nv, radius, rg, len(rg), rule,
                                 out, threads)
    return out


def voxel_flags(plant, q, voxels, radius, robot_geoms, rule=RULE_GPU, threads=0):
    qa = _aos(q, plant.dof)
    v, nv = _vox(voxels)
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    out = np.zeros((qa.shape[0], nv), np.uint8)
    lib().orc_voxel_flags_batch(plant.ref, qa, qa.shape[0], v, nv, radius, rg, len(rg), rule,
                                out.reshape(-1), threads)
    return out


def min_sd(plant, q, voxels=None, radius=0.0, robot_geoms=(), threads=0):
    qa = _aos(q, plant.dof)
    v, nv = _vox(voxels)
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    out = np.zeros(qa.shape[0], np.float64)
    lib().orc_min_sd_batch(plant.ref, qa, qa.shape[0], v, nv, radius, rg, len(rg), out, threads)
    return out


def fk_error(plant, q, threads=0):
    qa = _aos(q, plant.dof)
    out = np.zeros((qa.shape[0], plant.num_links), np.float64)
    lib().orc_fk_error_batch(plant.ref, qa, qa.shape[0], out.reshape(-1), threads)
    return out


def check_edge(plant, q1, q2, step=0.01, rule=RULE_GPU):
    return bool(lib().orc_check_edge(plant.ref, np.ascontiguousarray(q1, np.float32),
                                     np.ascontiguousarray(q2, np.float32), step, rule))


def edge_num_steps(q1, q2, step):
    q1 = np.ascontiguousarray(q1, np.float32)
    return lib().orc_edge_num_steps(q1, np.ascontiguousarray(q2, np.float32), len(q1), step)


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32), np.ascontiguousarray(key, np.uint32), out)
    return out


def hit_and_run(A, b, start, n, mixing, seed=1337, stream=0, round_=0, chain_offset=0):
    A = np.ascontiguousarray(A, np.float32)
    F, dim = A.shape
    out = np.zeros((n, dim), np.float32)
    lib().orc_hit_and_run(A, np.ascontiguousarray(b, np.float32), F, dim,
                          np.ascontiguousarray(start, np.float32), n, chain_offset, mixing, seed,
                          stream, round_, out)
    return out


def hit_and_run_segment(A, b, ls, le, n, mixing, seed=1337, stream=0, round_=0, chain_offset=0,
                        total=None):
    """Chains spread along ls -> le (orc_hit_and_run_segment): chain c of
    `total` (default n) starts at fraction (c + 1/2) / total."""
    A = np.ascontiguousarray(A, np.float32)
    F, dim = A.shape
    out = np.zeros((n, dim), np.float32)
    lib().orc_hit_and_run_segment(A, np.ascontiguousarray(b, np.float32), F, dim,
                                  np.ascontiguousarray(ls, np.float32),
                                  np.ascontiguousarray(le, np.float32), n, chain_offset,
                                  n if total is None else total, mixing, seed, stream, round_, out)
    return out


def num_samples_unadaptive(eps, delta, tau):
    return lib().orc_num_samples_unadaptive(eps, delta, tau)


def delta_k(delta, it):
    return lib().orc_delta_k(delta, it)


def project(p, q1, q2):
    p = np.ascontiguousarray(p, np.float32)
    proj = np.zeros_like(p)
    d = C.c_float(0)
    lib().orc_project_onto_segment(p, np.ascontiguousarray(q1, np.float32),
                                   np.ascontiguousarray(q2, np.float32), len(p), proj,
                                   C.cast(C.pointer(d), C.c_void_p))
    return proj, d.value


def bisect(plant, particle, proj, steps, rule=RULE_GPU):
    sc = np.zeros(plant.dof, np.float32)
    lib().orc_bisect(plant.ref, np.ascontiguousarray(particle, np.float32),
                     np.ascontiguousarray(proj, np.float32), plant.dof, steps, None, 0, 0.0,
                     None, 0, rule, sc)
    return sc


def place_faces(A, b, nf, cap, scol, proj, ls, le, margin, max_faces, metric=None):
    """A [cap][dim], b [cap] float32 arrays modified in place; returns
    (added, faces). metric: SPD dim x dim (None = identity, the reference)."""
    n = C.c_int32(nf)
    dim = A.shape[1]
    M = None if metric is None else np.ascontiguousarray(metric, np.float64)
    added = lib().orc_place_faces_metric(A.reshape(-1), b, C.byref(n), cap, dim,
                                         np.ascontiguousarray(scol, np.float32).reshape(-1),
                                         np.ascontiguousarray(proj, np.float32).reshape(-1),
                                         len(scol), np.ascontiguousarray(ls, np.float32),
                                         np.ascontiguousarray(le, np.float32), margin, max_faces,
                                         None if M is None else M.ctypes.data_as(C.c_void_p))
    return added, n.value


def bisect_voxels(plant, particle, proj, steps, voxels, radius, robot_geoms, rule=RULE_GPU):
    """orc_bisect with the full check (plant pairs + robot geometries vs voxels)."""
    sc = np.zeros(plant.dof, np.float32)
    v, nv = _vox(voxels)
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    lib().orc_bisect(plant.ref, np.ascontiguousarray(particle, np.float32),
                     np.ascontiguousarray(proj, np.float32), plant.dof, steps,
                     v.ctypes.data_as(C.c_void_p) if nv else None, nv, radius,
                     rg.ctypes.data_as(C.c_void_p) if len(rg) else None, len(rg), rule, sc)
    return sc


def edit_region(A, b, scol, proj, ls, le, margin):
    """Faces of one region after editing (orc_edit_region): A [F][dim], b [F];
    scol, proj [u][dim]. Returns (A', b', faces added)."""
    A = np.asarray(A, np.float32)
    F, dim = A.shape
    scol = np.ascontiguousarray(scol, np.float32).reshape(-1, dim)
    u = scol.shape[0]
    cap = F + u
    Ab = np.zeros((max(cap, 1), dim), np.float32)
    bb = np.zeros(max(cap, 1), np.float32)
    Ab[:F], bb[:F] = A, np.asarray(b, np.float32)
    n = C.c_int32(F)
    added = lib().orc_edit_region(Ab.reshape(-1), bb, C.byref(n), max(cap, 1), dim,
                                  scol.reshape(-1),
                                  np.ascontiguousarray(proj, np.float32).reshape(-1), u,
                                  np.ascontiguousarray(ls, np.float32),
                                  np.ascontiguousarray(le, np.float32), margin)
    assert added >= 0
    return Ab[:n.value].copy(), bb[:n.value].copy(), added


def edit_regions(plant, collisions, segment_idxs, line_starts, line_ends, regions, robot_geoms,
                 voxels, radius, margin=0.01, steps=9, rule=RULE_GPU):
    """EditRegionsCuda restated (planning/cuda_edit_regions.cu:44-300).

    collisions (dim, C); segment_idxs [C]; line_starts / line_ends (dim, R);
    regions: list of (A [F][dim], b [F]). Returns (list of (A, b),
    projections (dim, C), optimised collisions (dim, C)), the collisions
    grouped by segment in order of appearance."""
    col = np.asarray(collisions, np.float32)
    seg = np.asarray(segment_idxs, np.int64)
    ls = np.asarray(line_starts, np.float32)
    le = np.asarray(line_ends, np.float32)
    order = np.argsort(seg, kind="stable")
    col, seg = col[:, order], seg[order]
    n = col.shape[1]
    proj = np.zeros_like(col)
    scol = np.zeros_like(col)
    for k in range(n):
        proj[:, k], _ = project(col[:, k], ls[:, seg[k]], le[:, seg[k]])
        scol[:, k] = bisect_voxels(plant, col[:, k], proj[:, k], steps, voxels, radius,
                                   robot_geoms, rule)
    out = []
    for r, (A, b) in enumerate(regions):
        m = seg == r
        if not m.any():
            out.append((np.asarray(A, np.float32), np.asarray(b, np.float32)))
            continue
        A2, b2, _ = edit_region(A, b, scol[:, m].T, proj[:, m].T, ls[:, r], le[:, r], margin)
        out.append((A2, b2))
    return out, proj, scol


def eizo(plant, robot_geoms, A_dom, b_dom, ls, le, opts, cap=512, stream=0, rule=RULE_GPU,
         spread_starts=False):
    dim = plant.dof
    A = np.zeros((cap, dim), np.float32)
    b = np.zeros(cap, np.float32)
    F = A_dom.shape[0]
    A[:F] = A_dom
    b[:F] = b_dom
    nf = C.c_int32(F)
    o = EizoOptions(*opts)
    info = EizoInfo()
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    rc = lib().orc_eizo(plant.ref, rg, len(rg), None, 0, 0.0, A.reshape(-1), b, C.byref(nf), cap,
                        np.ascontiguousarray(ls, np.float32), np.ascontiguousarray(le, np.float32),
                        C.byref(o), stream, rule, int(spread_starts), C.byref(info))
    if rc != 0:
        raise RuntimeError("oracle EI-ZO face capacity exceeded")
    return A[:nf.value], b[:nf.value], info


def eizo_batch(plant, robot_geoms, A_dom, b_dom, ls, le, opts, cap=512, rule=RULE_GPU,
               spread_starts=False, threads=0, fast=False):
    """orc_eizo over the columns of ls / le (dim x R), one region per OpenMP
    thread. Returns (A [R][cap][dim], b [R][cap], faces [R], status [R],
    infos [R])."""
    dim = plant.dof
    R = ls.shape[1]
    A = np.zeros((R, cap, dim), np.float32)
    b = np.zeros((R, cap), np.float32)
    nf = np.zeros(R, np.int32)
    st = np.zeros(R, np.int32)
    infos = (EizoInfo * R)()
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    o = EizoOptions(*opts)
    L = lib_fast() if fast else lib()
    L.orc_eizo_batch(plant.ref, rg, len(rg), np.ascontiguousarray(A_dom, np.float32),
                     np.ascontiguousarray(b_dom, np.float32), A_dom.shape[0], cap, R,
                     np.ascontiguousarray(ls.T, np.float32), np.ascontiguousarray(le.T, np.float32),
                     C.byref(o), rule, int(spread_starts), A.reshape(-1), b.reshape(-1), nf, st,
                     C.cast(infos, C.c_void_p), threads)
    return A, b, nf, st, list(infos)


FAR = 0x3fffffff


def voxelize(points, origin, voxel, G):
    """uint8 occupancy [z][y][x] of points (3, M) (orc_voxelize)."""
    p = np.ascontiguousarray(np.asarray(points, np.float32).T)
    occ = np.zeros(G * G * G, np.uint8)
    lib().orc_voxelize(p.reshape(-1), p.shape[0], np.ascontiguousarray(origin, np.float32),
                       voxel, G, occ)
    return occ.reshape(G, G, G)


def edt_brute(occ, threads=0):
    """Squared distance in voxels to the nearest occupied voxel, [z][y][x]
    int32, FAR where nothing is occupied (orc_edt_brute)."""
    occ = np.ascontiguousarray(occ, np.uint8)
    G = occ.shape[0]
    d2 = np.zeros(occ.size, np.int32)
    lib().orc_edt_brute(occ.reshape(-1), G, d2, threads)
    return d2.reshape(occ.shape)


def roadmap_grid(center, size, res):
    """(nx, ny, nz) and the voxel centres (3, V), ids ix + iy nx + iz nx ny:
    per axis c - size/2 + i 2 res + res with i 2 res in float and the sum in
    double, rounded to float (roadmap_builder.cpp:709-741)."""
    res = np.float32(res)
    n = [int(np.ceil(np.float32(s) / (np.float32(2) * res))) for s in size]
    axes = []
    for k in range(3):
        i = np.arange(n[k])
        axes.append((np.float64(np.float32(center[k])) - np.float64(np.float32(size[k])) / 2.0
                     + ((i * 2).astype(np.float32) * res).astype(np.float64)
                     + np.float64(res)).astype(np.float32))
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return n, np.stack([x.ravel(), y.ravel(), z.ravel()])


def roadmap_edge_configs(q1, q2, step):
    """Interior configurations of an edge by the roadmap rule
    (roadmap_builder.cpp:526-541): n = max(0, floor(|q2 - q1| / step) - 2),
    t = (k + 1) / n (double, rounded to float), q = t q2 + (1 - t) q1 in
    float; |q2 - q1| as a sequential float sum of squares. (dim, n)."""
    q1 = np.asarray(q1, np.float32)
    q2 = np.asarray(q2, np.float32)
    s = np.float32(0)
    for d in range(len(q1)):
        e = np.float32(q2[d] - q1[d])
        s = np.float32(s + np.float32(e * e))
    dist = np.float32(np.sqrt(s))
    n = max(0, int(np.floor(np.float32(dist / np.float32(step)))) - 2)
    out = np.zeros((len(q1), n), np.float32)
    for k in range(n):
        t = np.float32((k + 1) / (1.0 * n))
        out[:, k] = np.float32(t) * q2 + np.float32(np.float32(1) - t) * q1
    return out


def voxel_map(plant, nodes, edges, step, robot_geoms, center, size, res, radius_delta=0.0):
    """{voxel: sorted items} -- every voxel of the map checked against every
    configuration of every item with voxel_flags (the reference's
    BuildCollisionMap / BuildEdgeCollisionMap without their chunking).
    nodes (N, dim); edges (E, 2) or None (items are then the nodes)."""
    _, centres = roadmap_grid(center, size, res)
    radius = float(np.float32(np.sqrt(3.0) * np.float32(res))) + radius_delta
    nodes = np.asarray(nodes, np.float32)
    if edges is None:
        configs = [nodes[i][:, None] for i in range(len(nodes))]
    else:
        configs = [roadmap_edge_configs(nodes[a], nodes[b], step) for a, b in edges]
    out = {}
    for item, q in enumerate(configs):
        if q.shape[1] == 0:
            continue
        flags = voxel_flags(plant, q, centres, radius, robot_geoms)
        for v in np.nonzero((flags == 0).any(axis=0))[0]:
            out.setdefault(int(v), []).append(item)
    return out


def build_roadmap(plant, X, max_neighbors, max_config, max_task, step, rule=RULE_GPU):
    """RoadmapBuilder::BuildRoadmap restated (roadmap_builder.cpp:137-311):
    per node, candidates with FLT_EPSILON < d^2 < max_config in (d^2, index)
    order (d^2 a sequential float sum); skip if any link frame moves more than
    max_task or |d| > max_config; accept if no interior point of the edge
    (roadmap_edge_configs) collides; at most max_neighbors; then the
    symmetric fix-up. X (N, dim); returns {node index: [node indices]}."""
    X = np.asarray(X, np.float32)
    N, dim = X.shape
    T = np.stack([fk(plant, X[k])[:, :3, 3] for k in range(N)])
    adj = {}
    for i in range(N):
        d2 = np.zeros(N, np.float32)
        for d in range(dim):
            e = (X[:, d] - X[i, d]).astype(np.float32)
            d2 = (d2 + (e * e).astype(np.float32)).astype(np.float32)
        cand = [j for j in np.lexsort((np.arange(N), d2))
                if d2[j] > np.float32(1.1920928955078125e-07) and d2[j] < np.float32(max_config)]
        acc = []
        for j in cand:
            if len(acc) >= max_neighbors:
                break
            dt = np.sqrt(((T[i] - T[j]) ** 2).sum(-1, dtype=np.float32)).max()
            if dt > np.float32(max_task):
                continue
            e = (X[i] - X[j]).astype(np.float32)
            if np.sqrt((e * e).sum(dtype=np.float32)) > np.float32(max_config):
                continue
            q = roadmap_edge_configs(X[i], X[j], step)
            if q.shape[1] == 0 or check(plant, q, rule).all():
                acc.append(int(j))
        if acc:
            adj[i] = acc
    missing = [(b, a) for a, nb in adj.items() for b in nb if a not in adj.get(b, [])]
    for b, a in missing:
        adj.setdefault(b, []).append(a)
    return adj
