"""ctypes binding of the C-ABI (include/csdecomp_amd.h) for callers that
hold device pointers (e.g. torch tensors on ``cuda:0``): the batched checker,
FK, sampler, edge validation, occupancy and k-NN with inputs already resident
in HBM.
"""
import ctypes as C
import os

import numpy as np

from . import LIB_PATH

_VP, _I32, _I64, _F, _U32, _U64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_uint32, C.c_uint64


class PlantDesc(C.Structure):
    _fields_ = [
        ("num_links", _I32), ("num_joints", _I32), ("num_configuration_variables", _I32),
        ("joint_type", _VP), ("joint_parent_link", _VP), ("joint_child_link", _VP),
        ("joint_config_idx", _VP), ("traversal", _VP), ("joint_X_PL_J", _VP), ("joint_axis", _VP),
        ("num_geometries", _I32), ("geom_type", _VP), ("geom_link", _VP), ("geom_group", _VP),
        ("geom_dims", _VP), ("geom_X_L_B", _VP), ("num_collision_pairs", _I32),
        ("collision_pairs", _VP),
    ]


_SIGS = {
    "csd_status_string": (C.c_char_p, [C.c_int]),
    "csd_last_error": (C.c_char_p, []),
    "csd_init": (C.c_int, [C.c_int]),
    "csd_plant_create": (C.c_int, [C.POINTER(PlantDesc), C.POINTER(_VP)]),
    "csd_plant_destroy": (C.c_int, [_VP]),
    "csd_plant_info": (C.c_int, [_VP, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32)]),
    "csd_forward_kinematics": (C.c_int, [_VP, _VP, _I64, _I64, _I64, _VP, _VP]),
    "csd_check_collision_free": (C.c_int, [_VP, _VP, _I64, _I64, _I64, _VP, _VP]),
    "csd_check_collision_free_voxels": (C.c_int, [_VP, _VP, _I64, _I64, _I64, _VP, _I32, _F, _VP, _I32, _VP, _VP]),
    "csd_voxel_flags": (C.c_int, [_VP, _VP, _I64, _I64, _I64, _VP, _I32, _F, _VP, _I32, _VP, _VP]),
    "csd_hit_and_run": (C.c_int, [_VP, _VP, _VP, _I32, _I32, _I32, _VP, _I64, _I32, _U64, _U32, _VP, _VP]),
    "csd_hit_and_run_segments": (C.c_int, [_VP, _VP, _VP, _I32, _I32, _I32, _VP, _VP, _I64, _I32,
                                           _U64, _U32, _U32, _VP, _VP]),
    "csd_bisect": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _I32, _I32, _VP, _I32, _F, _VP, _I32, _VP, _VP, _VP]),
    "csd_place_faces": (C.c_int, [_VP, _VP, _VP, _I32, _I32, _VP, _VP, _I32, _VP, _VP, _F, _I32, _VP, _VP]),
    "csd_place_faces_noalloc": (C.c_int, [_VP, _VP, _VP, _I32, _I32, _VP, _VP, _I32, _VP, _VP,
                                          _F, _I32, _VP, _VP, _VP]),
    "csd_edit_regions": (C.c_int, [_VP, _VP, _VP, _VP, _I32, _VP, _VP, _I32, _VP, _VP, _VP, _I32,
                                   _VP, _I32, _F, _VP, _I32, _F, _I32, _VP, _VP, _VP, _VP]),
    "csd_validate_edges": (C.c_int, [_VP, _VP, _I32, _VP, _I64, _F, _I32, _VP, _VP]),
    "csd_occupancy": (C.c_int, [_VP, _VP, _I64, _I64, _I64, _VP, _I32, _VP, _VP, _F, _VP, _VP]),
    "csd_knn": (C.c_int, [_VP, _I32, _I32, _I32, _VP, _VP, _VP]),
    "csd_knn_page": (C.c_int, [_VP, _I32, _I32, _VP, _I32, _I32, _VP, _VP, _VP, _VP, _VP]),
    "csd_voxel_collision_map": (C.c_int, [_VP, _VP, _I32, _VP, _I64, _F, _VP, _I32, _VP, _VP, _F,
                                          _VP, _VP, _I64, C.POINTER(_I64), _VP]),
    "csd_stream_synchronize": (C.c_int, [_VP]),
    "csd_distance_field_create": (C.c_int, [_VP, _I32, _F, C.POINTER(_VP)]),
    "csd_distance_field_destroy": (C.c_int, [_VP]),
    "csd_distance_field_reset_occupancy": (C.c_int, [_VP, _VP]),
    "csd_distance_field_set_occupancy": (C.c_int, [_VP, _VP, _I64, _VP]),
    "csd_distance_field_compute": (C.c_int, [_VP, _VP]),
    "csd_distance_field_get_distance": (C.c_int, [_VP, _VP, _I64, _VP, _VP]),
    "csd_distance_field_squared": (C.c_int, [_VP, C.POINTER(_VP)]),
    "csd_mlp_create": (C.c_int, [_I32, _VP, _VP, _U64, C.POINTER(_VP)]),
    "csd_mlp_destroy": (C.c_int, [_VP]),
    "csd_mlp_num_params": (C.c_int, [_VP, C.POINTER(_I64)]),
    "csd_mlp_get_params": (C.c_int, [_VP, _VP]),
    "csd_mlp_set_params": (C.c_int, [_VP, _VP]),
    "csd_mlp_predict": (C.c_int, [_VP, _VP, _I64, _I64, _I64, _VP, _VP]),
    "csd_mlp_gradients": (C.c_int, [_VP, _VP, _VP, _I32, _VP, _VP, _VP]),
    "csd_mlp_train_step": (C.c_int, [_VP, _VP, _VP, _I32, _F, _VP, _VP]),
    "csd_signed_distance": (C.c_int, [_VP, _VP, _I64, _I64, _I64, _VP, _VP]),
    "csd_tune_check": (C.c_int, [_VP, _VP, _I64, _I64, _I64, _VP, _VP, C.POINTER(_I32)]),
}

_lib = None


def lib():
    """The loaded C-ABI library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; build with `python csdecomp_amd/build.py`")
        _lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


class CsdError(RuntimeError):
    def __init__(self, status, what):
        l = lib()
        super().__init__(f"{what}: {l.csd_status_string(status).decode()} "
                         f"({l.csd_last_error().decode()})")
        self.status = status


def check(status, what="csd call"):
    if status != 0:
        raise CsdError(status, what)


def init(device=0):
    check(lib().csd_init(device), "csd_init")


def _ptr(a):
    """Pointer to a HOST array (numpy), or None."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    raise TypeError(f"expected a host numpy array, got {type(a).__name__}")


def _dptr(a, what="device array"):
    """Pointer to DEVICE memory: a torch tensor on the GPU, or a raw address
    (int) the caller vouches for. Host memory is refused: a kernel given a
    host pointer would fault."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        raise TypeError(f"{what}: a host numpy array was passed where device memory is required")
    if hasattr(a, "data_ptr"):  # torch tensor
        if not a.is_cuda:
            raise TypeError(f"{what}: the tensor is on {a.device}, not on the GPU")
        return C.c_void_p(a.data_ptr())
    return C.c_void_p(int(a))


def _dtensor(a, what, dtype, min_numel):
    """Checks a torch tensor handed to a kernel that reads it densely:
    on the GPU, contiguous, of `dtype`, with at least `min_numel` elements."""
    import torch
    if not isinstance(a, torch.Tensor):
        raise TypeError(f"{what}: expected a torch tensor on the GPU")
    if not a.is_cuda:
        raise TypeError(f"{what}: the tensor is on {a.device}, not on the GPU")
    if a.dtype != dtype:
        raise TypeError(f"{what}: dtype {a.dtype}, expected {dtype}")
    if not a.is_contiguous():
        raise ValueError(f"{what}: the tensor must be contiguous")
    if a.numel() < min_numel:
        raise ValueError(f"{what}: {a.numel()} elements, at least {min_numel} needed")
    return C.c_void_p(a.data_ptr())


def _span(a, n, stride_n, stride_j, dim, what="configurations"):
    """For a contiguous torch tensor, checks that every element the two
    strides address lies inside it."""
    if n <= 0 or not hasattr(a, "numel") or not a.is_contiguous():
        return
    last = (n - 1) * stride_n + (dim - 1) * stride_j
    if stride_n < 0 or stride_j < 0 or last >= a.numel():
        raise ValueError(f"{what}: strides ({stride_n}, {stride_j}) over {n} x {dim} address "
                         f"element {last}, past the tensor's {a.numel()}")


def _stream(s):
    if s is None:
        return None
    if hasattr(s, "cuda_stream"):
        return C.c_void_p(s.cuda_stream)
    return C.c_void_p(int(s))


class DevicePlant:
    """Device image of a MinimalPlant, owned by this object."""

    def __init__(self, minimal_plant):
        mp = minimal_plant
        t = mp.kin_tree
        J = t.num_joints
        joints = t.joints
        geoms = mp.scene_collision_geometries
        G = len(geoms)
        self._keep = [
            np.array([int(joints[j].type) for j in range(J)], np.int32),
            np.array([joints[j].parent_link for j in range(J)], np.int32),
            np.array([joints[j].child_link for j in range(J)], np.int32),
            np.array(list(t.joint_idx_to_config_idx)[:J], np.int32),
            np.array(list(t.joint_traversal_sequence)[:J], np.int32),
            np.ascontiguousarray([np.asarray(joints[j].X_PL_J, np.float32).T.reshape(16) for j in range(J)],
                                 np.float32).reshape(-1),
            np.ascontiguousarray([np.asarray(joints[j].axis, np.float32) for j in range(J)], np.float32).reshape(-1),
            np.array([int(g.type) for g in geoms], np.int32),
            np.array([g.link_index for g in geoms], np.int32),
            np.array([int(g.group) for g in geoms], np.int32),
            np.ascontiguousarray([np.asarray(g.dimensions, np.float32) for g in geoms], np.float32).reshape(-1),
            np.ascontiguousarray([np.asarray(g.X_L_B, np.float32).T.reshape(16) for g in geoms],
                                 np.float32).reshape(-1),
            np.ascontiguousarray(mp.collision_pairs_flat, np.int32),
        ]
        k = self._keep
        d = PlantDesc(t.num_links, J, t.num_configuration_variables,
                      *[x.ctypes.data for x in k[:7]], G, *[x.ctypes.data for x in k[7:12]],
                      mp.num_collision_pairs, k[12].ctypes.data)
        h = C.c_void_p()
        check(lib().csd_plant_create(C.byref(d), C.byref(h)), "csd_plant_create")
        self.handle = h
        self.dof = t.num_configuration_variables
        self.num_links = t.num_links

    def info(self):
        L, dof, P, fast = _I32(), _I32(), _I32(), _I32()
        check(lib().csd_plant_info(self.handle, C.byref(L), C.byref(dof), C.byref(P), C.byref(fast)))
        return {"num_links": L.value, "dof": dof.value, "device_pairs": P.value, "fast_path": bool(fast.value)}

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().csd_plant_destroy(self.handle)
        except Exception:
            pass


def check_collision_free(plant, q, n, stride_n, stride_j, out, stream=None):
    _span(q, n, stride_n, stride_j, plant.dof)
    check(lib().csd_check_collision_free(plant.handle, _dptr(q), n, stride_n, stride_j, _dptr(out),
                                         _stream(stream)), "csd_check_collision_free")


def check_collision_free_voxels(plant, q, n, stride_n, stride_j, voxels, num_voxels, radius,
                                robot_geoms, out, stream=None):
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    _span(q, n, stride_n, stride_j, plant.dof)
    check(lib().csd_check_collision_free_voxels(plant.handle, _dptr(q), n, stride_n, stride_j,
                                                _dptr(voxels), num_voxels, radius, _ptr(rg), len(rg),
                                                _dptr(out), _stream(stream)),
          "csd_check_collision_free_voxels")


def voxel_flags(plant, q, n, stride_n, stride_j, voxels, num_voxels, radius, robot_geoms, out,
                stream=None):
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    _span(q, n, stride_n, stride_j, plant.dof)
    check(lib().csd_voxel_flags(plant.handle, _dptr(q), n, stride_n, stride_j, _dptr(voxels),
                                num_voxels, radius, _ptr(rg), len(rg), _dptr(out), _stream(stream)),
          "csd_voxel_flags")


def forward_kinematics(plant, q, n, stride_n, stride_j, out, stream=None):
    _span(q, n, stride_n, stride_j, plant.dof)
    check(lib().csd_forward_kinematics(plant.handle, _dptr(q), n, stride_n, stride_j, _dptr(out),
                                       _stream(stream)), "csd_forward_kinematics")


def hit_and_run(A, b, nf, face_stride, dim, num_polytopes, starts, n, mixing, seed, round_, out,
                stream=None):
    check(lib().csd_hit_and_run(_dptr(A), _dptr(b), _dptr(nf), face_stride, dim, num_polytopes,
                                _dptr(starts), n, mixing, seed, round_, _dptr(out), _stream(stream)),
          "csd_hit_and_run")


def hit_and_run_segments(A, b, nf, face_stride, dim, num_polytopes, seg_start, seg_end, n,
                         mixing, seed, stream_offset, round_, out, stream=None):
    """The EI-ZO round's sampler (csd_hit_and_run_segments): chain i of
    polytope p starts at (i + 1/2) / n along segment p, stream p +
    stream_offset."""
    check(lib().csd_hit_and_run_segments(_dptr(A), _dptr(b), _dptr(nf), face_stride, dim,
                                         num_polytopes, _dptr(seg_start), _dptr(seg_end), n,
                                         mixing, seed, stream_offset, round_, _dptr(out),
                                         _stream(stream)), "csd_hit_and_run_segments")


def bisect(plant, particles, ls, le, seg_idx, n, steps, scol, proj, voxels=None, num_voxels=0,
           radius=0.0, robot_geoms=(), stream=None):
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    check(lib().csd_bisect(plant.handle, _dptr(particles), _dptr(ls), _dptr(le), _dptr(seg_idx), n,
                           steps, _dptr(voxels), num_voxels, radius, _ptr(rg), len(rg), _dptr(scol),
                           _dptr(proj), _stream(stream)), "csd_bisect")


def place_faces(A, b, nf, cap, dim, scol, proj, n, ls, le, margin, max_faces, added, stream=None):
    check(lib().csd_place_faces(_dptr(A), _dptr(b), _dptr(nf), cap, dim, _dptr(scol), _dptr(proj), n,
                                _dptr(ls), _dptr(le), margin, max_faces, _dptr(added), _stream(stream)),
          "csd_place_faces")


def place_faces_noalloc(A, b, nf, cap, dim, scol, proj, n, ls, le, margin, max_faces, added,
                        status=None, stream=None):
    """csd_place_faces_noalloc: the same faces without allocating (fixed
    scratch of the library, at most 65536 particles); asynchronous, the
    status (0, or 2 for face capacity exhausted) goes to `status`."""
    check(lib().csd_place_faces_noalloc(_dptr(A), _dptr(b), _dptr(nf), cap, dim, _dptr(scol),
                                        _dptr(proj), n, _dptr(ls), _dptr(le), margin, max_faces,
                                        _dptr(added), _dptr(status), _stream(stream)),
          "csd_place_faces_noalloc")


def edit_regions(plant, collisions, seg_idx, offsets, num_collisions, ls, le, num_regions, A, b,
                 nf, cap, margin, steps, scol, proj, status, voxels=None, num_voxels=0,
                 radius=0.0, robot_geoms=(), stream=None):
    """csd_edit_regions on device arrays: collisions grouped by region
    (offsets [R+1]); A [R][cap][dim], b [R][cap], nf [R] edited in place."""
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    check(lib().csd_edit_regions(plant.handle, _dptr(collisions), _dptr(seg_idx), _dptr(offsets),
                                 num_collisions, _dptr(ls), _dptr(le), num_regions, _dptr(A), _dptr(b),
                                 _dptr(nf), cap, _dptr(voxels), num_voxels, radius, _ptr(rg), len(rg),
                                 margin, steps, _dptr(scol), _dptr(proj), _dptr(status),
                                 _stream(stream)),
          "csd_edit_regions")


def validate_edges(plant, nodes, num_nodes, edges, num_edges, step, rule, out, stream=None):
    check(lib().csd_validate_edges(plant.handle, _dptr(nodes), num_nodes, _dptr(edges), num_edges,
                                   step, rule, _dptr(out), _stream(stream)), "csd_validate_edges")


def occupancy(plant, q, n, stride_n, stride_j, robot_geoms, center, size, resolution, grid,
              stream=None):
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    c = np.ascontiguousarray(center, np.float32)
    s = np.ascontiguousarray(size, np.float32)
    _span(q, n, stride_n, stride_j, plant.dof)
    check(lib().csd_occupancy(plant.handle, _dptr(q), n, stride_n, stride_j, _ptr(rg), len(rg),
                              _ptr(c), _ptr(s), resolution, _dptr(grid), _stream(stream)),
          "csd_occupancy")


def voxel_collision_map(plant, nodes, num_nodes, edges, num_edges, step, robot_geoms, center,
                        size, resolution, offsets, items=None, capacity=0, stream=None):
    """csd_voxel_collision_map; returns the number of entries (total). With
    items=None only the offsets are written."""
    rg = np.ascontiguousarray(robot_geoms, np.int32)
    c = np.ascontiguousarray(center, np.float32)
    sz = np.ascontiguousarray(size, np.float32)
    total = _I64(0)
    check(lib().csd_voxel_collision_map(plant.handle, _dptr(nodes), num_nodes, _dptr(edges),
                                        num_edges, step, _ptr(rg), len(rg), _ptr(c), _ptr(sz),
                                        resolution, _dptr(offsets), _dptr(items), capacity,
                                        C.byref(total), _stream(stream)),
          "csd_voxel_collision_map")
    return int(total.value)


def knn_page(nodes, num_nodes, dim, queries, num_queries, k, after_d2, after_idx, out_idx,
             out_d2, stream=None):
    """csd_knn_page: the next k neighbours of each query node after
    (after_d2, after_idx) (both None: the first k)."""
    check(lib().csd_knn_page(_dptr(nodes), num_nodes, dim, _dptr(queries), num_queries, k,
                             _dptr(after_d2), _dptr(after_idx), _dptr(out_idx), _dptr(out_d2),
                             _stream(stream)), "csd_knn_page")


def knn(nodes, num_nodes, dim, k, out_idx, out_d2=None, stream=None):
    check(lib().csd_knn(_dptr(nodes), num_nodes, dim, k, _dptr(out_idx), _dptr(out_d2),
                        _stream(stream)), "csd_knn")


def synchronize(stream=None):
    check(lib().csd_stream_synchronize(_stream(stream)), "csd_stream_synchronize")


def tune_check(plant, q, n, stride_n, stride_j, out, stream=None, force=0):
    """Autotunes the checker on this batch (see csd_tune_check) and returns
    the chosen shape (1, 2, 4: one pass with that many waves per 64
    configurations; 11, 12, 14: two passes; 20: screened). force: record
    that shape instead of timing."""
    s = _I32(force)
    _span(q, n, stride_n, stride_j, plant.dof)
    check(lib().csd_tune_check(plant.handle, _dptr(q), n, stride_n, stride_j, _dptr(out),
                               _stream(stream), C.byref(s)), "csd_tune_check")
    return int(s.value)


def signed_distance(plant, q, n, stride_n, stride_j, out, stream=None):
    """Smallest FP32 signed distance over the collision pairs (labels)."""
    _span(q, n, stride_n, stride_j, plant.dof)
    check(lib().csd_signed_distance(plant.handle, _dptr(q), n, stride_n, stride_j, _dptr(out),
                                    _stream(stream)), "csd_signed_distance")


class Mlp:
    """The signed-distance surrogate network (csd_mlp_* of the C-ABI)."""

    def __init__(self, in_dim, lower, upper, seed=0):
        lo = np.ascontiguousarray(lower, np.float32)
        hi = np.ascontiguousarray(upper, np.float32)
        h = _VP()
        check(lib().csd_mlp_create(in_dim, _ptr(lo), _ptr(hi), seed, C.byref(h)), "csd_mlp_create")
        self.handle = h
        self.in_dim = in_dim
        cnt = _I64()
        check(lib().csd_mlp_num_params(self.handle, C.byref(cnt)), "csd_mlp_num_params")
        self.num_params = int(cnt.value)

    @property
    def address(self):
        """The csd_mlp* as an integer (for CudaEdgeInflator.set_guidance)."""
        return int(self.handle.value)

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        check(lib().csd_mlp_get_params(self.handle, _ptr(p)), "csd_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        check(lib().csd_mlp_set_params(self.handle, _ptr(p)), "csd_mlp_set_params")

    def predict(self, q, n, stride_n, stride_j, out, stream=None):
        _span(q, n, stride_n, stride_j, self.in_dim)
        check(lib().csd_mlp_predict(self.handle, _dptr(q, "q"), n, stride_n, stride_j,
                                    _dptr(out, "out"), _stream(stream)), "csd_mlp_predict")

    def _batch_args(self, q, target, batch, loss):
        """q [batch][in_dim] and target [batch] as the kernels read them:
        contiguous float32 on the GPU."""
        import torch
        if batch <= 0 or batch % 128:
            raise ValueError("batch must be a positive multiple of 128")
        if q.dim() != 2 or q.shape[0] < batch or q.shape[1] != self.in_dim:
            raise ValueError(f"q: shape {tuple(q.shape)}, expected ({batch}, {self.in_dim})")
        return (_dtensor(q, "q", torch.float32, batch * self.in_dim),
                _dtensor(target, "target", torch.float32, batch),
                None if loss is None else _dtensor(loss, "loss", torch.float32, 1))

    def gradients(self, q, target, batch, grads=None, loss=None, stream=None):
        import torch
        qp, tp, lp = self._batch_args(q, target, batch, loss)
        gp = None if grads is None else _dtensor(grads, "grads", torch.float32, self.num_params)
        check(lib().csd_mlp_gradients(self.handle, qp, tp, batch, gp, lp, _stream(stream)),
              "csd_mlp_gradients")

    def train_step(self, q, target, batch, lr, loss=None, stream=None):
        qp, tp, lp = self._batch_args(q, target, batch, loss)
        check(lib().csd_mlp_train_step(self.handle, qp, tp, batch, lr, lp, _stream(stream)),
              "csd_mlp_train_step")

    def __del__(self):
        # csd_mlp_destroy refuses while an EI-ZO workspace holds the model
        # (it is then left allocated rather than freed under the workspace)
        try:
            if getattr(self, "handle", None) and _lib is not None:
                if _lib.csd_mlp_destroy(self.handle) == 0:
                    self.handle = None
        except Exception:
            pass


class DistanceField:
    """csd_distance_field_*: exact distance field of a G^3 voxel grid."""

    def __init__(self, origin, grid_size, voxel_size):
        o = np.ascontiguousarray(origin, np.float32)
        h = _VP()
        check(lib().csd_distance_field_create(o.ctypes.data_as(_VP), grid_size, voxel_size,
                                              C.byref(h)), "csd_distance_field_create")
        self.handle, self.grid_size = h, grid_size

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.csd_distance_field_destroy(self.handle)
            self.handle = None

    def reset_occupancy(self, stream=None):
        check(lib().csd_distance_field_reset_occupancy(self.handle, _stream(stream)),
              "csd_distance_field_reset_occupancy")

    def set_occupancy(self, points, n, stream=None):
        check(lib().csd_distance_field_set_occupancy(self.handle, _dptr(points), n,
                                                     _stream(stream)),
              "csd_distance_field_set_occupancy")

    def compute(self, stream=None):
        check(lib().csd_distance_field_compute(self.handle, _stream(stream)),
              "csd_distance_field_compute")

    def get_distance(self, points, n, out, stream=None):
        check(lib().csd_distance_field_get_distance(self.handle, _dptr(points), n, _dptr(out),
                                                    _stream(stream)),
              "csd_distance_field_get_distance")

    def squared_ptr(self):
        p = _VP()
        check(lib().csd_distance_field_squared(self.handle, C.byref(p)),
              "csd_distance_field_squared")
        return p.value
