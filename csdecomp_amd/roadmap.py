"""Dynamic-roadmap (DRM) offline maps with csdecomp's names
(planning/roadmap.h, planning/roadmap_builder.cpp, pybind drm_bindings.h).

The voxel collision maps -- which nodes, and which edges, each workspace
voxel blocks -- are computed on the GPU by csd_voxel_collision_map (one
workgroup per node or edge, LDS bitmap of the grid). Host code keeps the
DRM's dictionaries and applies the reference's post-processing (voxels that
block every node or every edge are dropped).
"""
import math

import numpy as np

from . import _pycsdecomp as _c


class RoadmapOptions:
    """planning/roadmap.h:55-75 (same fields and defaults)."""

    def __init__(self):
        self.robot_map_size_x = 1.0
        self.robot_map_size_y = 1.0
        self.robot_map_size_z = 1.0
        self.map_center = np.zeros(3, np.float32)
        self.nodes_processed_before_debug_statement = 1000
        self.max_task_space_distance_between_nodes = 0.25
        self.max_configuration_distance_between_nodes = 1.0
        self.offline_voxel_resolution = 0.05
        self.edge_step_size = 0.01

    def get_offline_voxel_radius(self):
        """sqrt(3) * resolution (roadmap.h:68), in float."""
        return float(np.float32(math.sqrt(3.0) * np.float32(self.offline_voxel_resolution)))

    def grid(self):
        """(nx, ny, nz): ceil(size / (2 resolution)) per axis."""
        r = np.float32(self.offline_voxel_resolution)
        return tuple(int(math.ceil(np.float32(s) / (np.float32(2) * r))) for s in
                     (self.robot_map_size_x, self.robot_map_size_y, self.robot_map_size_z))

    def __repr__(self):
        c = np.asarray(self.map_center, np.float32)
        return (f"RoadmapOptions(robot_map_size_x={self.robot_map_size_x}, "
                f"robot_map_size_y={self.robot_map_size_y}, robot_map_size_z={self.robot_map_size_z}, "
                f"map_center=[{c[0]}, {c[1]}, {c[2]}], "
                f"offline_voxel_resolution={self.offline_voxel_resolution}, "
                f"edge_step_size={self.edge_step_size})")


def GetCollisionVoxelCenter(index, options):
    """Centre of voxel `index` (roadmap_builder.cpp:709-741): per axis
    c - size / 2 + i 2 res + res, the products in float, the sum in double,
    rounded to float."""
    if index < 0:
        raise RuntimeError(f"[RoadmapBuilder:GetCollisionVoxelCenter] invalid index {index}\n.")
    nx, ny, _ = options.grid()
    iz = index // (nx * ny)
    iy = (index - nx * ny * iz) // nx
    ix = index - nx * ny * iz - nx * iy
    res = np.float32(options.offline_voxel_resolution)
    c = np.asarray(options.map_center, np.float32)
    sizes = (options.robot_map_size_x, options.robot_map_size_y, options.robot_map_size_z)
    out = []
    for k, i in enumerate((ix, iy, iz)):
        v = (float(c[k]) - float(np.float32(sizes[k])) / 2.0 + float(np.float32(i * 2) * res)
             + float(res))
        out.append(np.float32(v))
    return np.array(out, np.float32)


def GetCollisionVoxelId(robot_p_voxel, options):
    """Voxel id of a workspace point, -1 outside the map
    (roadmap_builder.cpp:681-707)."""
    nx, ny, nz = options.grid()
    res = np.float32(options.offline_voxel_resolution)
    p = np.asarray(robot_p_voxel, np.float32) - np.asarray(options.map_center, np.float32)
    sizes = (options.robot_map_size_x, options.robot_map_size_y, options.robot_map_size_z)
    idx = []
    for k, n in enumerate((nx, ny, nz)):
        l_true = np.float32(np.float32(n * 2) * res)
        idx.append(math.floor(n * (float(p[k]) + float(np.float32(sizes[k])) / 2.0) / float(l_true)))
    if min(idx) < 0 or idx[0] > nx - 1 or idx[1] > ny - 1 or idx[2] > nz - 1:
        return -1
    return idx[0] + idx[1] * nx + idx[2] * nx * ny


class DRM:
    """planning/roadmap.h DRM: the roadmap and its offline maps."""

    def __init__(self):
        self.collision_map = {}          # voxel id -> [node ids]
        self._edge_collision_map = {}    # voxel id -> {(id1, id2), (id2, id1)}
        self.node_adjacency_map = {}     # node id -> [node ids]
        self.id_to_node_map = {}         # node id -> configuration
        self.id_to_pose_map = {}         # node id -> 4x4 pose of the target link
        self.options = RoadmapOptions()
        self.target_link_index_ = -1
        self.target_link_name_ = "TARGET_LINK_IS_UNINITIALIZED"

    @property
    def edge_collision_map(self):
        """voxel id -> list of (id1, id2) pairs, both directions."""
        return {k: sorted(v) for k, v in self._edge_collision_map.items()}

    def Clear(self):
        self.__init__()

    def GetWorkspaceCorners(self):
        """The 8 corners of the map box (roadmap.cpp:40-60), in the
        reference's order (-x -y -z, -x -y +z, -x +y -z, ...)."""
        o = self.options
        h = [np.float32(o.robot_map_size_x / 2.0), np.float32(o.robot_map_size_y / 2.0),
             np.float32(o.robot_map_size_z / 2.0)]
        c = np.asarray(o.map_center, np.float32)
        return [c + np.array([sx * h[0], sy * h[1], sz * h[2]], np.float32)
                for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]

    # -- cereal binary archive of DRM::serialize (roadmap.h:103-110) --------
    # Layout of cereal's BinaryOutputArchive: arithmetic values as raw
    # little-endian bytes; container sizes as uint64 size tags; a map as its
    # size then key, value pairs; a vector of int32 as its size then the raw
    # block; a string as its size then its bytes; Eigen matrices as rows and
    # cols (size_t each) then the column-major data (roadmap.h:32-53).
    def Write(self, file_name):
        import struct
        out = bytearray()
        u64 = lambda n: out.extend(struct.pack("<Q", n))  # noqa: E731
        i32s = lambda v: out.extend(np.asarray(v, "<i4").tobytes())  # noqa: E731

        def mat(m, rows, cols):
            out.extend(struct.pack("<QQ", rows, cols))
            out.extend(np.asarray(m, "<f4").reshape(rows, cols).T.tobytes())  # column-major

        def vec_map(m):
            u64(len(m))
            for k in sorted(m):
                i32s([k])
                u64(len(m[k]))
                i32s(m[k])

        vec_map(self.collision_map)
        u64(len(self._edge_collision_map))
        for k in sorted(self._edge_collision_map):
            i32s([k])
            es = sorted(self._edge_collision_map[k])
            u64(len(es))
            for e in es:
                u64(2)
                i32s(e)
        vec_map(self.node_adjacency_map)
        u64(len(self.id_to_node_map))
        for k in sorted(self.id_to_node_map):
            v = np.asarray(self.id_to_node_map[k], np.float32).reshape(-1)
            i32s([k])
            mat(v, len(v), 1)
        u64(len(self.id_to_pose_map))
        for k in sorted(self.id_to_pose_map):
            i32s([k])
            mat(self.id_to_pose_map[k], 4, 4)
        o = self.options
        mat(np.asarray(o.map_center, np.float32), 3, 1)
        i32s([o.nodes_processed_before_debug_statement])
        out.extend(np.array([o.max_task_space_distance_between_nodes,
                             o.max_configuration_distance_between_nodes,
                             o.offline_voxel_resolution, o.edge_step_size, o.robot_map_size_x,
                             o.robot_map_size_y, o.robot_map_size_z], "<f4").tobytes())
        name = self.target_link_name_.encode()
        u64(len(name))
        out.extend(name)
        i32s([self.target_link_index_])
        with open(file_name, "wb") as f:
            f.write(bytes(out))

    def Read(self, file_name):
        import struct
        self.Clear()
        with open(file_name, "rb") as f:
            data = f.read()
        pos = [0]

        def take(n):
            if pos[0] + n > len(data):
                raise RuntimeError(f"truncated DRM archive: {file_name}")
            b = data[pos[0]:pos[0] + n]
            pos[0] += n
            return b

        u64 = lambda: struct.unpack("<Q", take(8))[0]  # noqa: E731
        i32 = lambda: struct.unpack("<i", take(4))[0]  # noqa: E731
        i32s = lambda n: np.frombuffer(take(4 * n), "<i4").astype(np.int32)  # noqa: E731

        def mat():
            r, c = struct.unpack("<QQ", take(16))
            return np.frombuffer(take(4 * r * c), "<f4").reshape(c, r).T.astype(np.float32)

        for _ in range(u64()):
            k = i32()
            self.collision_map[k] = i32s(u64()).tolist()
        for _ in range(u64()):
            k = i32()
            s = set()
            for _ in range(u64()):
                s.add(tuple(i32s(u64()).tolist()))
            self._edge_collision_map[k] = s
        for _ in range(u64()):
            k = i32()
            self.node_adjacency_map[k] = i32s(u64()).tolist()
        for _ in range(u64()):
            k = i32()
            self.id_to_node_map[k] = mat().reshape(-1)
        for _ in range(u64()):
            k = i32()
            self.id_to_pose_map[k] = mat()
        o = RoadmapOptions()
        o.map_center = mat().reshape(-1)
        o.nodes_processed_before_debug_statement = i32()
        f = np.frombuffer(take(28), "<f4")
        (o.max_task_space_distance_between_nodes, o.max_configuration_distance_between_nodes,
         o.offline_voxel_resolution, o.edge_step_size, o.robot_map_size_x, o.robot_map_size_y,
         o.robot_map_size_z) = [float(x) for x in f]
        self.options = o
        self.target_link_name_ = take(u64()).decode()
        self.target_link_index_ = i32()


def _voxel_map(mplant, robot_geometry_ids, nodes, edges, step, options):
    """{voxel: sorted item indices} from csd_voxel_collision_map."""
    import torch
    from . import capi
    nx, ny, nz = options.grid()
    V = nx * ny * nz
    dev = torch.device("cuda", torch.cuda.current_device())
    dp = capi.DevicePlant(mplant)
    dn = torch.from_numpy(np.ascontiguousarray(nodes, np.float32)).to(dev)
    de = None if edges is None else torch.from_numpy(np.ascontiguousarray(edges, np.int32)).to(dev)
    off = torch.empty(V + 1, dtype=torch.int64, device=dev)
    center = np.asarray(options.map_center, np.float32)
    size = [options.robot_map_size_x, options.robot_map_size_y, options.robot_map_size_z]
    ne = 0 if edges is None else len(edges)
    args = (dp, dn, len(nodes), de, ne, float(step), list(robot_geometry_ids), center, size,
            float(options.offline_voxel_resolution), off)
    total = capi.voxel_collision_map(*args)
    items = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    total = capi.voxel_collision_map(*args, items=items, capacity=total)
    o = off.cpu().numpy()
    it = items.cpu().numpy()[:total]
    out = {}
    for v in np.nonzero(o[1:] - o[:-1])[0]:
        out[int(v)] = np.sort(it[o[v]:o[v + 1]])
    return out


class RoadmapBuilder:
    """planning/roadmap_builder.h RoadmapBuilder: nodes, roadmap, pose map
    and the offline voxel maps."""

    def __init__(self, plant, target_link_name, options):
        self._plant = plant
        self._mplant = plant.getMinimalPlant()
        self._tree = plant.getKinematicTree()
        self._robot_ids = list(plant.getSceneInspector().robot_geometry_ids)
        self._dim = plant.numConfigurationVariables()
        self._lo = np.asarray(plant.getPositionLowerLimits(), np.float32)
        self._hi = np.asarray(plant.getPositionUpperLimits(), np.float32)
        self._raw = {}
        self._drm = DRM()
        self._drm.target_link_name_ = target_link_name
        self._drm.target_link_index_ = self._tree.get_link_index(target_link_name)
        self._drm.options = options

    def get_drm(self):
        return self._drm

    def reset(self):
        """Clears the DRM and restores default options (roadmap_builder.cpp:127-130)."""
        self._drm.Clear()

    def write(self, file_path):
        self._drm.Write(file_path)

    def read(self, file_path):
        self._drm.Read(file_path)

    def get_random_samples(self, num_samples):
        """num_samples hit-and-run chains of 50 steps from the centre of the
        joint-limit box, on the GPU sampler; the collision-free ones become
        nodes 1, 2, ... where those ids are free (roadmap_builder.cpp:53-69)."""
        dom = _c.HPolyhedron()
        dom.MakeBox(self._lo, self._hi)
        c = ((self._lo + self._hi) / 2).astype(np.float32)
        q = np.asarray(_c.UniformSampleInHPolyhedraCuda([dom], c.reshape(-1, 1), num_samples, 50)[0])
        free = np.asarray(_c.CheckCollisionFreeCuda(np.asfortranarray(q), self._mplant), bool)
        k = 1
        for j in np.nonzero(free)[0]:
            self._raw.setdefault(k, q[:, j].astype(np.float32).copy())
            k += 1

    def add_nodes_manual(self, nodes_to_add):
        """Adds the collision-free columns of (dim, N) that lie in the joint
        limits (roadmap_builder.cpp:71-121); returns the number added. Node
        ids continue from the current count."""
        q = np.asarray(nodes_to_add, np.float32)
        if q.ndim != 2 or q.shape[0] != self._dim:
            raise ValueError("nodes must be a (dim, N) array")
        inside = np.all((q >= self._lo[:, None] - 1e-9) & (q <= self._hi[:, None] + 1e-9), axis=0)
        q = q[:, inside]
        if q.shape[1] == 0:
            return 0
        free = np.asarray(_c.CheckCollisionFreeCuda(np.asfortranarray(q), self._mplant), bool)
        base = len(self._raw)
        added = 0
        for k in np.nonzero(free)[0]:
            added += 1
            self._raw[base + added] = q[:, k].copy()
        return added

    def build_roadmap(self, max_neighbors):
        """The roadmap over the added nodes (roadmap_builder.cpp:137-311).
        For each node, candidates in (squared distance, index) order with
        FLT_EPSILON < d^2 < max_configuration_distance_between_nodes; a
        candidate is accepted unless some link frame moves more than
        max_task_space_distance_between_nodes, the distance exceeds
        max_configuration_distance_between_nodes, or an interior point of the
        edge (roadmap rule: floor(d / edge_step_size) - 2 points, the
        neighbour included) collides; at most max_neighbors per node. Nodes
        with a neighbour enter the DRM, and the adjacency is made symmetric.
        Neighbour lists come page by page from csd_knn_page, edge checks from
        csd_validate_edges (rule 2), link frames from the GPU FK."""
        import torch
        from . import capi
        ids = sorted(self._raw)
        N = len(ids)
        opts = self._drm.options
        if N == 0 or max_neighbors <= 0:
            return
        dim = self._dim
        X = np.array([self._raw[i] for i in ids], np.float32).reshape(N, dim)
        dev = torch.device("cuda", torch.cuda.current_device())
        dX = torch.from_numpy(X).to(dev)
        dp = capi.DevicePlant(self._mplant)
        L = self._tree.num_links()
        tf = np.asarray(_c.computeForwardKinematicsCuda(np.asfortranarray(X.T), self._mplant.kin_tree))
        T = tf[:3, 3::4].T.reshape(N, L, 3)  # link translations
        maxd = np.float32(opts.max_configuration_distance_between_nodes)
        maxt = np.float32(opts.max_task_space_distance_between_nodes)
        K = 16
        accepted = [[] for _ in range(N)]
        pending = np.arange(N, dtype=np.int32)
        after_d2 = after_idx = None
        while pending.size:
            Q = len(pending)
            out_i = torch.empty(Q * K, dtype=torch.int32, device=dev)
            out_d = torch.empty(Q * K, dtype=torch.float32, device=dev)
            capi.knn_page(dX, N, dim, torch.from_numpy(pending).to(dev), Q, K,
                          None if after_d2 is None else torch.from_numpy(after_d2).to(dev),
                          None if after_idx is None else torch.from_numpy(after_idx).to(dev),
                          out_i, out_d)
            idx = out_i.cpu().numpy().reshape(Q, K)
            d2 = out_d.cpu().numpy().reshape(Q, K)
            inside = (idx >= 0) & (d2 < maxd)
            rows, cols = np.nonzero(inside)
            a, b = pending[rows], idx[rows, cols]
            dt = np.sqrt(((T[a] - T[b]) ** 2).sum(-1, dtype=np.float32)).max(-1)
            e = (X[a] - X[b]).astype(np.float32)
            dc = np.sqrt((e * e).sum(-1, dtype=np.float32))
            ok = (dt <= maxt) & (dc <= maxd)
            free = np.zeros(len(a), bool)
            if ok.any():
                edges = np.stack([a[ok], b[ok]], 1).astype(np.int32)
                flags = torch.empty(len(edges), dtype=torch.uint8, device=dev)
                capi.validate_edges(dp, dX, N, torch.from_numpy(edges).to(dev), len(edges),
                                    float(opts.edge_step_size), 2, flags)
                free[np.nonzero(ok)[0]] = flags.cpu().numpy().astype(bool)
            good = np.zeros((Q, K), bool)
            good[rows, cols] = free
            nxt, nd, ni = [], [], []
            for r in range(Q):
                node = pending[r]
                for c in range(K):
                    if len(accepted[node]) >= max_neighbors or not inside[r, c]:
                        break
                    if good[r, c]:
                        accepted[node].append(int(idx[r, c]))
                else:  # a full page inside the radius: the list may go on
                    if len(accepted[node]) < max_neighbors:
                        nxt.append(node)
                        nd.append(d2[r, K - 1])
                        ni.append(idx[r, K - 1])
            pending = np.array(nxt, np.int32)
            after_d2, after_idx = np.array(nd, np.float32), np.array(ni, np.int32)
        drm = self._drm
        for k, nb in enumerate(accepted):
            if nb:
                drm.id_to_node_map[ids[k]] = X[k].copy()
                drm.node_adjacency_map[ids[k]] = [ids[j] for j in nb]
        missing = [(b, a) for a, nb in drm.node_adjacency_map.items() for b in nb
                   if a not in drm.node_adjacency_map.get(b, [])]
        for b, a in missing:
            drm.node_adjacency_map.setdefault(b, []).append(a)

    def build_collision_map(self):
        """drm.collision_map: voxel -> nodes whose robot geometry touches the
        voxel sphere, without voxels that touch every node
        (roadmap_builder.cpp:326-482)."""
        ids = sorted(self._drm.id_to_node_map)
        nodes = np.array([self._drm.id_to_node_map[i] for i in ids], np.float32).reshape(-1, self._dim)
        m = _voxel_map(self._mplant, self._robot_ids, nodes, None, 1.0, self._drm.options)
        n = len(ids)
        self._drm.collision_map = {v: [ids[k] for k in lst] for v, lst in m.items() if len(lst) < n}

    def build_edge_collision_map(self, step_size_edge_collision_map):
        """drm.edge_collision_map: voxel -> edges (both directions) with an
        interior configuration touching the voxel sphere, without voxels that
        block every edge (roadmap_builder.cpp:484-679)."""
        ids = sorted(self._drm.id_to_node_map)
        pos = {i: k for k, i in enumerate(ids)}
        nodes = np.array([self._drm.id_to_node_map[i] for i in ids], np.float32).reshape(-1, self._dim)
        edges = [(pos[a], pos[b]) for a, nbrs in sorted(self._drm.node_adjacency_map.items())
                 for b in nbrs if a < b]
        directed = sum(len(v) for v in self._drm.node_adjacency_map.values())
        out = {}
        if edges:
            m = _voxel_map(self._mplant, self._robot_ids, nodes, np.array(edges, np.int32),
                           step_size_edge_collision_map, self._drm.options)
            for v, lst in m.items():
                s = set()
                for k in lst:
                    a, b = ids[edges[k][0]], ids[edges[k][1]]
                    s.add((a, b))
                    s.add((b, a))
                if len(s) != directed:
                    out[v] = s
        self._drm._edge_collision_map = out

    def build_pose_map(self):
        """drm.id_to_pose_map: pose of the target link at every node."""
        ids = sorted(self._drm.id_to_node_map)
        if not ids:
            self._drm.id_to_pose_map = {}
            return
        q = np.asfortranarray(np.array([self._drm.id_to_node_map[i] for i in ids], np.float32).T)
        tf = np.asarray(_c.computeForwardKinematicsCuda(q, self._mplant.kin_tree))
        L = self._tree.num_links()
        t = self._drm.target_link_index_
        self._drm.id_to_pose_map = {i: tf[:, 4 * (k * L + t):4 * (k * L + t) + 4].copy()
                                    for k, i in enumerate(ids)}
