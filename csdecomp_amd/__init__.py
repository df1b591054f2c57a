"""csdecomp_amd -- MI355X (gfx950) implementation of csdecomp's data-parallel
hot path: batched forward kinematics, sphere/box/capsule collision checking
and EI-ZO edge inflation, behind csdecomp's Python names.

The compute runs only on the GPU through ``lib/libcsdecomp_hip.so``; there
is no CPU fallback. Importing fails loudly if the native module is missing
(build it with ``python csdecomp_amd/build.py``).
"""
import os as _os

# PyTorch-ROCm wheels ship their own libamdhip64 (SONAME libamdhip64.so.7,
# requested by torch as "libamdhip64.so"). If this package's library loaded
# first, torch would later map a second HIP runtime and see no GPU. Loading
# torch first (when it is installed) makes both share one runtime.
try:
    import torch as _torch  # noqa: F401
except ImportError:
    pass

try:
    from . import _pycsdecomp
except ImportError as _e:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "csdecomp_amd native module is not built (run `python csdecomp_amd/build.py`): "
        f"{_e}") from _e

from ._pycsdecomp import *  # noqa: F401,F403  (the reference's binding names)
from ._pycsdecomp import _visibility_graph_edges
from .urdf import URDFParser  # noqa: F401
from .roadmap import (DRM, GetCollisionVoxelCenter, GetCollisionVoxelId,  # noqa: F401
                      RoadmapBuilder, RoadmapOptions)

LIB_PATH = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "lib", "libcsdecomp_hip.so")


def VisibilityGraph(configurations, plant, step_size, max_configs_per_batch=10000000):
    """Visibility graph over the columns of ``configurations`` (dim x N).

    Mirrors csdecomp.VisibilityGraph (planning/cuda_visibility_graph.cu:175-375):
    returns an N x N scipy.sparse CSC matrix of uint8 with a 1 at (i, j),
    i < j, for every mutually visible pair (edge points: both ends plus
    (uint)(|qi-qj|/step + 2) - 2 interior points).
    """
    import numpy as np
    import scipy.sparse as sp

    rows, cols = _visibility_graph_edges(configurations, plant, step_size, max_configs_per_batch)
    n = np.asarray(configurations).shape[1]
    data = np.ones(len(rows), dtype=np.uint8)
    return sp.csc_matrix((data, (rows, cols)), shape=(n, n), dtype=np.uint8)
