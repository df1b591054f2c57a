"""Argument handling of EditRegionsCuda's host layer (the checks of
planning/cuda_edit_regions.cu:55-123); every case fails before any device
call, so these run without a GPU."""
import numpy as np
import pytest

from test_oracle_edit import COL, LE, LS, VOX, reference_case


@pytest.fixture(scope="module")
def case(csd):
    plant, collisions, seg = reference_case()
    dom = csd.HPolyhedron()
    dom.MakeBox(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    return plant.getMinimalPlant(), list(plant.getSceneInspector().robot_geometry_ids), dom, \
        collisions, seg


def test_edit_regions_argument_errors(csd, case):
    mp, rg, dom, collisions, seg = case
    opts = csd.EditRegionsOptions()
    run = lambda **kw: csd.EditRegionsCuda(  # noqa: E731
        kw.get("collisions", collisions), kw.get("seg", seg), kw.get("ls", LS), kw.get("le", LE),
        kw.get("regions", [dom] * 3), mp, kw.get("rg", rg), VOX, 0.05, opts)
    with pytest.raises(RuntimeError, match="No collisions"):
        run(collisions=np.zeros((2, 0), np.float32), seg=[])
    small = csd.HPolyhedron()
    small.MakeBox([-0.5, -0.5], [0.5, 0.5])
    with pytest.raises(RuntimeError, match="line start is not contained in region 0"):
        run(regions=[small, dom, dom])
    with pytest.raises(RuntimeError, match="line end is not contained in region 1"):
        run(regions=[dom, small, dom])
    with pytest.raises(IndexError, match="has no region"):
        run(seg=seg[:-1] + [3])
    with pytest.raises(IndexError, match="fewer line segment indices"):
        run(seg=seg[:-1])
    with pytest.raises(ValueError, match="Number of line segments"):
        run(regions=[dom] * 2)
    with pytest.raises(ValueError, match="Robot geometry IDs cannot be empty"):
        run(rg=[])
    with pytest.raises(ValueError, match="collisions must live"):
        run(collisions=np.zeros((3, 12), np.float32))
    with pytest.raises(ValueError, match="at most 4096"):
        run(collisions=np.tile(COL[:, :1], (1, 4097)), seg=[0] * 4097)


def test_edit_regions_options(csd):
    o = csd.EditRegionsOptions()
    o.configuration_margin, o.bisection_steps, o.verbose = 0.02, 5, True
    assert (o.configuration_margin, o.bisection_steps, o.verbose) == (pytest.approx(0.02), 5, True)
