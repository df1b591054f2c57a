"""GPU runs on the Kinova + Robotiq + table scene loaded from the reference's
model directives (all-box robot geometry: the generic checker)."""
import os

import numpy as np
import pytest

from conftest import assert_flags_match
from test_directives import ASSETS, DIRECTIVES

pytestmark = pytest.mark.gpu


def _kinova(csd):
    p = csd.URDFParser()
    assert p.register_package("test_assets", ASSETS)
    assert p.parse_directives(os.path.join(DIRECTIVES, "kinova_sens_on_table.yml"))
    return p


def test_kinova_flags_parity(gpu, orc, csd):
    p = _kinova(csd)
    plant = p.build_plant()
    mp = plant.getMinimalPlant()
    op = orc.Plant(mp)
    rng = np.random.default_rng(4)
    lo = np.asarray(plant.getPositionLowerLimits(), np.float32)
    hi = np.asarray(plant.getPositionUpperLimits(), np.float32)
    q = np.asfortranarray(rng.uniform(lo, hi, (20000, 7)).astype(np.float32).T)
    got = np.asarray(csd.CheckCollisionFreeCuda(q, mp))
    ref = orc.check(op, q, orc.RULE_GPU)
    assert_flags_match(got, ref, orc.min_sd(op, q), "kinova")
    assert 0.05 < got.mean() < 0.95  # both outcomes are exercised


def test_kinova_edge_inflation(gpu, csd):
    """edge_inflation_zero_order_test.cpp:432-500 (FCITestKinova): the region
    holds both ends and 1000 samples from its Chebyshev centre collide less
    than 1.2 epsilon of the time."""
    p = _kinova(csd)
    kt = p.get_kinematic_tree()
    mp = p.get_minimal_plant()
    ids = p.get_scene_inspector().robot_geometry_ids
    ls = np.zeros(7, np.float32)
    le = np.array([0, 1, -1, 0.5, 0, 0, 0], np.float32)
    dom = csd.HPolyhedron()
    dom.MakeBox(kt.get_position_lower_limits(), kt.get_position_upper_limits())
    opts = csd.EizoOptions()
    opts.num_particles = 2000
    opts.max_iterations = 100
    opts.verbose = False
    mpoly, info = csd.EizoCuda(ls, le, np.zeros((3, 0), np.float32), 0.01, dom, mp, ids, opts)
    region = csd.HPolyhedron(mpoly)
    assert info.converged
    assert region.PointInSet(ls) and region.PointInSet(le)
    c = np.asarray(region.ChebyshevCenter(), np.float32)
    samples = csd.UniformSampleInHPolyhedraCuda([region], c.reshape(7, 1), 1000, 50)[0]
    free = np.asarray(csd.CheckCollisionFreeCuda(samples, mp))
    assert (1 - free.mean()) < opts.epsilon * 1.2
