"""GPU parity of the distance field (csd_distance_field_*, DistanceField)
against the oracle: squared distances bit-exact on every voxel."""
import ctypes

import numpy as np
import pytest

from test_oracle_field import REF_G, REF_ORIGIN, REF_VOXEL, fixture_points

pytestmark = pytest.mark.gpu


def _squared(field, G):
    """The field's squared-distance grid, copied to the host."""
    from csdecomp_amd import capi
    out = np.zeros(G * G * G, np.int32)
    capi.check(capi.lib().csd_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(field.squared_ptr()), out.nbytes,
                                         None), "copy")
    capi.check(capi.lib().csd_stream_synchronize(None), "sync")
    return out.reshape(G, G, G)


def test_reference_fixture_gpu(gpu, orc, csd):
    """cuda_distance_field_test.cpp:16-79 through DistanceField, plus every
    voxel's squared distance equal to the oracle's."""
    pts = fixture_points()
    f = csd.DistanceField(REF_ORIGIN, REF_G, float(REF_VOXEL))
    f.setOccupancy(pts)
    f.computeEDT()
    idx = np.arange(0, REF_G, 4)
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    P = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32) * REF_VOXEL + REF_ORIGIN
    got = np.asarray(f.getDistance(np.ascontiguousarray(P.T)))
    brute = np.sqrt(((P[:, None, :] - pts.T[None, :, :]) ** 2).sum(-1)).min(1)
    np.testing.assert_allclose(got, brute, atol=1e-3)
    occ = orc.voxelize(pts, REF_ORIGIN, REF_VOXEL, REF_G)
    ref = orc.edt_brute(occ)
    # the same through the C-ABI: whole grid, exact
    from csdecomp_amd import capi
    import torch
    g = capi.DistanceField(REF_ORIGIN, REF_G, float(REF_VOXEL))
    g.set_occupancy(torch.from_numpy(np.ascontiguousarray(pts.T)).to(gpu), pts.shape[1])
    g.compute()
    np.testing.assert_array_equal(_squared(g, REF_G), ref)
    # outside the grid: 0
    far = np.array([[5.0, -5.0], [0.0, 0.0], [0.0, 0.0]], np.float32)
    assert (np.asarray(f.getDistance(far)) == 0).all()


@pytest.mark.parametrize("G,density,seed", [(1, 1.0, 0), (37, 0.001, 1), (64, 0.02, 2),
                                            (100, 0.0005, 3), (48, 0.6, 4)])
def test_random_grids(gpu, orc, G, density, seed):
    import torch
    from csdecomp_amd import capi
    rng = np.random.default_rng(seed)
    occ = (rng.random((G, G, G)) < density).astype(np.uint8)
    occ[rng.integers(G), rng.integers(G), rng.integers(G)] = 1
    zz, yy, xx = np.nonzero(occ)
    vs = 0.05
    origin = np.array([0.3, -1.0, 2.0], np.float32)
    pts = (np.stack([xx, yy, zz], -1).astype(np.float32) * np.float32(vs) + origin)
    f = capi.DistanceField(origin, G, vs)
    f.set_occupancy(torch.from_numpy(np.ascontiguousarray(pts)).to(gpu), len(pts))
    f.compute()
    np.testing.assert_array_equal(_squared(f, G), orc.edt_brute(occ))


def test_empty_reset_and_large(gpu, orc):
    import torch
    from csdecomp_amd import capi
    G = 256
    f = capi.DistanceField([0, 0, 0], G, 0.01)
    f.compute()
    d = torch.empty(2, device=gpu)
    q = torch.tensor([[0.5, 0.5, 0.5], [3.0, 3.0, 3.0]], device=gpu)  # the grid spans 2.56 m
    f.get_distance(q, 2, d)
    torch.cuda.synchronize()
    assert np.isinf(d[0].item()) and d[1].item() == 0.0  # empty field; outside the grid
    rng = np.random.default_rng(7)
    pts = rng.uniform(0, 2.55, (60, 3)).astype(np.float32)
    f.set_occupancy(torch.from_numpy(pts).to(gpu), 60)
    f.compute()
    occ = orc.voxelize(pts.T, np.zeros(3, np.float32), np.float32(0.01), G)
    np.testing.assert_array_equal(_squared(f, G), orc.edt_brute(occ))
    f.reset_occupancy()
    f.set_occupancy(torch.from_numpy(pts[:1]).to(gpu), 1)
    f.compute()
    one = orc.voxelize(pts[:1].T, np.zeros(3, np.float32), np.float32(0.01), G)
    np.testing.assert_array_equal(_squared(f, G), orc.edt_brute(one))
    with pytest.raises(capi.CsdError):
        capi.DistanceField([0, 0, 0], 1025, 0.01)
