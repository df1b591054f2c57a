"""csd_place_faces_noalloc (no counterpart in the reference; for callers
that cannot allocate device memory) against csd_place_faces, through the
C-ABI: the same faces bit for bit on two calls in a row, the second with
fewer particles, plus the capacity and size limits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIM, DOM = 7, 14


def _inputs(seed, n):
    rng = np.random.default_rng(seed)
    ls = rng.uniform(-0.2, 0.2, DIM).astype(np.float32)
    le = rng.uniform(-0.2, 0.2, DIM).astype(np.float32)
    proj = (ls + rng.uniform(0, 1, (n, 1)) * (le - ls)).astype(np.float32)
    scol = (proj + rng.normal(0, 0.5, (n, DIM))).astype(np.float32)
    A0 = np.zeros((DOM, DIM), np.float32)
    b0 = np.full(DOM, 3.0, np.float32)
    for i in range(DIM):
        A0[2 * i, i], A0[2 * i + 1, i] = -1, 1
    return ls, le, proj, scol, A0, b0


def _run(gpu, fn, seed, n, cap, max_faces):
    import torch
    from csdecomp_amd import capi
    ls, le, proj, scol, A0, b0 = _inputs(seed, n)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(gpu)  # noqa: E731
    dA = torch.zeros(cap * DIM, device=gpu)
    dA[:DOM * DIM] = t(A0.reshape(-1))
    db = torch.zeros(cap, device=gpu)
    db[:DOM] = t(b0)
    dnf = torch.tensor([DOM], dtype=torch.int32, device=gpu)
    added = torch.full((1,), -7, dtype=torch.int32, device=gpu)
    status = torch.full((1,), -7, dtype=torch.int32, device=gpu)
    args = (dA, db, dnf, cap, DIM, t(scol), t(proj), n, t(ls), t(le), 0.01, max_faces, added)
    err = None
    try:
        if fn == "noalloc":
            capi.place_faces_noalloc(*args, status=status)
        else:
            capi.place_faces(*args)
    except capi.CsdError as e:
        err = e.status
    torch.cuda.synchronize()
    return (dA.cpu().numpy(), db.cpu().numpy(), int(dnf.item()), int(added.item()),
            int(status.item()), err)


def test_noalloc_matches_allocating_entry_point(gpu):
    first = True
    for seed, n in ((5, 3000), (6, 500)):  # second call: fewer particles
        ref = _run(gpu, "alloc", seed, n, 64, 10)
        got = _run(gpu, "noalloc", seed, n, 64, 10)
        assert ref[5] is None and got[5] is None and got[4] == 0
        assert got[2] == ref[2] == DOM + 10 and got[3] == ref[3] == 10
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
        first = False
    assert not first


def test_noalloc_capacity_and_limits(gpu):
    import torch
    from csdecomp_amd import capi
    # face capacity runs out after 2 faces: the allocating entry point
    # returns CSD_ERR_CAPACITY, the allocation-free one reports status 2
    ref = _run(gpu, "alloc", 7, 800, DOM + 2, 10)
    got = _run(gpu, "noalloc", 7, 800, DOM + 2, 10)
    assert ref[5] == 2 and got[5] is None and got[4] == 2
    assert got[2] == ref[2] == DOM + 2
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])
    # no particles: nothing placed
    got = _run(gpu, "noalloc", 8, 0, 64, 10)
    assert got[2] == DOM and got[3] == 0 and got[4] == 0
    # above the fixed scratch: refused before any launch
    dummy = torch.zeros(1, device=gpu)
    n = torch.zeros(1, dtype=torch.int32, device=gpu)
    with pytest.raises(capi.CsdError) as e:
        capi.place_faces_noalloc(dummy, dummy, n, 64, DIM, dummy, dummy, 65537, dummy, dummy,
                                 0.01, 10, n)
    assert e.value.status == 2 and "65536" in str(e.value)
