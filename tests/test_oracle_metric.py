"""Pins the oracle's face placement under an ellipsoid metric
(orc_place_faces_metric). The reference places faces in the Euclidean
metric (cuda_set_builder_utils.cu:359-424: a = (scol - proj) / |scol -
proj|); with an SPD metric M the particles are taken in ascending
|e|_M = sqrt(e' M e) and the face normal is M e / |M e| (the hyperplane
tangent at scol to the M-ellipsoid centred at proj). Checked against:
the identity metric (bit for bit the reference path), a closed form for an
axis-scaled metric, the change of order a metric causes, invariance to
scaling M, and an independent route: Euclidean placement in the
coordinates y = L' x (M = L L') mapped back to x."""
import numpy as np

DIM = 5


def _place(orc, scol, proj, ls, le, metric=None, margin=0.01, max_faces=10, F0=None):
    dim = scol.shape[1]
    cap = 64
    A = np.zeros((cap, dim), np.float32)
    b = np.zeros(cap, np.float32)
    nf = 0
    if F0 is not None:
        A[:len(F0[1])], b[:len(F0[1])] = F0
        nf = len(F0[1])
    added, n = orc.place_faces(A, b, nf, cap, scol, proj, ls, le, margin, max_faces, metric)
    return A[:n], b[:n], added


def _particles(seed, n=300, dim=DIM, spread=0.5):
    rng = np.random.default_rng(seed)
    ls = rng.uniform(-0.2, 0.2, dim).astype(np.float32)
    le = rng.uniform(-0.2, 0.2, dim).astype(np.float32)
    proj = (ls + rng.uniform(0, 1, (n, 1)) * (le - ls)).astype(np.float32)
    scol = (proj + rng.normal(0, spread, (n, dim))).astype(np.float32)
    return scol, proj, ls, le


def _spd(seed, dim=DIM):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(dim, dim))
    return G @ G.T + 0.3 * np.eye(dim)


def test_identity_metric_is_the_reference_path(orc):
    for seed in range(5):
        scol, proj, ls, le = _particles(seed)
        a = _place(orc, scol, proj, ls, le, None)
        b = _place(orc, scol, proj, ls, le, np.eye(DIM))
        assert a[2] == b[2] == 10
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


def test_axis_scaled_closed_form(orc):
    # one particle, e = (1, 1) from proj = 0, M = diag(4, 1): normal (4, 1)/sqrt(17),
    # b = a.scol - margin (the segment, at the origin, stays inside: no relaxation)
    scol = np.array([[1.0, 1.0]], np.float32)
    proj = np.zeros((1, 2), np.float32)
    ls = le = np.zeros(2, np.float32)
    A, b, added = _place(orc, scol, proj, ls, le, np.diag([4.0, 1.0]), margin=0.01)
    assert added == 1
    a = np.array([4.0, 1.0]) / np.sqrt(17.0)
    np.testing.assert_allclose(A[0], a, rtol=0, atol=1e-7)
    assert abs(b[0] - (A[0].astype(np.float64) @ [1.0, 1.0] - 0.01)) < 2e-7
    assert b[0] >= A[0].astype(np.float64) @ [1.0, 1.0] - 0.01  # rounded up


def test_metric_changes_the_order(orc):
    # A: e = (0.2, 0), |e|_M = 2; B: e = (0, 1), |e|_M = 1 under M = diag(100, 1)
    scol = np.array([[0.2, 0.0], [0.0, 1.0]], np.float32)
    proj = np.zeros((2, 2), np.float32)
    z = np.zeros(2, np.float32)
    A, _, _ = _place(orc, scol, proj, z, z, np.diag([100.0, 1.0]), max_faces=1)
    np.testing.assert_array_equal(A[0], [0.0, 1.0])  # B first
    A, _, _ = _place(orc, scol, proj, z, z, None, max_faces=1)
    np.testing.assert_array_equal(A[0], [1.0, 0.0])  # A first in the Euclidean metric


def test_metric_scale_invariance(orc):
    scol, proj, ls, le = _particles(11)
    M = _spd(3)
    a = _place(orc, scol, proj, ls, le, M)
    b = _place(orc, scol, proj, ls, le, 7.25 * M)
    np.testing.assert_allclose(a[0], b[0], atol=1e-7)
    np.testing.assert_allclose(a[1], b[1], atol=1e-6)


def test_metric_equals_euclidean_in_transformed_coordinates(orc):
    """With M = L L', |e|_M = |L' e| and M e = L (L' e): placing faces with
    the metric in x must give the hyperplanes that Euclidean placement gives
    in y = L' x, mapped back ({y: n.y <= c} = {x: (L n).x <= c}). Margin 0
    (a margin is a length along the normal, which the map does not keep)."""
    for seed in range(4):
        scol, proj, ls, le = _particles(20 + seed, n=200, spread=0.4)
        M = _spd(40 + seed)
        L = np.linalg.cholesky(M)
        Ax, bx, nx = _place(orc, scol, proj, ls, le, M, margin=0.0, max_faces=12)
        T = lambda x: (np.asarray(x, np.float64) @ L).astype(np.float32)  # noqa: E731  rows: L' x
        Ay, by, ny = _place(orc, T(scol), T(proj), T(ls), T(le), None, margin=0.0, max_faces=12)
        assert nx == ny and nx >= 5
        Ln = Ay.astype(np.float64) @ L.T  # rows L n
        s = np.linalg.norm(Ln, axis=1)
        np.testing.assert_allclose(Ax, Ln / s[:, None], atol=2e-5)
        np.testing.assert_allclose(bx, by / s, rtol=2e-5, atol=2e-5)
