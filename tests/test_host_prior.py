"""Host-side planar-prior construction and edge map (libhpm_host.so), on the
CPU.  Pins:
  * Delaunay: a subset of scipy.spatial.Delaunay's (Qhull) triangles on random
    points in general position, missing only thin hull triangles whose
    circumcircle reaches the bounding triangle (as cv::Subdiv2D); on
    co-circular grid points every triangle is counter-clockwise and no input
    point lies strictly inside any circumcircle;
  * plane through three points: the null vector of [X 1] from numpy's SVD,
    scaled to |n| = 1 and w >= 0;
  * support points: per 5 x 5 block the first minimum of cost - confidence,
    kept below 0.1, and the same with +0.2 (HPM.cpp:1026-1073), built by hand;
  * raster: the reference's barycentric loops only write pixels of the
    triangle's closed hull and cover its interior;
  * edge map: a vertical step edge gives one edge column, a constant image none.
"""
import numpy as np
import pytest

pytest.importorskip("scipy")
from scipy.spatial import Delaunay as QhullDelaunay  # noqa: E402

from hpm_mi355x import pipeline as P  # noqa: E402


def _tri_set(tris):
    return {tuple(sorted(map(tuple, np.asarray(t).reshape(3, 2).tolist()))) for t in tris}


@pytest.mark.parametrize("seed,n", [(0, 30), (1, 200), (2, 1000)])
def test_delaunay_matches_qhull(seed, n):
    """Every triangle is one of Qhull's; the only Qhull triangles missing are
    thin ones on the convex hull whose circumcircle reaches the bounding
    triangle (centre +- 3 x the larger side), as with cv::Subdiv2D."""
    rng = np.random.default_rng(seed)
    W, H = 4000, 3000
    pts = np.unique(rng.integers(0, [W, H], size=(n, 2)), axis=0)
    ours = _tri_set(P.delaunay(W, H, pts))
    ref = QhullDelaunay(pts.astype(np.float64))
    ref_set = {tuple(sorted(map(tuple, pts[s].tolist()))) for s in ref.simplices}
    assert ours <= ref_set
    hull = set(map(tuple, pts[ref.convex_hull.ravel()].tolist()))
    big = 3 * max(W, H)
    for t in ref_set - ours:
        a, b, c = (np.array(v, np.float64) for v in t)
        assert any(v in hull for v in t)
        d = 2 * (a[0] * (b[1] - c[1]) + b[0] * (c[1] - a[1]) + c[0] * (a[1] - b[1]))
        ux = ((a @ a) * (b[1] - c[1]) + (b @ b) * (c[1] - a[1]) + (c @ c) * (a[1] - b[1])) / d
        uy = ((a @ a) * (c[0] - b[0]) + (b @ b) * (a[0] - c[0]) + (c @ c) * (b[0] - a[0])) / d
        r = np.hypot(a[0] - ux, a[1] - uy)
        corners = [(W / 2 + big, H / 2), (W / 2, H / 2 + big), (W / 2 - big, H / 2 - big)]
        assert any(np.hypot(x - ux, y - uy) < r for x, y in corners)
    assert len(ref_set - ours) <= max(3, len(ref_set) // 50)


def test_delaunay_properties_on_grid_points():
    """Co-circular grid points: any valid Delaunay triangulation is accepted;
    check the empty-circumcircle property with exact integer arithmetic."""
    rng = np.random.default_rng(5)
    pts = np.unique(rng.integers(0, 40, size=(150, 2)) * 5, axis=0)
    dup = np.concatenate([pts, pts[:20]])  # duplicates are ignored
    tris = P.delaunay(200, 200, dup)
    assert len(tris) > 0
    # the triangles tile (almost all of) the convex hull without overlap:
    # areas sum to at most the hull area, short only by thin hull triangles
    from scipy.spatial import ConvexHull
    hull_area = ConvexHull(pts.astype(float)).volume
    T = tris.reshape(-1, 3, 2).astype(np.float64)
    area = 0.5 * np.abs((T[:, 1, 0] - T[:, 0, 0]) * (T[:, 2, 1] - T[:, 0, 1])
                        - (T[:, 1, 1] - T[:, 0, 1]) * (T[:, 2, 0] - T[:, 0, 0]))
    assert area.sum() <= hull_area + 1e-6 and area.sum() >= 0.98 * hull_area
    used = {tuple(v) for v in T.reshape(-1, 2).astype(int).tolist()}
    assert used <= set(map(tuple, pts.tolist()))
    P3 = [np.asarray(t, np.int64).reshape(3, 2) for t in tris]
    for a, b, c in P3:
        assert (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]) > 0
    for a, b, c in P3:
        d = pts.astype(np.int64)
        ad, bd, cd = a - d, b - d, c - d
        det = (ad[:, 0] * (bd[:, 1] * (cd ** 2).sum(1) - (bd ** 2).sum(1) * cd[:, 1])
               - ad[:, 1] * (bd[:, 0] * (cd ** 2).sum(1) - (bd ** 2).sum(1) * cd[:, 0])
               + (ad ** 2).sum(1) * (bd[:, 0] * cd[:, 1] - bd[:, 1] * cd[:, 0]))
        assert np.all(det <= 0)


def test_plane_through_points_matches_svd():
    rng = np.random.default_rng(3)
    for _ in range(50):
        X = rng.uniform(-2, 2, (3, 3)).astype(np.float32)
        X[:, 2] += 4
        got = P.plane_through_points(*X)
        A = np.concatenate([X.astype(np.float64), np.ones((3, 1))], 1)
        n = np.linalg.svd(A)[2][-1]
        n = n / np.linalg.norm(n[:3])
        if n[3] < 0:
            n = -n
        np.testing.assert_allclose(got, n, rtol=2e-5, atol=2e-6)
        np.testing.assert_allclose(X.astype(np.float64) @ got[:3] + got[3], 0, atol=2e-5)


def test_support_points_by_hand():
    h, w = 10, 12
    costs = np.full((h, w), 2.0, np.float32)   # 2 = invalid everywhere
    conf = np.zeros((h, w), np.float32)
    costs[1, 2], conf[1, 2] = 0.05, 0.5        # block (0, 0): f = -0.45, f + 0.2 = -0.25 -> two points
    costs[3, 4], conf[3, 4] = 0.05, 0.5        # same block, later in (c, r) order, equal value: not taken
    costs[7, 6], conf[7, 6] = 0.25, 0.2        # block (5, 5): f = 0.05 < 0.1 kept; f + 0.2 = 0.25 dropped
    costs[6, 11], conf[6, 11] = 0.3, 0.0       # block (10, 5): 0.3 >= 0.1 -> nothing
    pts = P.support_points(costs, conf)
    assert [tuple(p) for p in pts] == [(2, 1), (2, 1), (6, 7)]


def test_raster_covers_the_triangle():
    tri = np.array([[0, 0, 20, 0, 0, 20]], np.int32)
    m = P.raster(tri, 32, 32)
    ys, xs = np.nonzero(m)
    assert set(np.unique(m)) <= {0.0, 1.0}
    assert np.all(xs + ys <= 20)
    for y in range(19):
        for x in range(19 - y):
            assert m[y, x] == 1.0, (x, y)
    # a second triangle overwrites shared pixels with its own index + 1
    m2 = P.raster(np.array([[0, 0, 20, 0, 0, 20], [0, 0, 10, 0, 0, 10]], np.int32), 32, 32)
    assert m2[2, 2] == 2.0 and m2[15, 2] == 1.0


def test_edge_map_step_and_flat():
    g = np.zeros((20, 30), np.uint8)
    g[:, 15:] = 200
    e = P.canny(g)
    cols = np.nonzero(e.any(0))[0]
    assert len(cols) == 1 and cols[0] in (14, 15)
    assert np.all(e[:, cols[0]] == 255)
    assert not P.canny(np.full((20, 30), 77, np.uint8)).any()


def test_edge_map_colour_takes_the_strongest_channel():
    """A colour image: equal channels give the grey map; an edge present in one
    channel only is found; the channel of largest |dx| + |dy| decides, so a
    weak edge in another channel does not hide it."""
    g = np.zeros((20, 30), np.uint8)
    g[:, 15:] = 200
    grey = P.canny(g)
    np.testing.assert_array_equal(P.canny(np.stack([g, g, g], -1)), grey)
    for c in range(3):
        rgb = np.zeros((20, 30, 3), np.uint8)
        rgb[..., c] = g
        rgb[..., (c + 1) % 3] = g // 20  # weak step (10 levels) below the low threshold on its own
        np.testing.assert_array_equal(P.canny(rgb), grey)
    # a horizontal step in one channel and a vertical one in another: both found
    rgb = np.zeros((40, 40, 3), np.uint8)
    rgb[:, 20:, 0] = 200
    rgb[20:, :, 2] = 200
    e = P.canny(rgb)
    assert e[5, 19:21].any() and e[19:21, 5].any()

