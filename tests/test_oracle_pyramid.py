"""Pins of the oracle's multi-scale pieces (oracle/oracle_pipeline.py,
oracle/oracle_host.cpp) by closed forms and special cases:

* the schedule helpers: ComputeMultiScaleSettings (main.cpp:202-235), the
  level sizes of PrepareScaleImages (main.cpp:125-133) and the prior
  distances of each level (main.cpp:1254-1262);
* the library's resize stand-in: constant maps, exact halving of a ramp,
  8-bit rounding half up;
* the coarse planar prior (main.cpp:443-541): a fronto-parallel plane at
  depth Z seen through intrinsics scaled by the factor gives depth Z and the
  camera's viewing axis as world normal inside the triangles, nothing
  outside them or outside [0.6 GetMinDepth, 1.2 GetMaxDepth];
* the per-pixel reload (main.cpp:553-574): the plane (R n, -(R n) . X) of
  the back-projected pixel X, and no prior outside [GetMinDepth, GetMaxDepth].
"""
import ctypes
import math

import numpy as np

import oracle_py as O
import oracle_pipeline as OP


def test_multiscale_settings_closed_forms():
    # 320 px halves twice to reach 80; 2000 px once to reach 1000 (not > 1000)
    assert OP.multiscale_settings([320, 2000, 1000, 1001], 1, 80)[1][0] == 2
    assert OP.multiscale_settings([2000], 1, 1000) == ([2000], [1])
    assert OP.multiscale_settings([1000, 1001], 1, 1000) == ([1000, 1001], [0, 1])
    # initial max side: integer division with the half added, never below 1
    assert OP.multiscale_settings([3073, 3070, 3], 4, 1000) == ([768, 768, 1], [0, 0, 0])
    assert OP.multiscale_settings([3074], 4, 1000)[0] == [769]


def test_level_size_closed_forms():
    assert OP.level_size(320, 240, 320) == (320, 240)  # fits: unchanged
    assert OP.level_size(320, 240, 160) == (160, 120)
    assert OP.level_size(320, 240, 80) == (80, 60)
    assert OP.level_size(240, 320, 80) == (60, 80)
    # the float factor 81 / 322 times 322 lands on 81 (rounded, not truncated)
    assert OP.level_size(322, 242, 81) == (81, 61)
    assert OP.level_size(5, 3, 1) == (1, 1)  # never below one texel


def test_prior_distances_of_a_three_level_pyramid():
    # levels 2, 1, 0 of a view with 2 halvings: num_downscale after the
    # level's decrement is 1, 0, -1
    assert OP.prior_distances(2, 2, [1]) == [[0]]
    assert OP.prior_distances(2, 1, [0]) == [[1], [0]]
    assert OP.prior_distances(2, 0, [-1]) == [[2], [1], [0]]
    # a view needing one halving fewer (its num_downscale is one lower) gets a
    # distance one higher; one needing more halvings than the level has left
    # gets -1, for which the prior pass runs nothing
    assert OP.prior_distances(2, 2, [1, 0]) == [[0, 1]]
    assert OP.prior_distances(2, 2, [1, 2]) == [[0, -1]]


def test_resize_linear_closed_forms():
    c = np.full((24, 32), 7.25, np.float32)
    for (w, h) in ((16, 12), (8, 6), (64, 48), (13, 7)):
        r = OP.resize_linear(c, w, h)
        assert r.shape == (h, w) and r.flags.c_contiguous and np.all(r == np.float32(7.25))
    # a ramp along x halved: destination d samples position 2 d + 0.5,
    # except the last column, whose left texel is the border (weight 0)
    ramp = np.tile(np.arange(32, dtype=np.float32), (4, 1))
    r = OP.resize_linear(ramp, 16, 4)
    want = 2 * np.arange(16, dtype=np.float32) + np.float32(0.5)
    np.testing.assert_array_equal(r[0], want)
    # doubling: destination d samples (d + 0.5) / 2 - 0.5, clamped at 0 and 31
    r = OP.resize_linear(ramp, 64, 4)
    pos = np.clip((np.arange(64) + 0.5) / 2 - 0.5, 0, 31).astype(np.float32)
    np.testing.assert_array_equal(r[0], pos)
    # 8 bits: 2 x 2 means of (0, 1) pairs are .5 and round up; (0, 0) stays 0
    u = np.zeros((2, 4), np.uint8)
    u[:, 1] = 1
    r = OP.resize_linear(u, 2, 1, round_u8=True)
    assert r.dtype == np.uint8 and r.tolist() == [[1, 0]]


def _rotation(ax, ay, az):
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return (Rz @ Ry @ Rx).astype(np.float32)


def _cam(w, h, R, dmin=1.0, dmax=5.0):
    K = np.array([[200.0, 0, w / 2], [0, 180.0, h / 2], [0, 0, 1]], np.float32)
    return O.camera(K, R, [0.1, -0.2, 0.3], w, h, dmin, dmax)


def _coarse(cam, depth, tris, factor):
    hh, hw = depth.shape
    pd = np.zeros((hh, hw), np.float32)
    pn = np.zeros((hh, hw, 4), np.float32)
    t = np.ascontiguousarray(tris, np.int32)
    O.lib().orc_prior_coarse(ctypes.byref(cam), O._p(np.ascontiguousarray(depth, np.float32)), hw, hh, O._p(t),
                             len(t), float(factor), O._p(pd), O._p(pn))
    return pd, pn


def test_prior_coarse_fronto_parallel_plane():
    R = _rotation(0.3, -0.2, 0.1)
    cam = _cam(80, 60, R)  # the view's own size; the coarse grid is 40 x 30 at factor 1/2
    Z = np.float32(2.5)
    depth = np.full((30, 40), Z, np.float32)
    tris = [[2, 3, 30, 4, 5, 25]]  # one triangle inside the coarse grid
    pd, pn = _coarse(cam, depth, tris, np.float32(0.5))
    inside = pd > 0
    assert 200 < inside.sum() < 40 * 30
    # every covered pixel: depth Z (to an ulp of the float division), world
    # normal R^T (0, 0, -1) (the plane n . X + w = 0 with w >= 0)
    np.testing.assert_allclose(pd[inside], Z, rtol=2e-7)
    want_n = (R.T.astype(np.float64) @ np.array([0.0, 0.0, -1.0])).astype(np.float32)
    np.testing.assert_allclose(pn[inside][:, :3], np.broadcast_to(want_n, (int(inside.sum()), 3)), atol=2e-6)
    assert np.all(pn[~inside] == 0) and np.all(pn[..., 3] == 0)
    # the raster of the triangle: its first corner and centroid covered, the
    # grid corners outside it not
    assert inside[3, 2] and inside[10, 12] and not inside[29, 39] and not inside[0, 39] and not inside[29, 0]


def test_prior_coarse_depth_bounds():
    """Kept when Z lies in [0.6 (0.6 d_min), 1.2 (1.2 d_max)] (the widened
    range of GetMinDepth / GetMaxDepth widened again), dropped outside."""
    R = np.eye(3, dtype=np.float32)
    cam = _cam(80, 60, R, dmin=1.0, dmax=5.0)
    tris = [[0, 0, 39, 0, 0, 29]]
    for Z, kept in ((0.37, True), (0.35, False), (7.19, True), (7.21, False)):
        pd, _ = _coarse(cam, np.full((30, 40), Z, np.float32), tris, np.float32(0.5))
        assert (pd > 0).any() == kept, Z


def test_prior_reload_closed_form():
    R = _rotation(-0.4, 0.25, 0.7)
    w, h = 32, 24
    cam = _cam(w, h, R, dmin=1.0, dmax=5.0)
    n_world = np.array([0.2, -0.3, -0.9], np.float32)
    n_world /= np.linalg.norm(n_world)
    up_d = np.full((h, w), 3.0, np.float32)
    up_d[0, :] = 0.5   # below GetMinDepth (0.6): no prior
    up_d[1, :] = 6.1   # above GetMaxDepth (6.0): no prior
    up_n = np.zeros((h, w, 4), np.float32)
    up_n[..., :3] = n_world
    prior = np.zeros((h, w, 4), np.float32)
    masks = np.zeros((h, w), np.uint32)
    O.lib().orc_prior_reload(ctypes.byref(cam), O._p(up_d), O._p(up_n), O._p(prior), O._p(masks))
    assert masks[0].sum() == 0 and masks[1].sum() == 0 and np.all(prior[:2] == 0)
    assert np.all(masks[2:] == 1)
    n_cam = R.astype(np.float64) @ n_world.astype(np.float64)
    np.testing.assert_allclose(prior[2:, :, :3], np.broadcast_to(n_cam, (h - 2, w, 3)), atol=1e-6)
    K = np.array(cam.K, np.float64).reshape(3, 3)
    ys, xs = np.mgrid[2:h, 0:w].astype(np.float64)
    X = np.stack([3.0 * (xs - K[0, 2]) / K[0, 0], 3.0 * (ys - K[1, 2]) / K[1, 1], np.full_like(xs, 3.0)], -1)
    np.testing.assert_allclose(prior[2:, :, 3], -(X @ n_cam), atol=2e-5)
    # the plane passes through the back-projected pixel: n . X + w = 0
    np.testing.assert_allclose(np.einsum("yxk,k->yx", X, n_cam) + prior[2:, :, 3], 0.0, atol=2e-5)


def test_prior_coarse_slanted_plane_uses_the_scaled_intrinsics():
    """A slanted plane rendered as a depth map through intrinsics scaled by
    the factor: the prior reproduces that depth map and the plane's world
    normal.  Back-projecting with unscaled intrinsics would fit another plane."""
    R = _rotation(0.1, 0.2, -0.3)
    cam = _cam(80, 60, R)
    f = np.float64(0.5)
    K = np.array(cam.K, np.float64).reshape(3, 3)
    n = np.array([0.25, -0.15, -1.0])
    n /= np.linalg.norm(n)
    w = 2.5
    ys, xs = np.mgrid[0:30, 0:40].astype(np.float64)
    # depth where the ray of coarse pixel (x, y) meets n . X + w = 0
    ray = np.stack([(xs - K[0, 2] * f) / (K[0, 0] * f), (ys - K[1, 2] * f) / (K[1, 1] * f), np.ones_like(xs)], -1)
    depth = (-w / (ray @ n)).astype(np.float32)
    pd, pn = _coarse(cam, depth, [[1, 1, 38, 2, 3, 28]], np.float32(0.5))
    inside = pd > 0
    assert inside.sum() > 300
    np.testing.assert_allclose(pd[inside], depth[inside], rtol=3e-6)
    want_n = R.T.astype(np.float64) @ n
    np.testing.assert_allclose(pn[inside][:, :3], np.broadcast_to(want_n, (int(inside.sum()), 3)), atol=3e-6)


def test_support_points_scan_equals_the_scan_of_the_submap():
    """The strided scan of the first w x h of a wider map is the scan of that
    part copied out (the distance-0 function, pinned in test_oracle_pins)."""
    rng = np.random.default_rng(3)
    costs = np.where(rng.random((37, 53)) < 0.7, rng.random((37, 53)) * 0.2, 2.0).astype(np.float32)
    conf = (rng.random((37, 53)) * 0.3).astype(np.float32)
    sw, sh = 47, 33

    def scan(c, f, stride, w, h):
        xy = np.zeros((2 * ((w + 4) // 5) * ((h + 4) // 5), 2), np.int32)
        cnt = ctypes.c_int(0)
        if stride is None:
            O.lib().orc_support_points(O._p(c), O._p(f), w, h, O._p(xy), len(xy), ctypes.byref(cnt))
        else:
            O.lib().orc_support_points_scan(O._p(c), O._p(f), stride, w, h, O._p(xy), len(xy), ctypes.byref(cnt))
        return xy[:cnt.value]

    full = scan(costs, conf, 53, sw, sh)
    sub = scan(np.ascontiguousarray(costs[:sh, :sw]), np.ascontiguousarray(conf[:sh, :sw]), None, sw, sh)
    assert len(full) > 20
    np.testing.assert_array_equal(full, sub)
    assert full[:, 0].max() < sw and full[:, 1].max() < sh
