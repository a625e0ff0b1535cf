"""Planar prior built on the device (hpm_prior_support_points /
hpm_prior_build) against the host construction of the same steps
(GetSupportPoints_Classify_Check, DelaunayTriangulation, raster and
GetPriorPlaneParams of HPM.cpp:1026-1167 and main.cpp:326-431), bit for bit:

* distance 0, unit level: support points equal the host's; prior planes and
  triangle masks equal the host raster of the host triangulation with the
  host plane fit and the depth range test restated in numpy float32 (the same
  IEEE operations in the same order);
* distances 1 and 2: support points on the maps resized as cv::resize
  INTER_LINEAR (restated in numpy), including the scan of (int)(width * f)
  columns when the resized width rounds up;
* the whole pyramid schedule (prior passes at distances 2, 1, 0) writes
  byte-identical files with the prior built on the device and on the host.
"""
import math
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


def _resize_linear(src, dw, dh):
    """The host's ResizeLinear of a float map, in numpy (positions in float64,
    weights and interpolation in float32)."""
    sh, sw = src.shape

    def axis(n, sn, s):
        f = (np.arange(n, dtype=np.float64) + 0.5) * s - 0.5
        i = np.floor(f).astype(np.int64)
        w = (f - i).astype(np.float32)
        lo = i < 0
        i[lo], w[lo] = 0, 0
        hi = i >= sn - 1
        i[hi], w[hi] = sn - 1, 0
        return i, w

    xi, ax = axis(dw, sw, sw / dw)
    yi, ay = axis(dh, sh, sh / dh)
    xb, yb = np.minimum(xi + 1, sw - 1), np.minimum(yi + 1, sh - 1)
    a, b = src[yi][:, xi], src[yi][:, xb]
    c, d = src[yb][:, xi], src[yb][:, xb]
    top = a + ax[None, :] * (b - a)
    bot = c + ax[None, :] * (d - c)
    return (top + ay[:, None] * (bot - top)).astype(np.float32)


def _prior_context(seed, width, height):
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=seed, n_cams=4, width=width, height=height, f=0.82 * width, n_src=3)
    rng = np.random.default_rng(seed)
    depth = sc.depths[0].astype(np.float32)
    costs = np.where(rng.random(depth.shape) < 0.7, rng.random(depth.shape) * 0.15, 2.0).astype(np.float32)
    conf = (rng.random(depth.shape) * 0.5).astype(np.float32)
    normals = rng.normal(size=depth.shape + (3,)).astype(np.float32)
    normals /= np.linalg.norm(normals, axis=2, keepdims=True)
    planes = np.concatenate([normals, depth[..., None]], axis=2)
    ctx = H.Context(0)
    cams = [H.camera_from_scene(c) for c in sc.cameras]
    ctx.set_views(cams, [im.astype(np.float32) for im in sc.images])
    ctx.reload_hypotheses(planes, costs)
    ctx.set_confidence(conf)
    return ctx, cams[0], depth, costs, conf


def test_prior_distance0_equals_host_construction():
    from hpm_mi355x import pipeline as P
    W, H_ = 322, 242
    ctx, cam, depth, costs, conf = _prior_context(3, W, H_)
    print("case: device prior, distance 0, 322x242", flush=True)
    sp = ctx.prior_support_points(0)
    np.testing.assert_array_equal(sp, P.support_points(costs, conf))
    assert len(sp) > 1000
    tris = P.delaunay(W, H_, sp)
    ctx.prior_build(tris)
    got_planes, got_masks = ctx.prior()

    K = np.array(cam.K, np.float32)
    f32 = np.float32
    planes = np.zeros((len(tris), 4), np.float32)
    for t, v in enumerate(tris):
        X = []
        for k in range(3):
            x, y = int(v[2 * k]), int(v[2 * k + 1])
            d = depth[y, x]
            X.append([d * (f32(x) - K[2]) / K[0], d * (f32(y) - K[5]) / K[4], d])
        planes[t] = P.plane_through_points(*X)
    m = P.raster(tris, W, H_).astype(np.uint32)
    pl = planes[np.maximum(m, 1) - 1]
    xs = np.arange(W, dtype=np.float32)[None, :]
    ys = np.arange(H_, dtype=np.float32)[:, None]
    den = (xs - K[2]) * pl[..., 0] + ((K[0] / K[4]) * (ys - K[5])) * pl[..., 1] + K[0] * pl[..., 2]
    d = (-pl[..., 3] * K[0]) / den
    dmin, dmax = f32(cam.depth_min) * f32(0.6), f32(cam.depth_max) * f32(1.2)
    keep = (m > 0) & (d <= dmax) & (d >= dmin)
    want_masks = np.where(keep, m, 0).astype(np.uint32)
    want_planes = np.where(keep[..., None], pl, 0).astype(np.float32)
    print(f"triangles {len(tris)}, prior pixels {int(keep.sum())} of {W * H_}")
    assert keep.mean() > 0.5
    np.testing.assert_array_equal(got_masks, want_masks)
    np.testing.assert_array_equal(got_planes.view(np.uint32), want_planes.view(np.uint32))


@pytest.mark.parametrize("distance", [1, 2])
def test_prior_support_points_on_the_coarser_grid(distance):
    from hpm_mi355x import pipeline as P
    W, H_ = 322, 242
    ctx, cam, depth, costs, conf = _prior_context(5, W, H_)
    print(f"case: device prior support points, distance {distance}, 322x242", flush=True)
    factor = np.float32(1.0 / (distance * 2))
    # std::round of the float product: halves away from zero (np.round would
    # round them to even)
    hw, hh = (int(math.floor(float(np.float32(n) * factor) + 0.5)) for n in (W, H_))
    sw, sh = min(int(np.float32(W) * factor), hw), min(int(np.float32(H_) * factor), hh)
    if distance == 2:
        assert (hw, sw) == (81, 80)  # 322 / 4 = 80.5 rounds up; the scan stops at 80
    cds, fds = _resize_linear(costs, hw, hh), _resize_linear(conf, hw, hh)
    want = P.support_points(np.ascontiguousarray(cds[:sh, :sw]), np.ascontiguousarray(fds[:sh, :sw]))
    np.testing.assert_array_equal(ctx.prior_support_points(distance), want)
    # the triangulation and the rest run on that grid
    ctx.prior_build(P.delaunay(hw, hh, want))
    _, masks = ctx.prior()
    assert set(np.unique(masks).tolist()) <= {0, 1}
    assert masks.mean() > 0.5


def test_prior_build_rejects_bad_input():
    import hpm_mi355x as H
    from hpm_mi355x import pipeline as P
    ctx, cam, depth, costs, conf = _prior_context(7, 96, 72)
    with pytest.raises(H.HpmError, match="support_points first"):
        ctx.prior_build(np.zeros((1, 6), np.int32))
    sp = ctx.prior_support_points(0)
    with pytest.raises(H.HpmError, match="outside"):
        ctx.prior_build(np.array([[0, 0, 96, 0, 0, 10]], np.int32))
    ctx.prior_build(P.delaunay(96, 72, sp))


def test_pyramid_files_equal_with_host_prior(tmp_path):
    """The pyramid schedule on 320x240 with a size bound of 100 (levels of
    80x60, 160x120, 320x240; prior passes at distances 2, 1 and 0 on the
    finest level) writes byte-identical files with the prior built on the
    device and on the host."""
    import filecmp
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=4, n_cams=5, width=320, height=240, f=262.5, n_src=4)
    runs = {}
    try:
        for on_device in (True, False):
            folder = str(tmp_path / ("device" if on_device else "host"))
            scenes.write_dense_folder(sc, folder)
            P.set_prior_on_device(on_device)
            print(f"case: pyramid pipeline 5 views 320x240 bound 100, prior on device {on_device}", flush=True)
            runs[on_device] = (folder, P.run_pyramid_pipeline(folder, initial_image_scale=1, size_bound=100,
                                                              precision="fp16", seed=9))
    finally:
        P.set_prior_on_device(True)
    assert runs[True][1] == runs[False][1] > 4 * 5000
    names = []
    for root, _, files in os.walk(os.path.join(runs[True][0], "HPM_MVS_plusplus")):
        for f in files:
            names.append(os.path.relpath(os.path.join(root, f), runs[True][0]))
    assert len(names) > 5 * 6
    for rel in names:
        assert filecmp.cmp(os.path.join(runs[True][0], rel), os.path.join(runs[False][0], rel), shallow=False), rel
