"""Forced device-allocation failures.  With HPM_FAULT_ALLOC_AFTER=n in the
environment when a context is created, that context's device allocation
number n (from 0) fails once, as if the device were out of memory.  Each flow
below is run once for every n up to the first that no allocation reaches:

* the call that meets the failure returns HPM_E_NOMEM (-4) with a message;
* it leaves no half-built state behind: the same call made again completes
  the flow;
* the flow's results are bit-identical to those of a run without faults.

Flows: a photometric FP16 run (column-pair images), a geometric FP32 pass (source
depths), a pooled FP16 run (the pool's shared pair images), the planar prior
at distance 1 (downsampled maps, scratch, prior buffers) and Gauss-Newton
refinement on smoothed views (smoothed images and pair images, gradient buffers)."""
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

SEED = 17
MAX_ALLOCS = 64


def _context(n):
    import hpm_mi355x as H
    if n is None:
        return H.Context(0)
    os.environ["HPM_FAULT_ALLOC_AFTER"] = str(n)
    try:
        return H.Context(0)
    finally:
        del os.environ["HPM_FAULT_ALLOC_AFTER"]


def _sweep(name, flow):
    """flow(ctx, call) -> tuple of arrays; call(what, f, *args) makes one API
    call and, when it fails with HPM_E_NOMEM, records it and calls again."""
    import hpm_mi355x as H
    clean = flow(_context(None), lambda what, f, *a: f(*a))
    hit = []
    for n in range(MAX_ALLOCS):
        failed = []

        def call(what, f, *a):
            try:
                return f(*a)
            except H.HpmError as e:
                assert "failed (-4)" in str(e), str(e)
                assert str(e).split("(-4):", 1)[1].strip(), "no message"
                failed.append(what)
                return f(*a)

        print(f"case: {name}, allocation {n} fails", flush=True)
        got = flow(_context(n), call)
        for a, b in zip(got, clean):
            assert a.dtype == b.dtype and a.shape == b.shape
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (name, n, failed)
        if not failed:
            break
        assert len(failed) == 1, failed
        hit.append(failed[0])
    else:
        pytest.fail(f"{name}: more than {MAX_ALLOCS} allocations")
    print(f"{name}: {len(hit)} allocations failed in turn, in {sorted(set(hit))}")
    return hit


def test_photometric_fp16_alloc_faults(small_scene):
    import hpm_mi355x as H
    _, _, scams, imgs = small_scene
    cams = [H.camera_from_scene(c) for c in scams]

    def flow(ctx, call):
        call("set_params", ctx.set_params, H.default_params(precision="fp16"))
        call("set_views", ctx.set_views, cams, imgs)
        call("run", ctx.run, SEED)
        out = (ctx.planes(), ctx.costs())
        ctx.close()
        return out

    hit = _sweep("photometric fp16", flow)
    assert set(hit) == {"set_views", "run"}  # views and pixel state; the pair images


def test_geometric_fp32_alloc_faults(small_scene):
    import hpm_mi355x as H
    sc, ids, scams, imgs = small_scene
    cams = [H.camera_from_scene(c) for c in scams]
    depths = [None] + [sc.depths[i] * np.float32(1.003) for i in ids[1:]]
    first = _context(None)
    first.set_params(H.default_params())
    first.set_views(cams, imgs)
    first.run(SEED)
    planes, costs = first.planes(), first.costs()
    first.close()

    def flow(ctx, call):
        call("set_params", ctx.set_params, H.default_params(geom_consistency=1))
        call("set_views", ctx.set_views, cams, imgs)
        call("set_src_depths", ctx.set_src_depths, depths)
        call("reload", ctx.reload_hypotheses, planes, costs)
        call("run", ctx.run, SEED + 1)
        out = (ctx.planes(), ctx.costs(), ctx.selected_views())
        ctx.close()
        return out

    hit = _sweep("geometric fp32", flow)
    assert hit.count("set_src_depths") == len(ids) - 1


def test_pooled_fp16_alloc_faults(small_scene):
    import hpm_mi355x as H
    _, _, scams, imgs = small_scene
    cams = [H.camera_from_scene(c) for c in scams]
    pool = H.ImagePool(cams, imgs)

    def flow(ctx, call):
        call("set_params", ctx.set_params, H.default_params(precision="fp16"))
        call("set_views_pooled", ctx.set_views_pooled, pool, list(range(len(cams))))
        call("run", ctx.run, SEED)
        out = (ctx.planes(), ctx.costs())
        ctx.close()
        return out

    # the pool's pair images are built by the first context that runs at
    # this precision: a fresh pool for each case, so that every case builds
    # them (and can meet the failure there)
    def fresh(ctx, call):
        nonlocal pool
        pool.close()
        pool = H.ImagePool(cams, imgs)
        return flow(ctx, call)

    hit = _sweep("pooled fp16", fresh)
    assert hit.count("run") == len(cams)
    pool.close()


def test_prior_alloc_faults():
    import hpm_mi355x as H
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    W, H_ = 162, 122
    sc = scenes.make_plane_scene(seed=5, n_cams=4, width=W, height=H_, f=0.82 * W, n_src=3)
    rng = np.random.default_rng(5)
    depth = sc.depths[0].astype(np.float32)
    costs = np.where(rng.random(depth.shape) < 0.7, rng.random(depth.shape) * 0.15, 2.0).astype(np.float32)
    conf = (rng.random(depth.shape) * 0.5).astype(np.float32)
    normals = rng.normal(size=depth.shape + (3,)).astype(np.float32)
    normals /= np.linalg.norm(normals, axis=2, keepdims=True)
    planes = np.concatenate([normals, depth[..., None]], axis=2)
    cams = [H.camera_from_scene(c) for c in sc.cameras]
    imgs = [im.astype(np.float32) for im in sc.images]
    tris = {}

    def flow(ctx, call):
        call("set_views", ctx.set_views, cams, imgs)
        call("reload", ctx.reload_hypotheses, planes, costs)
        call("set_confidence", ctx.set_confidence, conf)
        sp = call("support_points", ctx.prior_support_points, 1)
        if "t" not in tris:
            tris["t"] = P.delaunay(81, 61, sp)  # the distance-1 grid: round(162 / 2) x round(122 / 2)
        call("prior_build", ctx.prior_build, tris["t"])
        pp, pm = ctx.prior()
        ctx.close()
        return sp, pp, pm

    hit = _sweep("prior distance 1", flow)
    assert {"support_points", "prior_build"} <= set(hit)


def test_smoothed_refine_alloc_faults(small_scene):
    import hpm_mi355x as H
    sc, ids, scams, imgs = small_scene
    cams = [H.camera_from_scene(c) for c in scams]
    # (depth, theta, phi) of the ground-truth plane, 2 % and 0.05 rad off
    gt = sc.depths[ids[0]].astype(np.float64)
    nc = sc.normals[ids[0]].astype(np.float64) @ np.asarray(cams[0].R, np.float64).reshape(3, 3).T
    rng = np.random.default_rng(3)
    params = np.stack([gt * (1 + rng.normal(0, 0.02, gt.shape)),
                       np.arccos(np.clip(-nc[..., 2], -1, 1)) + rng.normal(0, 0.05, gt.shape),
                       np.arctan2(nc[..., 1], nc[..., 0]) + rng.normal(0, 0.05, gt.shape)], -1).astype(np.float32)
    weights = np.ones(gt.shape + (len(cams) - 1,), np.float32)

    def flow(ctx, call):
        call("set_params", ctx.set_params, H.default_params(precision="fp32fast"))
        call("set_views", ctx.set_views, cams, imgs)
        pp, cost = call("refine", ctx.refine, params, weights, 4, "gn_smoothed")
        ctx.close()
        return pp, cost

    hit = _sweep("refine gn_smoothed", flow)
    assert hit.count("refine") >= 2 * len(cams)
