"""Pins the FP64 oracle of the plane-cost gradient: the analytic gradient of
the aggregated bilateral NCC with respect to (depth, theta, phi) against
central finite differences of the FP64 aggregated cost.  The cost is piecewise
smooth (bilinear cells, the clamp), so the bounds are set on the bulk of the
pixels: with h = 1e-5 (relative in depth, radians in the angles) the two agree
to 1e-4 relative on >= 95 % of the pixels; with the coarser h = 1e-3 relative
in depth the median relative difference stays below 2e-2."""
import numpy as np

import oracle_py as O


def _problem():
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=6, n_cams=5, width=64, height=48, f=52.5, n_src=4)
    ids, cams, imgs = scenes.reference_view_problem(sc, 2, 4)
    oc = [O.camera_from_scene(c) for c in cams]
    prm = O.params(len(oc), oc[0])
    gt = sc.depths[ids[0]]
    # ground-truth normals in the reference frame -> angles; perturbed a little
    R = np.asarray(cams[0].R, np.float64)
    nw = sc.normals[ids[0]].astype(np.float64)
    nc = nw @ R.T
    th = np.arccos(np.clip(-nc[..., 2], -1, 1))
    ph = np.arctan2(nc[..., 1], nc[..., 0])
    rng = np.random.default_rng(0)
    params = np.stack([gt * (1 + rng.normal(0, 0.01, gt.shape)), th + rng.normal(0, 0.05, gt.shape),
                       ph + rng.normal(0, 0.05, gt.shape)], -1).astype(np.float32)
    weights = np.ones(gt.shape + (4,), np.float32)
    return oc, imgs, prm, params, weights


def test_oracle_gradient_matches_central_differences():
    oc, imgs, prm, params, weights = _problem()
    cost, grad = O.cost_grad(oc, imgs, prm, params, weights)
    base = params.astype(np.float64)
    c0 = O.cost_params(oc, imgs, prm, base, weights)
    np.testing.assert_allclose(c0, cost, rtol=1e-12, atol=1e-12)
    live = (cost < 1.5) & (cost > 0.02)
    assert live.mean() > 0.5
    def fd(k, h):
        hi, lo = base.copy(), base.copy()
        hi[..., k] += h
        lo[..., k] -= h
        return (O.cost_params(oc, imgs, prm, hi, weights) - O.cost_params(oc, imgs, prm, lo, weights)) / (2 * h)

    for k in range(3):
        h = 1e-5 * base[..., 0] if k == 0 else np.full(base.shape[:2], 1e-5)
        g = grad[..., k]
        rel = (np.abs(fd(k, h) - g) / (np.abs(g) + 1e-3 * np.median(np.abs(g[live]))))[live]
        print(f"param {k}: h=1e-5 median rel {np.median(rel):.2e}, q95 {np.quantile(rel, 0.95):.2e}")
        assert np.quantile(rel, 0.95) <= 1e-4
    g = grad[..., 0]
    live &= g != 0  # a zero depth derivative (a window on flat texture) has no relative error
    rel = np.abs(fd(0, 1e-3 * base[..., 0])[live] - g[live]) / np.abs(g[live])
    print(f"depth: h=1e-3 relative, median rel {np.median(rel):.2e}")
    assert np.median(rel) <= 2e-2
