"""Pins the oracle's cost core (ComputeHomography, bilinear sampling,
ComputeBilateralNCC, the top-k initial cost) on the CPU with closed forms and
an independent brute-force evaluation, so that a dropped term, a sign error or
a transposed tap index fails a test.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_py as O

W, H = 48, 40
K = np.array([[60.0, 0, 23.5], [0, 60.0, 19.0], [0, 0, 1]])


def cam(R=np.eye(3), t=np.zeros(3), w=W, h=H, Kmat=K):
    return O.camera(Kmat, R, t, w, h, 1.0, 5.0)


def rot(ax, deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    if ax == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def texture(seed, w=W, h=H):
    rng = np.random.default_rng(seed)
    base = rng.uniform(0, 255, (h // 4 + 2, w // 4 + 2))
    # smooth-ish random texture by bilinear up-sampling
    ys = np.linspace(0, base.shape[0] - 1.001, h)
    xs = np.linspace(0, base.shape[1] - 1.001, w)
    y0 = ys.astype(int); x0 = xs.astype(int)
    fy = (ys - y0)[:, None]; fx = (xs - x0)[None, :]
    img = (base[y0][:, x0] * (1 - fx) * (1 - fy) + base[y0][:, x0 + 1] * fx * (1 - fy)
           + base[y0 + 1][:, x0] * (1 - fx) * fy + base[y0 + 1][:, x0 + 1] * fx * fy)
    return np.ascontiguousarray(img, dtype=np.float32)


def fronto_planes(depth, w=W, h=H, Kmat=K):
    """n = (0, 0, 1), d = -depth everywhere (n.X + d = 0)."""
    p = np.zeros((h, w, 4), np.float32)
    p[..., 2] = 1.0
    p[..., 3] = -depth
    return p


def prm(n):
    p = O.orc_params()
    p.patch_size, p.radius_increment, p.sigma_spatial, p.sigma_color, p.top_k = 11, 2, 5.0, 3.0, 4
    p.num_images, p.depth_min, p.depth_max, p.num_iterations = n, 0.6, 6.0, 3
    return p


# ------------------------------------------------------------------ bilinear
def test_bilinear_exact_on_affine_ramp():
    L = O.lib()
    yy, xx = np.mgrid[0:H, 0:W]
    img = (3.0 * xx - 2.0 * yy + 7.0).astype(np.float32)
    rng = np.random.default_rng(0)
    for _ in range(200):
        u, v = rng.uniform(0, W - 1), rng.uniform(0, H - 1)
        got = L.orc_bilinear_f64(img.ctypes.data, W, H, u, v)
        assert abs(got - (3 * u - 2 * v + 7)) < 1e-9


def test_bilinear_border_clamp_and_texel_centres():
    L = O.lib()
    img = texture(3)
    # texel centres are integers
    for (x, y) in [(0, 0), (5, 7), (W - 1, H - 1)]:
        assert L.orc_bilinear(img.ctypes.data, W, H, float(x), float(y)) == img[y, x]
    # outside the image the nearest border texels are used
    assert L.orc_bilinear_f64(img.ctypes.data, W, H, -3.0, 4.0) == img[4, 0]
    assert L.orc_bilinear_f64(img.ctypes.data, W, H, W + 2.5, 4.0) == img[4, W - 1]
    got = L.orc_bilinear_f64(img.ctypes.data, W, H, -0.5, -7.0)
    assert got == img[0, 0]


# ------------------------------------------------------------------ homography
def _H(rc, sc, plane):
    out = np.zeros(9)
    pl = np.asarray(plane, np.float64)
    O.lib().orc_homography_f64(ctypes.byref(rc), ctypes.byref(sc), pl.ctypes.data, out.ctypes.data)
    return out.reshape(3, 3)


def test_homography_matches_world_reprojection():
    rng = np.random.default_rng(1)
    Rr, tr = rot("y", 4), np.array([0.1, -0.05, 0.3])
    Rs, ts = rot("y", -6) @ rot("x", 3), np.array([-0.3, 0.02, 0.25])
    rc, sc = cam(Rr, tr), cam(Rs, ts)
    Rr32 = np.array(rc.R[:], np.float64).reshape(3, 3); tr32 = np.array(rc.t[:], np.float64)
    Rs32 = np.array(sc.R[:], np.float64).reshape(3, 3); ts32 = np.array(sc.t[:], np.float64)
    K32 = np.array(rc.K[:], np.float64).reshape(3, 3)
    for _ in range(50):
        n = rng.normal(size=3); n[2] = -abs(n[2]) - 1.0; n /= np.linalg.norm(n)
        px, py = rng.uniform(0, W), rng.uniform(0, H)
        depth = rng.uniform(1.5, 4.0)
        X = depth * np.linalg.solve(K32, [px, py, 1.0])
        d = -n @ X
        Hm = _H(rc, sc, [n[0], n[1], n[2], d])
        q = Hm @ [px, py, 1.0]
        Xw = Rr32.T @ (X - tr32)
        xs = K32 @ (Rs32 @ Xw + ts32)
        assert np.allclose(q[:2] / q[2], xs[:2] / xs[2], atol=1e-7)


def test_homography_pure_translation_shift():
    # src centre = ref centre + (b, 0, 0); fronto-parallel plane at depth z -> u_src = u - fx b / z
    b, z = 0.2, 3.0
    rc, sc = cam(), cam(t=np.array([-b, 0.0, 0.0]))
    Hm = _H(rc, sc, [0.0, 0.0, 1.0, -z])
    for (u, v) in [(3.0, 4.0), (20.0, 30.0)]:
        q = Hm @ [u, v, 1.0]
        b32 = float(np.float32(b))  # the camera stores float32 (main.h:39-47)
        assert abs(q[0] / q[2] - (u - 60.0 * b32 / z)) < 1e-9
        assert abs(q[1] / q[2] - v) < 1e-9


def test_plane_depth_round_trip():
    L = O.lib()
    rc = cam()
    rng = np.random.default_rng(2)
    for _ in range(50):
        n = rng.normal(size=3); n[2] = -abs(n[2]) - 0.5; n /= np.linalg.norm(n)
        px, py, depth = int(rng.integers(0, W)), int(rng.integers(0, H)), rng.uniform(1, 5)
        d = L.orc_distance_to_origin_f64(ctypes.byref(rc), px, py, depth, n.ctypes.data)
        back = L.orc_depth_from_plane_f64(ctypes.byref(rc), np.array([n[0], n[1], n[2], d]).ctypes.data, px, py)
        assert abs(back - depth) < 1e-5 * depth


# ------------------------------------------------------------------ NCC closed forms
def _costs(images, cams, planes, dtype=np.float64):
    return O.cost_fixed(cams, images, prm(len(cams)), planes, dtype)


def test_ncc_identical_views_is_zero():
    ref = texture(4)
    c = _costs([ref, ref], [cam(), cam()], fronto_planes(2.0))
    inner = c[6:-6, 6:-6, 0]
    assert np.all(inner < 1e-9)


@pytest.mark.parametrize("a,b,expect", [(0.5, 30.0, 0.0), (2.0, -100.0, 0.0), (-1.0, 255.0, 2.0)])
def test_ncc_affine_invariance(a, b, expect):
    ref = texture(5)
    c = _costs([ref, (a * ref + b).astype(np.float32)], [cam(), cam()], fronto_planes(2.0))
    assert np.all(np.abs(c[6:-6, 6:-6, 0] - expect) < 1e-7)


def test_ncc_flat_window_and_outside_centre():
    ref = texture(6)
    flat = np.full_like(ref, 100.0)
    c = _costs([flat, ref], [cam(), cam()], fronto_planes(2.0))
    assert np.all(c == 2.0)
    # source camera shifted far to the side: every centre projects outside
    c = _costs([ref, ref], [cam(), cam(t=np.array([-10.0, 0, 0]))], fronto_planes(2.0))
    assert np.all(c == 2.0)


def test_ncc_translation_minimum_at_true_depth():
    # src = ref shifted by 4 px to the left <=> camera baseline b with fx b / z = 4
    ref = texture(7)
    src = np.empty_like(ref)
    src[:, :-4] = ref[:, 4:]
    src[:, -4:] = ref[:, -1:]
    b = 0.2
    z_true = 60.0 * b / 4.0
    cams = [cam(), cam(t=np.array([-b, 0.0, 0.0]))]
    costs = {z: _costs([ref, src], cams, fronto_planes(z))[8:-8, 10:-12, 0].mean() for z in (z_true * 0.8, z_true, z_true * 1.25)}
    assert costs[z_true] < 1e-7
    assert costs[z_true * 0.8] > 0.05 and costs[z_true * 1.25] > 0.05


def _brute_ncc(ref, src, Hm, px, py, ss=5.0, sc=3.0):
    """Independent FP64 evaluation of the bilateral NCC of HPM.cu:375-448."""
    def tex(img, x, y):
        return float(img[min(max(y, 0), img.shape[0] - 1), min(max(x, 0), img.shape[1] - 1)])

    def bil(img, u, v):
        x0, y0 = math.floor(u), math.floor(v)
        a, b = u - x0, v - y0
        return ((1 - a) * (1 - b) * tex(img, x0, y0) + a * (1 - b) * tex(img, x0 + 1, y0)
                + (1 - a) * b * tex(img, x0, y0 + 1) + a * b * tex(img, x0 + 1, y0 + 1))

    q = Hm @ [px, py, 1.0]
    if not (0 <= q[0] / q[2] < src.shape[1] and 0 <= q[1] / q[2] < src.shape[0]):
        return 2.0
    c = tex(ref, px, py)
    ws, rs, ss_, = [], [], []
    for dx in range(-5, 6, 2):
        for dy in range(-5, 6, 2):
            r = tex(ref, px + dx, py + dy)
            qq = Hm @ [px + dx, py + dy, 1.0]
            s = bil(src, qq[0] / qq[2], qq[1] / qq[2])
            w = math.exp(-math.hypot(dx, dy) / (2 * ss * ss) - abs(r - c) / (2 * sc * sc))
            ws.append(w); rs.append(r); ss_.append(s)
    w, r, s = map(np.array, (ws, rs, ss_))
    W_ = w.sum()
    mr, ms = (w * r).sum() / W_, (w * s).sum() / W_
    vr = (w * r * r).sum() / W_ - mr * mr
    vs = (w * s * s).sum() / W_ - ms * ms
    if vr < 1e-5 or vs < 1e-5:
        return 2.0
    cov = (w * r * s).sum() / W_ - mr * ms
    return max(0.0, min(2.0, 1 - cov / math.sqrt(vr * vs)))


def test_ncc_brute_force_general_cameras():
    rng = np.random.default_rng(8)
    ref, src = texture(9), texture(10)
    src = (0.6 * src + 0.4 * ref).astype(np.float32)
    rc = cam(rot("y", 2), np.array([0.05, 0.0, 0.1]))
    sc = cam(rot("y", -5) @ rot("x", 2), np.array([-0.25, 0.03, 0.1]))
    planes = np.zeros((H, W, 4), np.float32)
    for y in range(H):
        for x in range(W):
            n = rng.normal(size=3); n[2] = -abs(n[2]) - 1.5; n /= np.linalg.norm(n)
            depth = rng.uniform(1.5, 4.0)
            X = depth * np.array([(x - 23.5) / 60.0, (y - 19.0) / 60.0, 1.0])
            planes[y, x] = [n[0], n[1], n[2], -n @ X]
    c = _costs([ref, src], [rc, sc], planes)
    for (x, y) in [(3, 2), (20, 15), (30, 33), (45, 5), (11, 38), (24, 20)]:
        Hm = _H(rc, sc, planes[y, x].astype(np.float64))
        assert abs(c[y, x, 0] - _brute_ncc(ref, src, Hm, x, y)) < 1e-9


def test_init_cost_topk_and_mask():
    """ComputeMultiViewInitialCostandSelectedViews: mean of the k smallest valid
    costs and the mask of views at or below the k-th (HPM.cu:450-488)."""
    rng = np.random.default_rng(13)
    ref = texture(14)
    srcs = [(ref * a + rng.normal(0, s, ref.shape)).astype(np.float32)
            for a, s in ((1.0, 5), (1.0, 30), (-1.0, 5), (1.0, 60), (1.0, 15), (0.5, 10))]
    cams = [cam()] * 7
    planes = fronto_planes(2.0)
    costs = O.cost_fixed(cams, [ref] + srcs, prm(7), planes, np.float32)
    ic, mask = O.init_cost(cams, [ref] + srcs, prm(7), planes)
    for (x, y) in [(10, 10), (20, 25), (35, 15)]:
        cv = costs[y, x]
        valid = cv[cv < 2.0]
        k = min(len(valid), 4)
        srt = np.sort(cv)
        assert abs(ic[y, x] - np.float32(srt[:k].sum(dtype=np.float32)) / np.float32(k)) <= 1e-6
        thr = srt[k - 1]
        expect = sum(1 << i for i in range(6) if cv[i] <= thr)
        assert mask[y, x] == expect
