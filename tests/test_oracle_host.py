"""Pins the oracle's restatement of the passes around PatchMatch on the CPU:
acosf, joint bilateral upsampling (JBU_cu / JBU_cu_prior, HPM.cu:3037-3197) and
the multi-view consistency check of fusion / confidence evaluation
(main.cpp:823-1109).

The pins are closed forms and invariants:
  * acosf within 1 ulp of long-double arccos;
  * JBU of a constant map is that constant; on a constant guide image JBU is
    plain Gaussian resampling with exp(-d^2 / (2 * 0.5^2)) weights; on a 0/255
    checkerboard guide only the offsets of equal guide value contribute;
  * with two translated cameras, a fronto-parallel plane and an integer
    disparity, fusion keeps exactly the reference view's pixels that see the
    other view (the second view's matches are masked; int(u + 0.5) truncation
    and the greedy masks decide the first columns), a 2 % depth error or a
    15 degree normal error rejects everything, and confidence maps mirror each
    other along the disparity.
"""
import math

import numpy as np

import oracle_py as O


def test_acosf_ulp_bound_and_specials():
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(0.999, 1, 2000), rng.uniform(-1, -0.999, 2000),
                         rng.uniform(-0.5, 0.5, 2000)]).astype(np.float32)
    got = np.array([O.acosf(float(x)) for x in xs], np.float64)
    ref = np.arccos(xs.astype(np.longdouble))
    ulp = np.spacing(np.abs(ref.astype(np.float32))).astype(np.float64)
    err = np.abs(got - ref.astype(np.float64)) / ulp
    assert err.max() <= 1.0, err.max()
    assert O.acosf(1.0) == 0.0
    assert O.acosf(-1.0) == np.float32(math.pi)
    assert O.acosf(0.0) == np.float32(math.pi / 2)
    assert math.isnan(O.acosf(1.0000001)) and math.isnan(O.acosf(float("nan")))


def _guide(h, w, seed):
    rng = np.random.default_rng(seed)
    return np.rint(rng.uniform(0, 255, (h, w))).astype(np.float32)


def test_jbu_constant_map_is_preserved():
    fine = _guide(40, 56, 1)
    coarse = np.full((20, 28), 3.25, np.float32)
    out = O.jbu_depth(fine, coarse)
    np.testing.assert_allclose(out, 3.25, rtol=2e-6)
    n = np.zeros((20, 28, 3), np.float32)
    n[..., 2] = -1.0
    d, on = O.jbu_prior(fine, coarse, n)
    np.testing.assert_allclose(d, 3.25, rtol=2e-6)
    np.testing.assert_allclose(on[..., :3], np.broadcast_to([0, 0, -1], on[..., :3].shape), atol=1e-7)
    assert np.all(on[..., 3] == 0)


def test_jbu_on_flat_guide_is_gaussian_resampling():
    """Constant guide: the range weight is 1, so JBU reduces to a Gaussian
    average of the clamped coarse window (scale 2: 5 x 5 window)."""
    h, w, s = 24, 32, 2
    fine = np.full((h, w), 100.0, np.float32)
    rng = np.random.default_rng(3)
    coarse = rng.uniform(1, 5, (h // s, w // s)).astype(np.float32)
    out = O.jbu_depth(fine, coarse)
    ch, cw = coarse.shape
    for (py, px) in [(0, 0), (5, 7), (11, 30), (23, 31), (12, 16)]:
        oy, ox = py * 0.5, px * 0.5
        num = den = 0.0
        for j in range(-2, 3):
            ry = min(max(int(oy + j), 0), ch - 1)
            for i in range(-2, 3):
                rx = min(max(int(ox + i), 0), cw - 1)
                g = math.exp(-((ox - rx) ** 2 + (oy - ry) ** 2) / 0.5)
                num += coarse[ry, rx] * g
                den += g
        assert abs(out[py, px] - num / den) <= 2e-6 * abs(num / den), (py, px)


def test_jbu_range_weight_on_checkerboard_guide():
    """A 0/255 checkerboard guide: the range weight of the fine neighbour
    (px + i, py + j) is 1 when i + j is even and exp(-255^2 / 1300.5) (about
    2e-22, below float resolution of the sums) when odd, so JBU is the Gaussian
    average over the even offsets only.  This also fixes that the guide
    neighbour is taken at the fine offset (i, j), not at the coarse sample."""
    h, w = 24, 32
    yy, xx = np.mgrid[0:h, 0:w]
    fine = np.where((xx + yy) % 2 == 0, 0.0, 255.0).astype(np.float32)
    rng = np.random.default_rng(4)
    coarse = rng.uniform(1, 5, (h // 2, w // 2)).astype(np.float32)
    out = O.jbu_depth(fine, coarse)
    ch, cw = coarse.shape
    for (py, px) in [(3, 3), (5, 8), (10, 21), (17, 12), (20, 26)]:
        oy, ox = py * 0.5, px * 0.5
        num = den = 0.0
        for j in range(-2, 3):
            ry = min(max(int(oy + j), 0), ch - 1)
            for i in range(-2, 3):
                if (i + j) % 2:
                    continue
                rx = min(max(int(ox + i), 0), cw - 1)
                g = math.exp(-((ox - rx) ** 2 + (oy - ry) ** 2) / 0.5)
                num += coarse[ry, rx] * g
                den += g
        assert abs(out[py, px] - num / den) <= 2e-6 * abs(num / den), (py, px)


def _stereo_pair(W=40, H=24, f=100.0, Z=5.0, disparity=8):
    """Two cameras translated along x by b = disparity * Z / f, looking down +z
    at a fronto-parallel plane z = Z (world = first camera frame)."""
    b = disparity * Z / f
    K = np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]], np.float32)
    R = np.eye(3, dtype=np.float32)
    cams = [O.camera(K, R, [0.0, 0.0, 0.0], W, H, 1.0, 10.0), O.camera(K, R, [-b, 0.0, 0.0], W, H, 1.0, 10.0)]
    depth = np.full((H, W), Z, np.float32)
    normal = np.zeros((H, W, 3), np.float32)
    normal[..., 2] = -1.0
    return cams, [depth, depth.copy()], [normal, normal.copy()]


def _expected_fusion_columns(W, dsp):
    """View-0 columns fused, in order, for the integer-disparity pair: column x
    sees view-1 column x - dsp.  int(u + 0.5) truncates toward zero, so column
    dsp - 1 (u = -1) lands on view-1 column 0 with a 1 px reprojection error
    (score e^-1 = 0.37 > 0.3) and is fused first; column dsp then finds column 0
    masked and is rejected; columns dsp + 1 .. W - 1 fuse.  Every view-1 pixel
    that could match was masked by view 0, the rest project outside view 0."""
    return [dsp - 1] + list(range(dsp + 1, W))


def test_fusion_two_view_closed_form():
    W, H, dsp = 40, 24, 8
    cams, depths, normals = _stereo_pair(W, H, disparity=dsp)
    pairs, pts = O.consistency(0, cams, depths, normals, [[1], [0]])
    expect = [(0, y * W + x) for y in range(H) for x in _expected_fusion_columns(W, dsp)]
    assert [tuple(map(int, p)) for p in pairs] == expect
    # points lie on the plane z = 5, at the back-projected pixel
    np.testing.assert_allclose(pts[:, 2], 5.0, rtol=1e-6)
    x = pairs[:, 1] % W
    np.testing.assert_allclose(pts[:, 0], (x - W / 2) * 5.0 / 100.0, atol=1e-5)
    np.testing.assert_allclose(pts[:, 3:], np.broadcast_to([0, 0, -1], pts[:, 3:].shape))


def test_fusion_rejects_depth_and_normal_errors():
    cams, depths, normals = _stereo_pair()
    d1 = [depths[0], depths[1] * np.float32(1.02)]
    pairs, _ = O.consistency(0, cams, d1, normals, [[1], [0]])
    assert len(pairs) == 0
    a = math.radians(15)
    n1 = normals[1].copy()
    n1[..., 0], n1[..., 2] = math.sin(a), -math.cos(a)
    pairs, _ = O.consistency(0, cams, depths, [normals[0], n1], [[1], [0]])
    assert len(pairs) == 0
    # an invalid (zero) depth never fuses
    pairs, _ = O.consistency(0, cams, [np.zeros_like(depths[0]), depths[1]], normals, [[1], [0]])
    assert all(p[0] == 1 for p in pairs)


def test_confidence_two_view_mirrors_along_disparity():
    W, H, dsp = 40, 24, 8
    cams, depths, normals = _stereo_pair(W, H, disparity=dsp)
    c0, c1 = O.consistency(1, cams, depths, normals, [[1], [0]])
    assert np.all(c0[:, :dsp - 1] == 0) and np.all(c0[:, dsp] == 0)
    np.testing.assert_allclose(c0[:, dsp - 1], math.exp(-1.0), rtol=1e-6)
    assert np.all(c0[:, dsp + 1:] > 0.99)
    # every fused view-0 pixel writes its score into the view-1 pixel it used
    np.testing.assert_array_equal(c1[:, 0], c0[:, dsp - 1])
    np.testing.assert_array_equal(c1[:, 1:W - dsp], c0[:, dsp + 1:])
    assert np.all(c1[:, W - dsp:] == 0)
    # depth range filter (main.cpp:1029,1057 use the reference camera's range
    # for both depths): view 0 with range [6, 10] fuses nothing as reference,
    # view 1 then fuses.  It writes its scores into the view-0 pixels it used,
    # but only in memory (main.cpp:1086): view 0's confidence.dmb was written
    # when view 0 was done (main.cpp:1095-1100), so that file stays zero
    K = np.array(cams[0].K).reshape(3, 3)
    far0 = O.camera(K, np.eye(3), [0, 0, 0], W, H, 6.0, 10.0)
    c0, c1 = O.consistency(1, [far0, cams[1]], depths, normals, [[1], [0]])
    assert np.all(c1[:, W - dsp:] == 0) and np.all(c1[:, :W - dsp] > 0.99)
    assert np.all(c0 == 0)
    # with the two images swapped in pair order, the scores reach the
    # [6, 10] view's map before that view is done, so they are in its file
    c1s, c0s = O.consistency(1, [cams[1], far0], [depths[1], depths[0]], [normals[1], normals[0]], [[1], [0]])
    np.testing.assert_array_equal(c1s, c1)
    np.testing.assert_array_equal(c0s[:, dsp:], c1s[:, :W - dsp])
    assert np.all(c0s[:, :dsp] == 0)
    far1 = O.camera(K, np.eye(3), [-dsp * 5.0 / 100.0, 0, 0], W, H, 6.0, 10.0)
    c0, c1 = O.consistency(1, [far0, far1], depths, normals, [[1], [0]])
    assert np.all(c0 == 0) and np.all(c1 == 0)


_REPROJ_SRC = r"""
#include <cmath>
// main.cpp:922 / 1064, compiled as main.cpp is (host C++11): c and r are int
// pixel coordinates, tmp_pt.x / tmp_pt.y float.
extern "C" void ref_reproj(const int* c, const int* r, const float* x, const float* y, int n, float* out) {
  for (int i = 0; i < n; ++i) {
    float reproj_error = sqrt(pow(c[i] - x[i], 2) + pow(r[i] - y[i], 2));
    out[i] = reproj_error;
  }
}
"""


def test_reproj_error_is_the_cxx11_double_expression(tmp_path):
    """The reprojection error of the consistency check equals the reference's
    expression compiled with g++ -std=c++11: pow(float, int) returns double
    there, so the squares, sum and square root are double.  A float
    evaluation (sqrtf(dx*dx + dy*dy)) differs on about 17 % of these inputs."""
    import ctypes
    import subprocess
    src = tmp_path / "reproj.cpp"
    src.write_text(_REPROJ_SRC)
    so = tmp_path / "libreproj.so"
    subprocess.run(["g++", "-std=c++11", "-O2", "-ffp-contract=off", "-fPIC", "-shared", str(src), "-o", str(so)],
                   check=True)
    L = ctypes.CDLL(str(so))
    rng = np.random.default_rng(7)
    n = 200000
    c = rng.integers(0, 1600, n).astype(np.int32)
    r = rng.integers(0, 1200, n).astype(np.int32)
    u = (c + rng.uniform(-2.5, 2.5, n)).astype(np.float32)
    v = (r + rng.uniform(-2.5, 2.5, n)).astype(np.float32)
    ref = np.empty(n, np.float32)
    L.ref_reproj(*(ctypes.c_void_p(a.ctypes.data) for a in (c, r, u, v)), ctypes.c_int(n),
                 ctypes.c_void_p(ref.ctypes.data))
    got = O.reproj_error(c, r, u, v)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    dx, dy = (c.astype(np.float32) - u), (r.astype(np.float32) - v)
    single = np.sqrt(dx * dx + dy * dy, dtype=np.float32)
    assert np.mean(single != ref) > 0.05  # the case distinguishes the two evaluations
