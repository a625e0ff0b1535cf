"""GPU parity of the passes around PatchMatch (needs an MI355X): joint
bilateral upsampling (hpm_jbu_depth / hpm_jbu_prior vs JBU_cu / JBU_cu_prior,
HPM.cu:3037-3197) and the fusion / confidence consistency check
(hpm_consistency_check vs RunFusion / ConfidenceEvaluation, main.cpp:823-1109).

Bar: bit-exact against the FP32 oracle (same float operations, same order, the
same exp / acos); fused-point indices identical and in the same order.
"""
import numpy as np
import pytest

import oracle_py as O
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


@pytest.mark.parametrize("h,w,s", [(48, 64, 2), (61, 83, 2), (64, 96, 4), (30, 40, 3)])
def test_jbu_depth_bitexact(h, w, s):
    import hpm_mi355x as H
    print(f"case: jbu_depth {w}x{h} scale {s}", flush=True)
    rng = np.random.default_rng(h * w + s)
    fine = np.rint(rng.uniform(0, 255, (h, w))).astype(np.float32)
    coarse = rng.uniform(1, 5, (h // s, w // s)).astype(np.float32)
    coarse[0, 0] = np.nan  # NaN propagates the same way on both sides
    got = H.jbu_depth(fine, coarse)
    ref = O.jbu_depth(fine, coarse)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("h,w,s", [(48, 64, 2), (61, 83, 2), (64, 96, 4)])
def test_jbu_prior_bitexact(h, w, s):
    import hpm_mi355x as H
    print(f"case: jbu_prior {w}x{h} scale {s}", flush=True)
    rng = np.random.default_rng(7 * h + w)
    fine = np.rint(rng.uniform(0, 255, (h, w))).astype(np.float32)
    coarse = rng.uniform(1, 5, (h // s, w // s)).astype(np.float32)
    n = rng.normal(size=(h // s, w // s, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    gd, gn = H.jbu_prior(fine, coarse, n)
    rd, rn = O.jbu_prior(fine, coarse, n)
    np.testing.assert_array_equal(gd, rd)
    np.testing.assert_array_equal(gn, rn)


def _noisy_maps(seed=11, n_cams=5, width=96, height=72):
    """Ground-truth depth / normal maps of 5 views with 0.3 % depth noise, 5 %
    outliers and a few invalid pixels, so that every branch of the test fires."""
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=2, n_cams=n_cams, width=width, height=height, f=78.75 * width / 96,
                                 n_src=n_cams - 1)
    rng = np.random.default_rng(seed)
    depths, normals = [], []
    for d, n in zip(sc.depths, sc.normals):
        d = d * (1 + rng.normal(0, 0.003, d.shape)).astype(np.float32)
        out = rng.random(d.shape) < 0.05
        d[out] *= rng.uniform(0.7, 1.3, out.sum()).astype(np.float32)
        d[rng.random(d.shape) < 0.01] = 0.0
        depths.append(d.astype(np.float32))
        nn = n + rng.normal(0, 0.05, n.shape).astype(np.float32)
        nn /= np.linalg.norm(nn, axis=-1, keepdims=True)
        normals.append(nn.astype(np.float32))
    sources = [[j for j, s in sc.pairs[i]] for i in range(len(sc.cameras))]
    return sc, depths, normals, sources


def test_fusion_bitexact():
    import hpm_mi355x as H
    print("case: consistency fusion 5 x 96x72", flush=True)
    sc, depths, normals, sources = _noisy_maps()
    hc = [H.camera_from_scene(c) for c in sc.cameras]
    oc = [O.camera_from_scene(c) for c in sc.cameras]
    gp, gx = H.consistency_check(H.CHECK_FUSION, hc, depths, normals, sources)
    rp, rx = O.consistency(0, oc, depths, normals, sources)
    assert len(rp) > 1000
    np.testing.assert_array_equal(gp, rp)
    np.testing.assert_array_equal(gx, rx)


def test_fusion_mask_mode_bitexact():
    """RunFusion_Sky_Strict's consistency pass (float acceptance threshold)."""
    import hpm_mi355x as H
    print("case: consistency fusion-mask 5 x 96x72", flush=True)
    sc, depths, normals, sources = _noisy_maps(17)
    hc = [H.camera_from_scene(c) for c in sc.cameras]
    oc = [O.camera_from_scene(c) for c in sc.cameras]
    gp, gx = H.consistency_check(H.CHECK_FUSION_MASK, hc, depths, normals, sources)
    rp, rx = O.consistency(2, oc, depths, normals, sources)
    assert len(rp) > 1000
    np.testing.assert_array_equal(gp, rp)
    np.testing.assert_array_equal(gx, rx)


def test_confidence_bitexact():
    import hpm_mi355x as H
    print("case: consistency confidence 5 x 96x72", flush=True)
    sc, depths, normals, sources = _noisy_maps(12)
    hc = [H.camera_from_scene(c) for c in sc.cameras]
    oc = [O.camera_from_scene(c) for c in sc.cameras]
    g = H.consistency_check(H.CHECK_CONFIDENCE, hc, depths, normals, sources)
    r = O.consistency(1, oc, depths, normals, sources)
    assert sum(int((x > 0).sum()) for x in r) > 1000
    for a, b in zip(g, r):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("mode", [0, 1])
def test_consistency_chunked_bitexact(mode):
    """Views of 1280 x 960 (1.23 Mpix) are resolved in two chunks of rows each
    (8 chunks, more than the pinned buffers that rotate among them), and a
    view without sources sits between them: fusion and confidence still
    equal the oracle's bit for bit."""
    import hpm_mi355x as H
    print(f"case: consistency mode {mode} 5 x 1280x960, chunked", flush=True)
    sc, depths, normals, sources = _noisy_maps(13 + mode, n_cams=5, width=1280, height=960)
    sources[1] = []
    hc = [H.camera_from_scene(c) for c in sc.cameras]
    oc = [O.camera_from_scene(c) for c in sc.cameras]
    g = H.consistency_check(H.CHECK_FUSION if mode == 0 else H.CHECK_CONFIDENCE, hc, depths, normals, sources)
    r = O.consistency(mode, oc, depths, normals, sources)
    if mode == 0:
        assert len(r[0]) > 100000
        np.testing.assert_array_equal(g[0], r[0])
        np.testing.assert_array_equal(g[1], r[1])
    else:
        assert sum(int((x > 0).sum()) for x in r) > 100000
        for a, b in zip(g, r):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_consistency_self_and_repeated_sources_bitexact(mode):
    """Source lists that name the image itself or one source twice (two
    slots then write the same mask and confidence locations, and an image
    masks its own pixels): fusion and confidence equal the oracle."""
    import hpm_mi355x as H
    print(f"case: consistency mode {mode} 5 x 96x72, self and repeated sources", flush=True)
    sc, depths, normals, sources = _noisy_maps(21 + mode)
    sources[0] = [0] + sources[0]
    sources[2] = sources[2][:2] + sources[2][:2] + sources[2][2:]
    hc = [H.camera_from_scene(c) for c in sc.cameras]
    oc = [O.camera_from_scene(c) for c in sc.cameras]
    kind = {0: H.CHECK_FUSION, 1: H.CHECK_CONFIDENCE, 2: H.CHECK_FUSION_MASK}[mode]
    g = H.consistency_check(kind, hc, depths, normals, sources)
    r = O.consistency(mode, oc, depths, normals, sources)
    if mode == 1:
        assert sum(int((x > 0).sum()) for x in r) > 1000
        for a, b in zip(g, r):
            np.testing.assert_array_equal(a, b)
    else:
        assert len(r[0]) > 1000
        np.testing.assert_array_equal(g[0], r[0])
        np.testing.assert_array_equal(g[1], r[1])


def test_fusion_two_view_closed_form_on_gpu():
    """The integer-disparity pair of the oracle pins, through the C ABI."""
    import hpm_mi355x as H
    W, Hh, f, Z, dsp = 40, 24, 100.0, 5.0, 8
    K = np.array([[f, 0, W / 2], [0, f, Hh / 2], [0, 0, 1]], np.float32)
    cams = [H.make_camera(K, np.eye(3), [0, 0, 0], W, Hh, 1, 10),
            H.make_camera(K, np.eye(3), [-dsp * Z / f, 0, 0], W, Hh, 1, 10)]
    d = np.full((Hh, W), Z, np.float32)
    n = np.zeros((Hh, W, 3), np.float32)
    n[..., 2] = -1
    pairs, _ = H.consistency_check(H.CHECK_FUSION, cams, [d, d], [n, n], [[1], [0]])
    cols = [dsp - 1] + list(range(dsp + 1, W))
    assert [tuple(map(int, p)) for p in pairs] == [(0, y * W + x) for y in range(Hh) for x in cols]
