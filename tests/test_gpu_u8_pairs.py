"""8-bit column-pair storage of the FP16 precision (needs an MI355X).

With every texel an integer in [0, 255] the 8-bit pairs feed the bilinear
sample the same exact operands as the FP16 column pairs, so the storage must
not change a bit of any result.  Each case runs the same problem with the
storage forced on (hpm_set_u8_pairs_min_bytes 0) and off (-1) and compares
planes, costs, view masks and texture maps bit for bit:
  * single-view RunPatchMatch, photometric, with and without hipGraph replay;
  * the cost at fixed random planes (every (pixel, view) pair);
  * contexts on a shared image pool, run as a fused batch (hpm_batch);
  * a fused batch whose views do not all qualify falls back to FP16 pairs
    for all of them;
  * images that are not integral (a quarter level added to random texels)
    stay on FP16 pairs.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


def _problem(w=160, h=120, n=6, seed=4, ref=2):
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=seed, n_cams=n, width=w, height=h, f=0.82 * w, n_src=n - 1)
    return sc, scenes.reference_view_problem(sc, ref, n - 1)


def _ctx(cams, imgs, u8, graph=False):
    import hpm_mi355x as H
    c = H.Context(0)
    c.set_params(H.default_params(precision="fp16"))
    c.set_u8_pairs_min_bytes(0 if u8 else -1)
    c.set_views([H.camera_from_scene(x) for x in cams], imgs)
    c.set_graph(graph)
    return c


def _same(a, b):
    np.testing.assert_array_equal(a.planes().view(np.uint32), b.planes().view(np.uint32))
    np.testing.assert_array_equal(a.costs().view(np.uint32), b.costs().view(np.uint32))
    np.testing.assert_array_equal(a.selected_views(), b.selected_views())
    np.testing.assert_array_equal(a.texture(), b.texture())


@pytest.mark.parametrize("graph", [False, True])
def test_u8_run_equals_fp16(graph):
    print(f"case: RunPatchMatch fp16 160x120 V=5, 8-bit vs FP16 pairs, graph={graph}", flush=True)
    sc, (ids, cams, imgs) = _problem()
    a, b = _ctx(cams, imgs, True, graph), _ctx(cams, imgs, False, graph)
    for seed in (7, 8):  # with graphs: capture, then replay
        a.run(seed)
        b.run(seed)
        _same(a, b)
    a.close(); b.close()


def test_u8_cost_fixed_equals_fp16():
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    print("case: cost_fixed fp16 160x120 V=5, 8-bit vs FP16 pairs", flush=True)
    sc, (ids, cams, imgs) = _problem()
    planes = scenes.random_planes_per_pixel(cams[0], seed=4)
    a, b = _ctx(cams, imgs, True), _ctx(cams, imgs, False)
    ca = a.cost_fixed(planes, precision=H.PRECISIONS["fp16"])
    cb = b.cost_fixed(planes, precision=H.PRECISIONS["fp16"])
    np.testing.assert_array_equal(ca.view(np.uint32), cb.view(np.uint32))
    assert (ca < 2.0).mean() > 0.5
    a.close(); b.close()


def test_u8_pooled_fused_batch_equals_fp16():
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    print("case: pooled fused batch fp16 128x96, 3 views, 8-bit vs FP16 pairs", flush=True)
    sc = scenes.make_plane_scene(seed=3, n_cams=7, width=128, height=96, f=105.0, n_src=4)
    pool = H.ImagePool([H.camera_from_scene(c) for c in sc.cameras], [im.astype(np.float32) for im in sc.images])
    runs = {}
    for u8 in (True, False):
        ctxs = []
        for ref in (2, 3, 4):
            ids, cams, imgs = scenes.reference_view_problem(sc, ref, 4)
            c = H.Context(0)
            c.set_params(H.default_params(precision="fp16"))
            c.set_u8_pairs_min_bytes(0 if u8 else -1)
            c.set_views_pooled(pool, ids)
            ctxs.append(c)
        b = H.Batch(ctxs)
        b.set_graph(True)
        for _ in range(2):
            b.run([51, 52, 53])
        runs[u8] = ctxs
    for x, y in zip(runs[True], runs[False]):
        _same(x, y)


def test_u8_mixed_batch_falls_back_to_fp16():
    """One view forced onto 8-bit pairs, one never: the batch runs both on
    FP16 pairs, with the bits of single FP16 runs."""
    import hpm_mi355x as H
    print("case: fused batch fp16, one view 8-bit-eligible and one not", flush=True)
    _, (i1, c1, m1) = _problem(160, 120, 6, 4, 2)
    _, (i2, c2, m2) = _problem(128, 96, 5, 9, 1)
    fused = [_ctx(c1, m1, True), _ctx(c2, m2, False)]
    single = [_ctx(c1, m1, False), _ctx(c2, m2, False)]
    b = H.Batch(fused)
    b.run([61, 62])
    single[0].run(61)
    single[1].run(62)
    for f, s in zip(fused, single):
        _same(f, s)


def test_non_integral_images_stay_on_fp16():
    print("case: RunPatchMatch fp16 160x120 with non-integral texels: no 8-bit pairs", flush=True)
    sc, (ids, cams, imgs) = _problem()
    rng = np.random.default_rng(0)
    frac = [(im + 0.25 * rng.integers(0, 2, im.shape)).astype(np.float32) for im in imgs]
    a, b = _ctx(cams, frac, True), _ctx(cams, frac, False)
    a.run(7)
    b.run(7)
    _same(a, b)
    c = _ctx(cams, imgs, False)
    c.run(7)
    assert not np.array_equal(c.costs(), b.costs())  # the quarter level does reach the costs
