"""Edge cases of the RunPatchMatch path (needs an MI355X), bit-exact against
the FP32 oracle and close to FP64 on the mixed-precision engine:
  * one source view (V = 1);
  * V = 20, the largest view count whose direction cost rows live in LDS
    (8 x 20 rows x 256 threads x 4 B = 160 KiB per workgroup);
  * V = 24, past that limit (rows in private memory);
  * odd image sizes (an odd last row is never visited, HPM.cu:3006);
  * the smallest accepted reference image (16 x 16);
  * the largest view the mixed-precision path accepts, and the first it
    refuses.
"""
import numpy as np
import pytest

import oracle_py as O
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


def _problem(w, h, nsrc, seed=11):
    from hpm_mi355x import scenes
    f = 0.82 * w
    sc = scenes.make_plane_scene(seed=seed, n_cams=nsrc + 1, width=w, height=h, f=f, n_src=nsrc)
    return scenes.reference_view_problem(sc, nsrc // 2, nsrc)


@pytest.mark.parametrize("w,h,nsrc", [(48, 40, 1), (40, 32, 20), (40, 32, 24), (37, 29, 3), (16, 16, 2)])
def test_exact_path_bitexact(w, h, nsrc):
    import hpm_mi355x as H
    ids, cams, imgs = _problem(w, h, nsrc)
    print(f"case: RunPatchMatch exact {w}x{h} V={len(cams) - 1}", flush=True)
    ctx = H.Context(0)
    ctx.set_params(H.default_params())
    ctx.set_views([H.camera_from_scene(c) for c in cams], imgs)
    ctx.run(9)
    oc = [O.camera_from_scene(c) for c in cams]
    r = O.run_patchmatch(oc, imgs, O.params(len(oc), oc[0]), 9)
    assert np.array_equal(ctx.planes().view(np.uint32), r["planes"].view(np.uint32))
    assert np.array_equal(ctx.costs().view(np.uint32), r["costs"].view(np.uint32))
    assert np.array_equal(ctx.selected_views(), r["selected"])


@pytest.mark.parametrize("w,h,nsrc", [(40, 32, 20), (40, 32, 24), (37, 29, 3)])
def test_fp16_path_close_to_fp64(w, h, nsrc):
    import hpm_mi355x as H
    ids, cams, imgs = _problem(w, h, nsrc)
    print(f"case: RunPatchMatch fp16 {w}x{h} V={len(cams) - 1}", flush=True)
    ctx = H.Context(0)
    ctx.set_params(H.default_params(precision="fp16"))
    ctx.set_views([H.camera_from_scene(c) for c in cams], imgs)
    ctx.run(9)
    oc = [O.camera_from_scene(c) for c in cams]
    r64 = O.run_patchmatch(oc, imgs, O.params(len(oc), oc[0]), 9, dtype=np.float64)
    d, d64 = ctx.planes()[..., 3].astype(np.float64), r64["planes"][..., 3]
    assert np.mean(np.abs(d - d64) / np.abs(d64) <= 0.01) >= 0.99


def test_fp16_largest_view_cost_vs_fp64():
    """The largest view the mixed-precision path accepts: 4095 x 4094, whose
    (w + 2) x (h + 1) column-pair grid holds 2^24 - 1 elements, so the cell
    addresses, formed in float, reach 2^24 - 2 and stay exact.  Cost at fixed
    random planes, FP16 against the FP64 oracle on the last 48 rows (the
    largest addresses): validity agrees on > 99.9 % of the pairs, max |dC| <=
    1e-5.  One column more (4096 x 4094, 2^24 + 4094 elements) is refused with
    an error before any launch; the exact FP32 path takes it."""
    import dataclasses

    import hpm_mi355x as H
    from hpm_mi355x import scenes
    W, Hh = 4095, 4094
    print(f"case: cost_fixed fp16 {W}x{Hh} V=2 vs FP64 band (largest accepted view)", flush=True)
    sc = scenes.make_plane_scene(seed=21, n_cams=3, width=W, height=Hh, f=0.82 * W, n_src=2)
    ids, cams, imgs = scenes.reference_view_problem(sc, 1, 2)
    assert (W + 2) * (Hh + 1) == (1 << 24) - 1
    planes = scenes.random_planes_per_pixel(cams[0], seed=4)
    y0, y1 = Hh - 48, Hh
    oc = [O.camera_from_scene(c) for c in cams]
    ref64 = O.cost_fixed(oc, imgs, O.params(len(oc), oc[0]), planes.astype(np.float64), np.float64,
                         rows=(y0, y1))[y0:y1]
    ctx = H.Context(0)
    ctx.set_params(H.default_params(precision="fp16"))
    ctx.set_views([H.camera_from_scene(c) for c in cams], imgs)
    got = ctx.cost_fixed(planes, precision=H.PRECISIONS["fp16"]).astype(np.float64)[y0:y1]
    ctx.close()
    same = (got == 2.0) == (ref64 == 2.0)
    valid = same & (ref64 != 2.0)
    err = np.abs(got - ref64)[same]
    print(f"validity agrees {same.mean():.6f}, valid pairs {valid.sum()}, |dC| max {err.max():.2e} "
          f"mean {err.mean():.2e}")
    assert same.mean() > 0.999 and valid.sum() > 100000
    assert err.max() <= 1e-5

    print(f"case: fp16 refuses {W + 1}x{Hh}", flush=True)
    wide = [dataclasses.replace(c, width=W + 1) for c in cams]
    wide_imgs = [np.ascontiguousarray(np.pad(im, ((0, 0), (0, 1)), mode="edge")) for im in imgs]
    ctx = H.Context(0)
    ctx.set_params(H.default_params(precision="fp16"))
    ctx.set_views([H.camera_from_scene(c) for c in wide], wide_imgs)
    band = np.ascontiguousarray(np.pad(planes, ((0, 0), (0, 1), (0, 0)), mode="edge"))
    with pytest.raises(H.HpmError, match="mixed-precision"):
        ctx.cost_fixed(band, precision=H.PRECISIONS["fp16"])
    exact = ctx.cost_fixed(band, precision=H.PRECISIONS["fp32"])
    ctx.close()
    assert exact.shape == (Hh, W + 1, 2) and np.isfinite(exact).all()
