"""Edge cases of the RunPatchMatch path (needs an MI355X), bit-exact against
the FP32 oracle and close to FP64 on the mixed-precision engine:
  * one source view (V = 1);
  * V = 20, the largest view count whose direction cost rows live in LDS
    (8 x 20 rows x 256 threads x 4 B = 160 KiB per workgroup);
  * V = 24, past that limit (rows in private memory);
  * odd image sizes (an odd last row is never visited, HPM.cu:3006);
  * the smallest accepted reference image (16 x 16).
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
