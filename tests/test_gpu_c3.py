"""The C3 scene at full size -- 3072 x 2048, 30 cameras on two rings, 24
plane patches and 6 spheres, V = 10 -- on three designated reference views
(7 on the low ring, 15 and 22 on the high one), against the FP64 oracle and
against ground truth.  The oracle is far too slow for whole 3072 x 2048
views, so it runs on a band of 256 rows:

* cost at fixed hypotheses (one random plane per pixel), FP16 storage
  against FP64 on the band (at fixed hypotheses a band holds exactly the
  full image's rows): |dC| <= 1e-5 at q999, mean <= 2e-6, the same
  valid / invalid decisions on more than 99.9 % of pairs;
* a photometric RunPatchMatch in FP16 over the whole view against the
  ray-cast ground truth, and against the FP64 oracle's RunPatchMatch of the
  band on the band's central 128 rows.  A band run lacks what propagates in
  from the rows outside it, and PatchMatch turns one different adoption into
  a different patch, so that agreement is bounded loosely (depth within 1 %
  of FP64 on more than 93 % of those pixels; 95-98 % measured on MI355X).
  The parity bound is the ground-truth accuracy of the two on the same rows:
  the GPU's within 0.02 of FP64's or better.  View 15, at the end of the high
  ring with every source to one side, is the hardest of the three: about
  0.76 of its pixels lie within 1 % of ground truth after one photometric
  pass (0.90 for view 7), for the FP64 oracle as for the GPU.
"""
import time

import numpy as np
import pytest

import oracle_py as O
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

REFS = (7, 15, 22)
SEED = 5


@pytest.fixture(scope="module")
def c3_scene():
    from hpm_mi355x import scenes
    t0 = time.perf_counter()
    sc = scenes.make_orbit_scene(seed=3, render_refs=REFS)
    print(f"C3 views rendered in {time.perf_counter() - t0:.1f} s", flush=True)
    return sc


@pytest.mark.parametrize("ref", REFS)
def test_c3_cost_vs_fp64_band(c3_scene, ref):
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    ids, cams, imgs = scenes.reference_view_problem(c3_scene, ref, 10)
    assert (cams[0].width, cams[0].height, len(cams)) == (3072, 2048, 11)
    print(f"case: C3 cost_fixed fp16 3072x2048 V=10, view {ref}, FP64 band", flush=True)
    planes = scenes.random_planes_per_pixel(cams[0], seed=4)
    y0, y1 = 896, 1152
    oc = [O.camera_from_scene(c) for c in cams]
    ref64 = O.cost_fixed(oc, imgs, O.params(len(oc), oc[0]), planes.astype(np.float64), np.float64,
                         rows=(y0, y1))[y0:y1]
    ctx = H.Context(0)
    ctx.set_params(H.default_params(precision="fp16"))
    ctx.set_views([H.camera_from_scene(c) for c in cams], imgs)
    got = ctx.cost_fixed(planes, precision=H.PRECISIONS["fp16"]).astype(np.float64)[y0:y1]
    same = (got == 2.0) == (ref64 == 2.0)
    err = np.abs(got - ref64)[same]
    print(f"view {ref}: validity agrees {same.mean():.6f}, |dC| max {err.max():.2e} "
          f"q999 {np.quantile(err, 0.999):.2e} mean {err.mean():.2e}")
    assert same.mean() > 0.999
    assert np.quantile(err, 0.999) <= 1e-5 and err.mean() <= 2e-6


@pytest.mark.parametrize("ref", REFS)
def test_c3_patchmatch_vs_gt_and_fp64_band(c3_scene, ref):
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    ids, cams, imgs = scenes.reference_view_problem(c3_scene, ref, 10)
    print(f"case: C3 RunPatchMatch fp16 3072x2048 V=10, view {ref}", flush=True)
    ctx = H.Context(0)
    ctx.set_params(H.default_params(precision="fp16"))
    ctx.set_views([H.camera_from_scene(c) for c in cams], imgs)
    ctx.run(SEED)  # first run: pair images, warm-up
    t0 = time.perf_counter()
    ctx.run(SEED)
    dt = time.perf_counter() - t0
    planes = ctx.planes()
    d = planes[..., 3].astype(np.float64)
    gt = c3_scene.depths[ref].astype(np.float64)
    hit = gt > 0
    gt_frac = float(np.mean(np.abs(d[hit] - gt[hit]) / gt[hit] <= 0.01))
    # the FP64 oracle on rows 896-1151, compared on the central 128 rows
    y0, y1, m = 896, 1152, 64
    oc = [O.camera_from_scene(c) for c in cams]
    t1 = time.perf_counter()
    r64 = O.run_patchmatch(oc, imgs, O.params(len(oc), oc[0]), SEED, dtype=np.float64, rows=(y0, y1))
    t_oracle = time.perf_counter() - t1
    band = slice(y0 + m, y1 - m)
    d64 = r64["planes"][band, :, 3]
    rel = np.abs(d[band] - d64) / np.abs(d64)
    n, n64 = planes[band, :, :3].astype(np.float64), r64["planes"][band, :, :3]
    cosang = np.sum(n * n64, -1) / (np.linalg.norm(n, axis=-1) * np.linalg.norm(n64, axis=-1))
    ang = np.degrees(np.arccos(np.clip(cosang, -1, 1)))
    gt_band = gt[band]
    hb = gt_band > 0
    gt64 = float(np.mean(np.abs(d64[hb] - gt_band[hb]) / gt_band[hb] <= 0.01))
    print(f"view {ref}: GPU RunPatchMatch {1e3 * dt:.1f} ms (oracle FP64 band {t_oracle:.1f} s); "
          f"within 1 % of GT {gt_frac:.3f} (whole view), band rows: GPU {np.mean(np.abs(d[band][hb] - gt_band[hb]) / gt_band[hb] <= 0.01):.3f}, "
          f"FP64 {gt64:.3f}; GPU vs FP64: depth within 1 % {np.mean(rel <= 0.01):.3f}, normal within 5 deg "
          f"{np.mean(ang <= 5.0):.3f}")
    gpu_band = float(np.mean(np.abs(d[band][hb] - gt_band[hb]) / gt_band[hb] <= 0.01))
    assert gt_frac > 0.70
    assert gpu_band >= gt64 - 0.02  # as close to the ground truth as FP64 on the same rows
    assert np.mean(rel <= 0.01) > 0.93
