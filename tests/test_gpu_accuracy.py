"""The headline view at full size -- RunPatchMatch (photometric) of a 1600 x
1200 reference with 10 sources, the bench's workload -- against the FP32 and
FP64 oracles on the whole view, same seed.

* The exact FP32 engine equals the FP32 oracle bit for bit: planes (world
  normal, depth), costs.
* Accuracy against FP64: the share of pixels whose depth lies within 1 % of
  the FP64 result and whose normal lies within 5 degrees of it.  PatchMatch
  turns a small difference in one cost into a different adoption, and the
  adopted hypothesis then spreads.  The exact FP32 path performs the
  reference's own FP32 arithmetic, whose E[x^2] - E[x]^2 window variance
  cancels (costs off by up to 5e-3); it agrees with FP64 on 0.958 of the
  depths and only 0.658 of the normals (MI355X, seed 109): a normal is
  decided between refinement candidates whose costs differ by less than that
  noise.  The mixed-precision paths (FP64 view-pair basis, FP32 tap offsets,
  shifted window sums; FP16, BF16 or FP32 texels) reach 0.998 of the depths
  and 0.984 of the normals.
* The floor: the FP64 oracle against itself, with every source camera's
  translation one float ulp up, agrees on 0.995 of the depths and 0.962 of
  the normals (seed 109).  A change of the inputs at the last bit moves the
  FP64 result further than the mixed-precision arithmetic does, so 99 % of
  normals within 5 degrees is not a bound any arithmetic can meet on this
  scene; the test asserts the floor instead.
Tolerances: depth >= 0.99 (depth within 1 % on 99 % of pixels); normals >=
0.975 (0.984 measured) and at least the FP64 floor; never worse than the
exact path.
"""
import time

import numpy as np
import pytest

import oracle_py as O
from conftest import gpu_available, nudged_cameras

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

SEED = 109


def agreement(planes, ref):
    d, d0 = planes[..., 3].astype(np.float64), ref[..., 3].astype(np.float64)
    n, n0 = planes[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    cosang = np.sum(n * n0, -1) / np.maximum(np.linalg.norm(n, axis=-1) * np.linalg.norm(n0, axis=-1), 1e-30)
    ang = np.degrees(np.arccos(np.clip(cosang, -1, 1)))
    return float(np.mean(np.abs(d - d0) <= 0.01 * np.abs(d0))), float(np.mean(ang <= 5.0))


def test_headline_view_against_fp32_and_fp64_oracles():
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    print("case: RunPatchMatch 1600x1200 V=10 (bench view), all precisions vs FP32 / FP64 oracle", flush=True)
    f = 525.0 * 1600 / 640
    sc = scenes.make_plane_scene(seed=1, n_cams=11, width=1600, height=1200, f=f, n_src=10)
    ids, cams, imgs = scenes.reference_view_problem(sc, 5, 10)
    oc = [O.camera_from_scene(c) for c in cams]
    prm = O.params(len(oc), oc[0])
    t0 = time.perf_counter()
    r32 = O.run_patchmatch(oc, imgs, prm, SEED)
    t1 = time.perf_counter()
    r64 = O.run_patchmatch(oc, imgs, prm, SEED, dtype=np.float64)
    t2 = time.perf_counter()
    print(f"oracle FP32 {t1 - t0:.1f} s, FP64 {t2 - t1:.1f} s", flush=True)
    r64u = O.run_patchmatch(nudged_cameras(oc), imgs, prm, SEED, dtype=np.float64)
    floor64 = agreement(r64u["planes"], r64["planes"])
    print(f"FP64 vs FP64 with source translations one ulp up: depth {floor64[0]:.4f}, normals {floor64[1]:.4f}",
          flush=True)
    ctx = H.Context(0)
    ctx.set_views([H.camera_from_scene(c) for c in cams], imgs)
    res = {}
    for prec in ("fp32", "fp16", "bf16", "fp32fast"):
        ctx.set_params(H.default_params(precision=prec))
        ctx.run(SEED)
        planes, costs = ctx.planes(), ctx.costs()
        if prec == "fp32":
            assert np.array_equal(planes.view(np.uint32), r32["planes"].view(np.uint32))
            assert np.array_equal(costs.view(np.uint32), r32["costs"].view(np.uint32))
        res[prec] = agreement(planes, r64["planes"])
        print(f"{prec}: depth within 1 % of FP64 {res[prec][0]:.4f}, normals within 5 deg {res[prec][1]:.4f}",
              flush=True)
    ctx.close()
    gt = sc.depths[ids[0]]
    print(f"within 1 % of ground truth: FP64 {np.mean(np.abs(r64['planes'][..., 3] - gt) / gt <= 0.01):.4f}, "
          f"FP32 {np.mean(np.abs(r32['planes'][..., 3] - gt) / gt <= 0.01):.4f}", flush=True)
    floor_d, floor_n = res["fp32"]
    assert floor_d >= 0.94 and floor_n >= 0.6
    for prec in ("fp16", "bf16", "fp32fast"):
        assert res[prec][0] >= 0.99 and res[prec][1] >= 0.975, prec
        assert res[prec][0] >= floor_d and res[prec][1] >= floor_n, prec
        assert res[prec][0] >= min(0.99, floor64[0]) and res[prec][1] >= min(0.99, floor64[1]), prec
