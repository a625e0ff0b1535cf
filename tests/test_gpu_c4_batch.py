"""The uneven-size batch (the C4 workload's shape): reference views of
640 x 480, 1024 x 768, 1600 x 1200 and 2048 x 1536 with 6 to 10 sources each,
over one scene, run as one fused batch (one launch per stage over every
view's tiles, hipGraph replay) with the plain FP32 engine, against the FP32
oracle's RunPatchMatch of each view, bit for bit:

* views up to 1024 x 768: the whole view (planes after the output step,
  costs, texture);
* the larger views: the rows [y0, y0 + 64) of a band.  The oracle computes
  rows [y0 - 180, y0 + 64 + 84) and initialises the rows its propagation
  reads around them; a pixel reads rows up to 29 above and 13 below its own
  (HPM.cu:1367-1607, the left-up extension included), so the band's edges
  reach at most 6 x 29 = 174 rows down and 6 x 13 = 78 rows up in the six
  half-passes, and the checked rows hold exactly the whole view's values.

12 views (scene seed 11): four of 640 x 480, three of 1024 x 768, three of
1600 x 1200, two of 2048 x 1536."""
import time

import numpy as np
import pytest

import oracle_py as O
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


def test_uneven_fused_batch_fp32_equals_oracle():
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    print("case: fused batch fp32, 12 uneven views 640x480..2048x1536, V 6..10 vs FP32 oracle", flush=True)
    t0 = time.perf_counter()
    sc = scenes.make_batch_scene(seed=11, n_views=12)
    print(f"scene rendered in {time.perf_counter() - t0:.1f} s", flush=True)
    sizes = {(c.width, c.height) for c in sc.cameras}
    assert sizes == {(640, 480), (1024, 768), (1600, 1200), (2048, 1536)}
    assert {len(p) for p in sc.pairs} >= {6, 10}
    pool = H.ImagePool([H.camera_from_scene(c) for c in sc.cameras], [im.astype(np.float32) for im in sc.images])
    ctxs, seeds = [], []
    for i in range(len(sc.cameras)):
        ids = [i] + [j for j, _ in sc.pairs[i]]
        c = H.Context(0)
        c.set_params(H.default_params(precision="fp32"))
        c.set_views_pooled(pool, ids)
        c.set_canny(np.zeros((sc.cameras[i].height, sc.cameras[i].width), np.uint32))
        ctxs.append(c)
        seeds.append(300 + i)
    b = H.Batch(ctxs)
    b.set_graph(True)
    b.run(seeds)
    checked = 0
    for i, c in enumerate(ctxs):
        ids = [i] + [j for j, _ in sc.pairs[i]]
        cams = [O.camera_from_scene(sc.cameras[k]) for k in ids]
        imgs = [sc.images[k].astype(np.float32) for k in ids]
        prm = O.params(len(cams), cams[0])
        w, h = cams[0].width, cams[0].height
        planes, costs = c.planes(), c.costs()
        if w * h <= 1024 * 768:
            r = O.run_patchmatch(cams, imgs, prm, seeds[i])
            rows = slice(0, h)
            assert np.array_equal(c.texture().view(np.uint32), r["texture"].view(np.uint32)), i
        else:
            y0 = h // 2 - 32
            r = O.run_patchmatch(cams, imgs, prm, seeds[i], rows=(y0 - 180, y0 + 64 + 84))
            rows = slice(y0, y0 + 64)
        same_p = np.array_equal(planes[rows].view(np.uint32), r["planes"][rows].view(np.uint32))
        same_c = np.array_equal(costs[rows].view(np.uint32), r["costs"][rows].view(np.uint32))
        print(f"view {i}: {w}x{h}, {len(ids) - 1} sources, rows {rows.start}-{rows.stop}: planes "
              f"{'equal' if same_p else 'DIFFER'}, costs {'equal' if same_c else 'DIFFER'}", flush=True)
        assert same_p and same_c, i
        checked += 1
    assert checked == len(ctxs)
    b.close()
    for c in ctxs:
        c.close()
    pool.close()
