"""The multi-scale pipeline (main.cpp:1183-1352: level images and cameras,
photometric pass at the coarsest level, JBU + hierarchy pass at the finer
ones, prior passes at distances 2, 1 and 0, three geometric passes per level,
RunFusion) through libhpm_host.so on the GPU in plain FP32, against the same
schedule restated on the CPU oracle (oracle/oracle_pipeline.py,
run_pyramid_pipeline): every .dmb file the library leaves -- the finest
level's depths, depths_geom, normals, costs, confidence and the texture map
of every level -- must equal the oracle's bit for bit, and the fused points
must be the oracle's, in order.

Case: 6 views of 320 x 240 (5 sources each) with a size bound of 80, so the
levels are 80 x 60, 160 x 120 and 320 x 240 and the prior passes run at
distances 0 (every level), 1 (the two finer levels) and 2 (the finest).  The
oracle takes the edge map and the Delaunay triangles from the library, as in
the single-level test."""
import os
import time

import numpy as np
import pytest

import oracle_pipeline as OP
from conftest import gpu_available
from test_gpu_c1_pipeline import read_ply_xyz

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

SEED = 23


def test_pyramid_pipeline_files_equal_oracle(tmp_path):
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=4, n_planes=8, n_cams=6, width=320, height=240, f=262.5, n_src=5)
    folder = str(tmp_path)
    scenes.write_dense_folder(sc, folder, depth_lo=1.0, depth_hi=5.0)
    print("case: pyramid pipeline 6 views 320x240, 3 levels (size bound 80), FP32 on the GPU vs the oracle",
          flush=True)
    t0 = time.perf_counter()
    n_gpu = P.run_pyramid_pipeline(folder, initial_image_scale=1, size_bound=80, precision="fp32", seed=SEED)
    t1 = time.perf_counter()
    files, (pairs, pts) = OP.run_pyramid_pipeline(folder, 1, 80, SEED, P.canny, P.delaunay)
    t2 = time.perf_counter()
    print(f"GPU pipeline {t1 - t0:.1f} s, oracle pipeline {t2 - t1:.1f} s, fused {n_gpu} / {len(pairs)}", flush=True)
    assert {name for _, name in files} == {"depths", "depths_geom", "normals", "costs", "confidence", "texture0",
                                           "texture1", "texture2"}
    out = os.path.join(folder, "HPM_MVS_plusplus")
    bad = []
    for (i, name), want in sorted(files.items()):
        got = P.read_dmb(os.path.join(out, f"2333_{i:08d}", name + ".dmb"))
        assert got.shape == want.shape, (i, name)
        diff = got.view(np.uint32) != np.ascontiguousarray(want, np.float32).view(np.uint32)
        if diff.any():
            bad.append((i, name, int(diff.sum())))
    assert not bad, f"files that differ from the oracle's (view, file, differing values): {bad[:12]}"
    xyz = read_ply_xyz(os.path.join(out, "HPM_MVS_plusplus.ply"))
    assert n_gpu == len(pairs) == len(xyz)
    np.testing.assert_array_equal(xyz.view(np.uint32), np.ascontiguousarray(pts[:, :3]).view(np.uint32))
