"""The single-level pipeline (main.cpp:1249-1291: photometric pass,
ConfidenceEvaluation, prior pass, three geometric passes, RunFusion) through
libhpm_host.so on the GPU in plain FP32, against the same pipeline restated
on the CPU oracle (oracle/oracle_pipeline.py): every .dmb file the library
writes -- depths, depths_geom, normals, costs, confidence, texture0 of every
view -- must equal the oracle's bit for bit, and the fused points (RunFusion's
PointCloud, read back from HPM_MVS_plusplus.ply) must be the oracle's, in
order.

Cases: a small folder (6 views of 160 x 120, 5 sources), and C1: 10 cameras
of 640 x 480 (fx = fy = 525, cx = 320, cy = 240) on a 30 degree arc over 8
textured planes, pair.txt with the top 9 sources, depth range [1, 5],
--scale 1 (one pyramid level).  The oracle takes two inputs from the
library: the edge map and the Delaunay triangles of its own support points
(neither is pinned against OpenCV; see the oracle module)."""
import os
import time

import numpy as np
import pytest

import oracle_pipeline as OP
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

SEED = 21
CASES = {
    "small": dict(seed=4, n_planes=8, n_cams=6, width=160, height=120, f=131.25, n_src=5),
    "c1": dict(seed=1, n_planes=8, n_cams=10, width=640, height=480, f=525.0, n_src=9),
}


def read_ply_xyz(path):
    """x y z of HPM_MVS_plusplus.ply (float x y z, uchar r g b per vertex)."""
    with open(path, "rb") as fh:
        data = fh.read()
    hdr, body = data.split(b"end_header\n", 1)
    n = int([l for l in hdr.decode().splitlines() if l.startswith("element vertex")][0].split()[-1])
    rec = np.dtype([("xyz", "<f4", 3), ("rgb", "u1", 3)])
    assert len(body) == n * rec.itemsize
    return np.frombuffer(body, rec)["xyz"]


@pytest.mark.parametrize("case", ["small", "c1"])
def test_pipeline_files_equal_oracle_pipeline(tmp_path, case):
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(**CASES[case])
    folder = str(tmp_path)
    scenes.write_dense_folder(sc, folder, depth_lo=1.0, depth_hi=5.0)
    print(f"case: single-level pipeline {case}, FP32 on the GPU vs the oracle pipeline", flush=True)
    t0 = time.perf_counter()
    n_gpu = P.run_pipeline(folder, precision="fp32", seed=SEED)
    t1 = time.perf_counter()
    files, (pairs, pts) = OP.run_pipeline(folder, SEED, P.canny, P.delaunay)
    t2 = time.perf_counter()
    print(f"GPU pipeline {t1 - t0:.1f} s, oracle pipeline {t2 - t1:.1f} s, fused {n_gpu} / {len(pairs)}", flush=True)
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
