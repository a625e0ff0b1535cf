"""The HPM facade as a drop-in for main.cpp's ProcessProblem: a C++ program
(tests/facade_process_problem.cpp) that runs ProcessProblem's call sequence
(main.cpp:257-631) -- InuputInitialization, CudaSpaceInitialization,
CudaCannyInitialization, RunPatchMatch, the per-pixel GetPlaneHypothesis /
GetCost / GetTexture loops, and in the prior pass ReadDmb of the previous
maps, GetSupportPoints_Classify_Check, DelaunayTriangulation, the raster,
GetPriorPlaneParams, GetDepthFromPlaneParam with GetMin/MaxDepth,
CudaPlanarPriorInitialization, CudaHypothesesReload, CudaPlanarPriorRelease,
CudaSpaceRelease, ReleaseProblemHostMemory -- against hpmhost::HPM writes
the same files, bit for bit, as the library's own ProcessProblem."""
import os
import subprocess

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "tests", "_build", "facade_process_problem")


def test_facade_call_sequence_writes_the_library_files(tmp_path):
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    assert os.path.exists(PROGRAM), "build the library first (hpm-mvs_plusplus_amd/build.py)"
    sc = scenes.make_plane_scene(seed=6, n_cams=5, width=160, height=120, f=131.25, n_src=4)
    lib_dir, fac_dir = str(tmp_path / "library"), str(tmp_path / "facade")
    for d in (lib_dir, fac_dir):
        scenes.write_dense_folder(sc, d)
    print("case: facade ProcessProblem sequence vs library, 5 views 160x120, photometric + prior", flush=True)
    seed = 40
    for i in range(5):
        P.process_problem(lib_dir, i, precision="fp32", seed=seed + i)
    P.confidence(lib_dir)
    for i in range(5):
        P.process_problem(lib_dir, i, prior=True, precision="fp32", seed=seed + 1000 + i)
    r = subprocess.run([PROGRAM, fac_dir, str(seed)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    for i in range(5):
        for name in ("depths", "normals", "costs", "texture0", "confidence"):
            a = P.read_dmb(os.path.join(lib_dir, "HPM_MVS_plusplus", f"2333_{i:08d}", name + ".dmb"))
            b = P.read_dmb(os.path.join(fac_dir, "HPM_MVS_plusplus", f"2333_{i:08d}", name + ".dmb"))
            assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32)), (i, name)
