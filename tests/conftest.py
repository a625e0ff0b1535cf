import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hpm-mvs_plusplus_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


_GPU = None


def gpu_available():
    """True when the HIP library loads and a HIP device answers (no torch import)."""
    global _GPU
    if _GPU is None:
        try:
            import hpm_mi355x as H
            H.Context(0).close()
            _GPU = True
        except Exception:
            _GPU = False
    return _GPU


def nudged_cameras(cams):
    """Copies of oracle cameras with every source camera's translation one
    float ulp up (the reference camera unchanged): two FP64 oracle runs on
    the original and the nudged cameras bound how closely any arithmetic can
    follow the FP64 result of a chaotic PatchMatch."""
    import ctypes
    import oracle_py as O
    out = []
    for i, c in enumerate(cams):
        d = O.orc_camera()
        ctypes.pointer(d)[0] = c
        if i > 0:
            for k in range(3):
                d.t[k] = float(np.nextafter(np.float32(c.t[k]), np.float32(np.inf)))
        out.append(d)
    return out


@pytest.fixture(scope="session")
def small_scene():
    """160x120, 5 cameras, reference view 2 with 4 source views."""
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=1, n_cams=5, width=160, height=120, f=131.25, n_src=4)
    ids, cams, imgs = scenes.reference_view_problem(sc, 2, 4)
    return sc, ids, cams, imgs


@pytest.fixture(scope="session")
def wide_scene():
    """128x96 with 10 source views (exercises V = 10)."""
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=3, n_cams=11, width=128, height=96, f=105.0, n_src=10)
    ids, cams, imgs = scenes.reference_view_problem(sc, 5, 10)
    return sc, ids, cams, imgs
