"""Faults a dense folder can hold, fed to the host pipeline's C ABI
(include/hpm_pipeline.h): a missing or truncated pair.txt
(GenerateSampleList, main.cpp:171-200), missing, corrupt or mis-shaped .dmb
maps (readDepthDmb / readNormalDmb, HPM.cpp:26-96), a missing or malformed
camera (ReadCamera, HPM.cpp:159-184), a source view that pair.txt does not
list, and out-of-range arguments.  Each returns HPM_E_INVALID with a message
that names the file or the argument, before any device work: none of these
needs a GPU.  The faults that surface only after the device is up (a
missing camera or a corrupt image in ProcessProblem, a source depth map of
another size in a geometric pass) are at the end, marked gpu."""
import ctypes
import os
import time

import numpy as np
import pytest

from conftest import gpu_available

HPM_E_INVALID = -1
W, H = 16, 12


def _write_dmb(path, a, type_flag=1, dims=None):
    a = np.ascontiguousarray(a, np.float32)
    h, w = a.shape[:2]
    nb = 1 if a.ndim == 2 else a.shape[2]
    hdr = np.array([type_flag] + list(dims or (h, w, nb)), np.int32)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(a.tobytes())


def _write_cam(path, f=20.0):
    with open(path, "w") as fh:
        fh.write("extrinsic\n1 0 0 0\n0 1 0 0\n0 0 1 0\n0 0 0 1\n\nintrinsic\n")
        fh.write(f"{f} 0 {W / 2}\n0 {f} {H / 2}\n0 0 1\n\n1.0 0.02 192 5.0\n")


def _write_pgm(path, w=W, h=H, header=None):
    with open(path, "wb") as fh:
        fh.write(header or f"P5\n{w} {h}\n255\n".encode())
        fh.write(np.full(w * h, 100, np.uint8).tobytes())


def _result(folder, i):
    return os.path.join(folder, "HPM_MVS_plusplus", f"2333_{i:08d}")


@pytest.fixture
def folder(tmp_path):
    """Three 16 x 12 views, each with the other two as sources, and the
    depth and normal maps a fusion reads."""
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "images"))
    os.makedirs(os.path.join(d, "cams"))
    with open(os.path.join(d, "pair.txt"), "w") as fh:
        fh.write("3\n0\n2 1 10.0 2 5.0\n1\n2 0 10.0 2 5.0\n2\n2 0 10.0 1 5.0\n")
    for i in range(3):
        _write_pgm(os.path.join(d, "images", f"{i:08d}.pgm"))
        _write_cam(os.path.join(d, "cams", f"{i:08d}_cam.txt"))
        rf = _result(d, i)
        os.makedirs(rf)
        _write_dmb(os.path.join(rf, "depths.dmb"), np.full((H, W), 3.0))
        _write_dmb(os.path.join(rf, "depths_geom.dmb"), np.full((H, W), 3.0))
        n = np.zeros((H, W, 3), np.float32)
        n[..., 2] = -1
        _write_dmb(os.path.join(rf, "normals.dmb"), n)
    return d


def _lib():
    from hpm_mi355x import pipeline as P
    return P.host_lib()


def _fusion(folder, geom=0):
    L = _lib()
    n = ctypes.c_int64(-7)
    rc = L.hpm_host_fusion(folder.encode() if folder is not None else None, geom, ctypes.byref(n))
    return rc, L.hpm_host_last_error().decode()


def _confidence(folder, geom=0):
    L = _lib()
    rc = L.hpm_host_confidence(folder.encode(), geom)
    return rc, L.hpm_host_last_error().decode()


def _invalid(rc_msg, *needles):
    rc, msg = rc_msg
    assert rc == HPM_E_INVALID, (rc, msg)
    for s in needles:
        assert s in msg, msg


# ------------------------------------------------------------------ pair.txt

def test_missing_pair_file(folder):
    os.remove(os.path.join(folder, "pair.txt"))
    _invalid(_fusion(folder), "cannot open", "pair.txt")
    _invalid(_confidence(folder), "cannot open", "pair.txt")
    L = _lib()
    n = ctypes.c_int64(0)
    rc = L.hpm_host_run_pyramid_pipeline(folder.encode(), 1, 1000, 0, 1, 0, ctypes.byref(n))
    assert rc == HPM_E_INVALID and "pair.txt" in L.hpm_host_last_error().decode()
    rc = L.hpm_host_run_pipeline(folder.encode(), 0, 1, ctypes.byref(n))
    assert rc == HPM_E_INVALID and "pair.txt" in L.hpm_host_last_error().decode()
    rc = L.hpm_host_fusion_mask(folder.encode(), 1, ctypes.byref(n))
    assert rc == HPM_E_INVALID and "pair.txt" in L.hpm_host_last_error().decode()


@pytest.mark.parametrize("text", [
    "3\n0\n2 1 10.0 2 5.0\n1\n2 0 10.0 2 5.0\n",  # a view fewer than announced
    "3\n0\n2 1 10.0 2 5.0\n1\n2 0 10.0\n2\n2 0 10.0 1 5.0\n",  # a source fewer than announced
    "3\n0\n2 1 ten 2 5.0\n",  # not a number
    "",  # empty
])
def test_truncated_pair_file(folder, text):
    with open(os.path.join(folder, "pair.txt"), "w") as fh:
        fh.write(text)
    _invalid(_fusion(folder), "truncated pair file")


def test_source_without_pair_entry(folder):
    with open(os.path.join(folder, "pair.txt"), "w") as fh:
        fh.write("3\n0\n2 1 10.0 7 5.0\n1\n2 0 10.0 2 5.0\n2\n2 0 10.0 1 5.0\n")
    _invalid(_fusion(folder), "source image without a pair.txt entry")
    with open(os.path.join(folder, "pair.txt"), "w") as fh:
        fh.write("3\n0\n2 1 10.0 -2 5.0\n1\n2 0 10.0 2 5.0\n2\n2 0 10.0 1 5.0\n")
    _invalid(_fusion(folder), "source image without a pair.txt entry")


def test_source_count_checked_before_any_pass(folder):
    """main.cpp:1209-1220 refuses a problem over MAX_IMAGES before any pass;
    a view whose sources all score <= 0 (GenerateSampleList drops them,
    main.cpp:193-195) has no source at all.  Both are refused up front by
    the pipelines and by ProcessProblem, naming the view."""
    L = _lib()
    n = ctypes.c_int64(0)
    srcs = " ".join(f"{j % 2 + 1} 1.0" for j in range(33))
    for text, needle in [
        ("3\n0\n2 1 10.0 2 5.0\n1\n2 0 0.0 2 -1.0\n2\n2 0 10.0 1 5.0\n",
         "view 1 has no source view with a positive score in pair.txt"),
        (f"3\n0\n33 {srcs}\n1\n2 0 10.0 2 5.0\n2\n2 0 10.0 1 5.0\n",
         "view 0 needs 34 images, more than HPM_MAX_IMAGES = 33"),
    ]:
        with open(os.path.join(folder, "pair.txt"), "w") as fh:
            fh.write(text)
        assert L.hpm_host_run_pyramid_pipeline(folder.encode(), 1, 1000, 0, 1, 0, ctypes.byref(n)) == HPM_E_INVALID
        assert L.hpm_host_last_error().decode() == needle
        assert L.hpm_host_run_pipeline(folder.encode(), 0, 1, ctypes.byref(n)) == HPM_E_INVALID
        assert L.hpm_host_last_error().decode() == needle
        idx = int(needle.split()[1])
        assert L.hpm_host_process_problem(folder.encode(), idx, 0, 0, 0, 0, 0, 1) == HPM_E_INVALID
        assert L.hpm_host_last_error().decode() == needle
    assert not os.path.exists(os.path.join(folder, "HPM_MVS_plusplus", "HPM_MVS_plusplus.ply"))


def test_problem_index_out_of_range(folder):
    L = _lib()
    for idx in (3, -1, 1 << 30):
        rc = L.hpm_host_process_problem(folder.encode(), idx, 0, 0, 0, 0, 0, 1)
        assert rc == HPM_E_INVALID
        assert L.hpm_host_last_error().decode() == "problem index out of range"


# ------------------------------------------------------------------ .dmb maps

def test_missing_dmb(folder):
    os.remove(os.path.join(_result(folder, 1), "depths.dmb"))
    _invalid(_fusion(folder), "cannot read", "2333_00000001/depths.dmb")
    _invalid(_confidence(folder), "cannot read", "2333_00000001/depths.dmb")
    os.remove(os.path.join(_result(folder, 2), "normals.dmb"))
    _invalid(_fusion(folder, geom=1), "cannot read", "2333_00000002/normals.dmb")


@pytest.mark.parametrize("kind", ["type", "short_body", "zero_dim", "negative_dim", "huge_dims", "header_only",
                                  "empty"])
def test_corrupt_dmb(folder, kind):
    path = os.path.join(_result(folder, 2), "depths.dmb")
    d = np.full((H, W), 3.0, np.float32)
    if kind == "type":
        _write_dmb(path, d, type_flag=2)
    elif kind == "short_body":
        _write_dmb(path, d[:-1])
        with open(path, "r+b") as fh:  # the header still says 12 rows
            fh.write(np.array([1, H, W, 1], np.int32).tobytes())
    elif kind == "zero_dim":
        _write_dmb(path, d, dims=(0, W, 1))
    elif kind == "negative_dim":
        _write_dmb(path, d, dims=(H, -W, 1))
    elif kind == "huge_dims":  # 60000 x 60000 x 3 floats (43 GB) claimed by a 784-byte file
        _write_dmb(path, d, dims=(60000, 60000, 3))
    elif kind == "header_only":
        with open(path, "wb") as fh:
            fh.write(np.array([1, H, W, 1], np.int32).tobytes())
    else:
        open(path, "wb").close()
    t0 = time.perf_counter()
    _invalid(_fusion(folder), "cannot read", "2333_00000002/depths.dmb")
    assert time.perf_counter() - t0 < 5.0  # refused from the header, nothing allocated for it


def test_dmb_trailing_bytes_accepted(folder):
    """readDepthDmb reads h * w * nb floats and stops: bytes after them do
    not make the file unreadable.  The call then gets past the maps (and
    fails on the device without a GPU, or fuses with one)."""
    path = os.path.join(_result(folder, 0), "depths.dmb")
    with open(path, "ab") as fh:
        fh.write(b"\0" * 64)
    rc, msg = _fusion(folder)
    assert "depths.dmb" not in msg


@pytest.mark.parametrize("kind", ["normals_one_channel", "normals_other_size", "depth_three_channels"])
def test_misshaped_maps(folder, kind):
    rf = _result(folder, 1)
    if kind == "normals_one_channel":
        _write_dmb(os.path.join(rf, "normals.dmb"), np.zeros((H, W), np.float32))
        _invalid(_fusion(folder), "2333_00000001/normals.dmb", "16x12x1 map where 16x12x3 is needed")
    elif kind == "normals_other_size":
        _write_dmb(os.path.join(rf, "normals.dmb"), np.zeros((H - 2, W, 3), np.float32))
        _invalid(_fusion(folder), "2333_00000001/normals.dmb", "16x10x3 map where 16x12x3 is needed")
    else:
        _write_dmb(os.path.join(rf, "depths.dmb"), np.zeros((H, W, 3), np.float32))
        _invalid(_fusion(folder), "2333_00000001/depths.dmb", "16x12x3 map where 16x12x1 is needed")
    _invalid(_confidence(folder), "2333_00000001/")


# ------------------------------------------------------------------ cameras and images

def test_missing_camera(folder):
    os.remove(os.path.join(folder, "cams", "00000002_cam.txt"))
    _invalid(_fusion(folder), "cannot open", "cams/00000002_cam.txt")


def test_malformed_camera(folder):
    with open(os.path.join(folder, "cams", "00000001_cam.txt"), "w") as fh:
        fh.write("extrinsic\n1 0 0 0\n0 1 0 0\n0 0 1 0\n0 0 0 1\n\nintrinsic\n20 0 8\n0 20\n")
    _invalid(_fusion(folder), "malformed camera file", "cams/00000001_cam.txt")


def test_missing_image(folder):
    os.remove(os.path.join(folder, "images", "00000000.pgm"))
    _invalid(_fusion(folder), "cannot read the size of image", "images/00000000")


# ------------------------------------------------------------------ arguments

def test_bad_arguments(folder):
    L = _lib()
    n = ctypes.c_int64(0)
    cases = [
        ("hpm_host_run_pyramid_pipeline", lambda: L.hpm_host_run_pyramid_pipeline(folder.encode(), 0, 1000, 0, 1, 0,
                                                                                 ctypes.byref(n))),
        ("hpm_host_run_pyramid_pipeline", lambda: L.hpm_host_run_pyramid_pipeline(folder.encode(), 1, 0, 0, 1, 0,
                                                                                 ctypes.byref(n))),
        ("hpm_host_run_pyramid_pipeline", lambda: L.hpm_host_run_pyramid_pipeline(folder.encode(), 1, 1000, 0, 1, 0,
                                                                                 None)),
        ("hpm_host_fusion", lambda: L.hpm_host_fusion(None, 1, ctypes.byref(n))),
        ("hpm_host_confidence", lambda: L.hpm_host_confidence(None, 1)),
        ("hpm_host_process_problem", lambda: L.hpm_host_process_problem(None, 0, 0, 0, 0, 0, 0, 1)),
        ("hpm_host_run_pipeline", lambda: L.hpm_host_run_pipeline(folder.encode(), 0, 1, None)),
        ("hpm_host_fusion_mask", lambda: L.hpm_host_fusion_mask(folder.encode(), 1, None)),
    ]
    for name, call in cases:
        assert call() == HPM_E_INVALID, name
        assert L.hpm_host_last_error().decode() == f"{name}: invalid argument"


def test_error_message_cleared_on_success():
    """A call that succeeds leaves no message behind."""
    from hpm_mi355x import pipeline as P
    L = _lib()
    assert L.hpm_host_fusion(None, 1, None) == HPM_E_INVALID
    assert L.hpm_host_last_error().decode() != ""
    P.canny(np.zeros((8, 8), np.uint8))
    assert L.hpm_host_last_error().decode() == ""


# ------------------------------------------------------------------ after the device is up

@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="no HIP device")
def test_gpu_problem_faults(folder):
    """ProcessProblem reads the cameras and images once the device context
    exists: a missing camera, an image whose header claims more pixels than
    the file holds, and in a geometric pass a source depth map of another
    size than its image."""
    L = _lib()

    def run(idx, geom=0):
        print(f"case: hpm_host_process_problem idx {idx} geom {geom}", flush=True)
        rc = L.hpm_host_process_problem(folder.encode(), idx, geom, 0, 0, 0, 0, 1)
        return rc, L.hpm_host_last_error().decode()

    os.rename(os.path.join(folder, "cams", "00000002_cam.txt"), os.path.join(folder, "cams", "x"))
    _invalid(run(0), "cannot open", "cams/00000002_cam.txt")
    os.rename(os.path.join(folder, "cams", "x"), os.path.join(folder, "cams", "00000002_cam.txt"))

    img = os.path.join(folder, "images", "00000001.pgm")
    _write_pgm(img, header=b"P5\n60000 60000\n255\n")
    _invalid(run(0), "cannot read image", "images/00000001")
    _write_pgm(img)

    rf = _result(folder, 2)
    _write_dmb(os.path.join(rf, "depths.dmb"), np.full((H, W - 4), 3.0))
    _invalid(run(0, geom=1), "2333_00000002/depths.dmb", "12x12x1 map where 16x12x1 is needed")
