"""ctypes binding of lib/libhpm_host.so: the dense-folder pipeline and the
planar-prior construction steps (include/hpm_pipeline.h)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import HpmError, lib as _core_lib

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB_PATH = os.environ.get("HPM_HOST_LIB", os.path.join(HERE, "lib", "libhpm_host.so"))

_lib = None


def host_lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HOST_LIB_PATH):
        raise HpmError(f"{HOST_LIB_PATH} is missing: run `python __graft_entry__.py build` (or build.py)")
    _core_lib()  # loads libhpm_mi355x.so first
    L = ctypes.CDLL(HOST_LIB_PATH)
    P, i32p, u8p, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint8), \
        ctypes.POINTER(ctypes.c_float)
    L.hpm_host_last_error.restype = ctypes.c_char_p
    L.hpm_host_last_error.argtypes = []
    L.hpm_host_process_problem.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 6 + [ctypes.c_uint64]
    L.hpm_host_confidence.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.hpm_host_fusion.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.hpm_host_run_pipeline.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64)]
    L.hpm_host_run_pyramid_pipeline.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.hpm_host_fusion_mask.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.hpm_host_stage_times.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
    L.hpm_host_reset_stage_times.argtypes = []
    L.hpm_host_reset_stage_times.restype = None
    L.hpm_host_set_verbose.argtypes = [ctypes.c_int]
    L.hpm_host_set_verbose.restype = None
    L.hpm_host_set_threads.argtypes = [ctypes.c_int]
    L.hpm_host_set_threads.restype = None
    L.hpm_host_set_prior_on_device.argtypes = [ctypes.c_int]
    L.hpm_host_set_prior_on_device.restype = None
    L.hpm_host_set_fused_min_pixels.argtypes = [ctypes.c_int64]
    L.hpm_host_set_fused_min_pixels.restype = None
    L.hpm_host_set_confidence_overlap.argtypes = [ctypes.c_int]
    L.hpm_host_set_confidence_overlap.restype = None
    L.hpm_host_decode_jpeg.argtypes = [u8p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int), u8p, ctypes.c_int64]
    L.hpm_host_support_points.argtypes = [fp, fp, ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]
    L.hpm_host_delaunay.argtypes = [ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, i32p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)]
    L.hpm_host_raster.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
    L.hpm_host_plane_through_points.argtypes = [fp, fp, fp, fp]
    L.hpm_host_canny.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, u8p]
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise HpmError(f"{what} failed ({rc}): {host_lib().hpm_host_last_error().decode()}")


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


PRECISIONS = {"fp32": 0, "fp16": 1, "bf16": 2, "fp32fast": 3}


def process_problem(folder, idx, geom=False, prior=False, mand=False, multi_geometry=False, precision="fp32", seed=1):
    _check(host_lib().hpm_host_process_problem(folder.encode(), idx, int(geom), int(prior), int(mand),
                                               int(multi_geometry), PRECISIONS[precision], seed),
           "hpm_host_process_problem")


def confidence(folder, geom=False):
    _check(host_lib().hpm_host_confidence(folder.encode(), int(geom)), "hpm_host_confidence")


def fusion(folder, geom=True):
    n = ctypes.c_int64(0)
    _check(host_lib().hpm_host_fusion(folder.encode(), int(geom), ctypes.byref(n)), "hpm_host_fusion")
    return n.value


def run_pipeline(folder, precision="fp32", seed=1):
    n = ctypes.c_int64(0)
    _check(host_lib().hpm_host_run_pipeline(folder.encode(), PRECISIONS[precision], seed, ctypes.byref(n)),
           "hpm_host_run_pipeline")
    return n.value


def run_pyramid_pipeline(folder, initial_image_scale=4, size_bound=1000, precision="fp32", seed=1, mask=False):
    """main.cpp's pyramid schedule and fusion (strict sky-mask fusion with
    mask=True) over a dense folder; returns the number of fused points (the
    PLY files are written under HPM_MVS_plusplus/)."""
    n = ctypes.c_int64(0)
    _check(host_lib().hpm_host_run_pyramid_pipeline(folder.encode(), initial_image_scale, size_bound,
                                                    PRECISIONS[precision], seed, int(mask), ctypes.byref(n)),
           "hpm_host_run_pyramid_pipeline")
    return n.value


def fusion_mask(folder, geom=True):
    """RunFusion_Sky_Strict over the folder's maps and masks/%08d.ppm; returns
    the number of points kept."""
    n = ctypes.c_int64(0)
    _check(host_lib().hpm_host_fusion_mask(folder.encode(), int(geom), ctypes.byref(n)), "hpm_host_fusion_mask")
    return n.value


STAGES = ("setup", "edges", "prior", "patchmatch", "write", "jbu", "confidence", "fusion")


def stage_times(reset=False):
    """{stage: (seconds, calls)} summed since the last reset."""
    n = len(STAGES)
    sec = (ctypes.c_double * n)()
    calls = (ctypes.c_int64 * n)()
    host_lib().hpm_host_stage_times(sec, calls, n)
    out = {k: (sec[i], calls[i]) for i, k in enumerate(STAGES)}
    if reset:
        host_lib().hpm_host_reset_stage_times()
    return out


def set_verbose(level=1):
    host_lib().hpm_host_set_verbose(int(level))


def set_threads(n=0):
    """Host threads for the order-free passes (0: default, 1: view order)."""
    host_lib().hpm_host_set_threads(int(n))


def set_fused_min_pixels(n=4000000):
    """Reference views of at least n pixels run as a one-view fused launch with
    16-column tiles (mixed precision only); the results are identical."""
    host_lib().hpm_host_set_fused_min_pixels(int(n))


def set_prior_on_device(enable=True):
    """Build the prior passes' planar prior on the device (default; only the
    Delaunay triangulation runs on the host) or entirely on the host; the
    output files are identical."""
    host_lib().hpm_host_set_prior_on_device(int(bool(enable)))


def set_confidence_overlap(enable=True):
    """Run each pass that follows a confidence evaluation alongside it (each
    view waits for its own map; default) or after it; same files."""
    host_lib().hpm_host_set_confidence_overlap(int(bool(enable)))


def decode_jpeg(data: bytes):
    """The host library's JPEG decoder (Huffman-coded sequential or progressive,
    8-bit): (h, w) grey or (h, w, 3) R G B uint8, as libjpeg's default decoding
    gives them."""
    buf = np.frombuffer(data, np.uint8)
    w, h, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L = host_lib()
    _check(L.hpm_host_decode_jpeg(_ptr(buf, ctypes.c_uint8), buf.size, ctypes.byref(w), ctypes.byref(h),
                                  ctypes.byref(c), None, 0), "decode_jpeg")
    out = np.zeros((h.value, w.value, c.value), np.uint8)
    _check(L.hpm_host_decode_jpeg(_ptr(buf, ctypes.c_uint8), buf.size, ctypes.byref(w), ctypes.byref(h),
                                  ctypes.byref(c), _ptr(out, ctypes.c_uint8), out.size), "decode_jpeg")
    return out[..., 0] if c.value == 1 else out


def support_points(costs, confidences):
    c = np.ascontiguousarray(costs, np.float32)
    f = np.ascontiguousarray(confidences, np.float32)
    h, w = c.shape
    cap = 2 * ((w + 4) // 5) * ((h + 4) // 5)
    out = np.zeros((max(cap, 1), 2), np.int32)
    n = ctypes.c_int(0)
    _check(host_lib().hpm_host_support_points(_ptr(c, ctypes.c_float), _ptr(f, ctypes.c_float), w, h,
                                              _ptr(out, ctypes.c_int32), cap, ctypes.byref(n)), "support_points")
    return out[:n.value]


def delaunay(w, h, points):
    xy = np.ascontiguousarray(np.asarray(points, np.int32).reshape(-1, 2))
    cap = 2 * len(xy) + 8
    out = np.zeros((cap, 6), np.int32)
    n = ctypes.c_int(0)
    _check(host_lib().hpm_host_delaunay(w, h, _ptr(xy, ctypes.c_int32), len(xy), _ptr(out, ctypes.c_int32), cap,
                                        ctypes.byref(n)), "delaunay")
    return out[:min(n.value, cap)]


def raster(tris, w, h):
    t = np.ascontiguousarray(np.asarray(tris, np.int32).reshape(-1, 6))
    mask = np.zeros((h, w), np.float32)
    _check(host_lib().hpm_host_raster(_ptr(t, ctypes.c_int32), len(t), w, h, _ptr(mask, ctypes.c_float)), "raster")
    return mask


def plane_through_points(X1, X2, X3):
    a = [np.ascontiguousarray(np.asarray(x, np.float32).reshape(3)) for x in (X1, X2, X3)]
    out = np.zeros(4, np.float32)
    _check(host_lib().hpm_host_plane_through_points(*[_ptr(x, ctypes.c_float) for x in a],
                                                    _ptr(out, ctypes.c_float)), "plane_through_points")
    return out


def canny(image, low=50.0, high=150.0):
    """Edge map of an (h, w) grey or (h, w, 3) R G B uint8 image; 255 = edge."""
    g = np.ascontiguousarray(image, np.uint8)
    channels = 1 if g.ndim == 2 else g.shape[2]
    out = np.zeros(g.shape[:2], np.uint8)
    _check(host_lib().hpm_host_canny(_ptr(g, ctypes.c_uint8), g.shape[1], g.shape[0], channels, low, high,
                                     _ptr(out, ctypes.c_uint8)), "canny")
    return out


def read_dmb(path):
    """A .dmb file as a (h, w) or (h, w, nb) float32 array."""
    with open(path, "rb") as fh:
        hdr = np.frombuffer(fh.read(16), np.int32)
        if hdr[0] != 1:
            raise HpmError(f"{path}: not a float dmb")
        h, w, nb = (int(x) for x in hdr[1:])
        data = np.frombuffer(fh.read(4 * h * w * nb), np.float32)
    return data.reshape((h, w) if nb == 1 else (h, w, nb))
