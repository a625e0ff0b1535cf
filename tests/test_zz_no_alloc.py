"""hpm_canny_noalloc, the entry point that allocates nothing, against
hpm_canny (the same edge map computed with per-call device buffers): two
calls in a row, the second on a smaller image, so that the static scratch
array's leftovers from the first call must not leak into the second; and the
size limit, refused before any launch."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


def _image(w, h, channels, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = 128 + 60 * np.sin(x / 9.0) * np.cos(y / 13.0)
    base[(x - w // 2) ** 2 + (y - h // 2) ** 2 < (min(w, h) // 4) ** 2] += 50
    img = base[..., None] + rng.normal(0, 6, (h, w, channels))
    img = np.clip(img, 0, 255).astype(np.uint8)
    return img[..., 0] if channels == 1 else img


def test_canny_noalloc_equals_canny_on_two_calls():
    import hpm_mi355x as H
    L = H.lib()
    u8p = ctypes.POINTER(ctypes.c_uint8)
    L.hpm_canny_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_void_p]
    L.hpm_canny_noalloc.restype = ctypes.c_int
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipSetDevice(0) == 0
    cases = [(640, 480, 1, 1), (320, 200, 3, 2)]  # the second call is smaller
    for (w, h, c, seed) in cases:
        print(f"case: hpm_canny_noalloc {w}x{h}x{c}", flush=True)
        img = np.ascontiguousarray(_image(w, h, c, seed))
        expect = H.canny(img)
        assert expect.any() and not expect.all()
        d_img, d_out = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(d_img), ctypes.c_size_t(img.nbytes)) == 0
        assert hip.hipMalloc(ctypes.byref(d_out), ctypes.c_size_t(w * h)) == 0
        try:
            assert hip.hipMemcpy(d_img, img.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(img.nbytes), 1) == 0
            assert hip.hipMemset(d_out, 7, ctypes.c_size_t(w * h)) == 0  # stale output bytes must be overwritten
            assert hip.hipDeviceSynchronize() == 0
            rc = L.hpm_canny_noalloc(d_img, w, h, c, 50.0, 150.0, d_out)
            assert rc == 0, rc
            got = np.empty((h, w), np.uint8)
            assert hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), d_out, ctypes.c_size_t(w * h), 2) == 0
            np.testing.assert_array_equal(got, expect)
        finally:
            hip.hipFree(d_img)
            hip.hipFree(d_out)
    # beyond the scratch array's size: refused with HPM_E_TOO_LARGE before any launch
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(256)) == 0
    try:
        assert L.hpm_canny_noalloc(d, 3072, 2049, 1, 50.0, 150.0, d) == -5
        assert L.hpm_canny_noalloc(None, 16, 16, 1, 50.0, 150.0, d) == -1
    finally:
        hip.hipFree(d)
