"""The host library's JPEG decoder (images/%08d.jpg, read by cv::imread in
main.cpp:125 and main.cpp:213) against Pillow's decoder, which is
libjpeg-turbo with the defaults OpenCV's imread also uses (accurate integer
IDCT, fancy upsampling, BT.601 YCbCr -> RGB): every sample must match
bit for bit.  Pillow also writes the files, so the cases cover what an
encoder emits: 4:4:4 / 4:2:2 / 4:2:0 / 4:1:1 subsampling, grey, optimised
Huffman tables, restart intervals, odd sizes (partial MCUs), extreme
quantisation, and progressive files (spectral selection and successive
approximation, end-of-band runs)."""
import ctypes
import io

import numpy as np
import pytest

from hpm_mi355x import pipeline as P

Image = pytest.importorskip("PIL.Image")
ImageFile = pytest.importorskip("PIL.ImageFile")


def _image(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    base = np.stack([(xx * 3 + yy) % 256, (yy * 5) % 256, ((xx * yy) // 7) % 256], -1)
    return np.clip(base + rng.integers(-40, 40, base.shape), 0, 255).astype(np.uint8)


def _encode(arr, **kw):
    buf = io.BytesIO()
    # Pillow's encoder holds a progressive file whole in a block of at most
    # ImageFile.MAXBLOCK bytes: give it room for noisy, high-quality images
    old = ImageFile.MAXBLOCK
    ImageFile.MAXBLOCK = max(old, 4 * arr.size + 65536)
    try:
        Image.fromarray(arr).save(buf, format="JPEG", **kw)
    finally:
        ImageFile.MAXBLOCK = old
    return buf.getvalue()


def _same_as_pillow(data):
    ref = np.asarray(Image.open(io.BytesIO(data)))
    got = P.decode_jpeg(data)
    assert got.dtype == np.uint8 and got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


SIZES = [(8, 8), (17, 9), (64, 48), (131, 77), (300, 401)]


@pytest.mark.parametrize("progressive", [False, True], ids=["sequential", "progressive"])
@pytest.mark.parametrize("subsampling", [0, 1, 2], ids=["444", "422", "420"])
@pytest.mark.parametrize("size", SIZES, ids=[f"{w}x{h}" for w, h in SIZES])
def test_colour_matches_pillow(size, subsampling, progressive):
    arr = _image(*size, seed=size[0] * 7 + subsampling)
    for q in (30, 90):
        _same_as_pillow(_encode(arr, quality=q, subsampling=subsampling, progressive=progressive))
    _same_as_pillow(_encode(arr, quality=75, subsampling=subsampling, progressive=progressive, optimize=True))


@pytest.mark.parametrize("progressive", [False, True], ids=["sequential", "progressive"])
def test_grey_and_restarts_match_pillow(progressive):
    arr = _image(131, 77, seed=5)
    grey = np.asarray(Image.fromarray(arr).convert("L"))
    _same_as_pillow(_encode(grey, quality=80, progressive=progressive))
    _same_as_pillow(_encode(arr, quality=85, progressive=progressive, restart_marker_blocks=3))
    _same_as_pillow(_encode(arr, quality=85, progressive=progressive, restart_marker_rows=1))


def test_quality_extremes_and_411():
    arr = _image(97, 61, seed=9)
    for q in (1, 5, 100):
        _same_as_pillow(_encode(arr, quality=q))
    # 4:1:1 (h 4, v 1): the luma MCU is 32 x 8 and chroma rows are replicated 4x
    _same_as_pillow(_encode(arr, quality=90, subsampling="4:1:1"))


def test_flat_image_decodes_to_the_constant():
    # a constant block is its DC term alone: the decoded value is the input
    # to within the DC quantisation step, identical in every pixel
    arr = np.full((24, 40, 3), (200, 30, 90), np.uint8)
    got = P.decode_jpeg(_encode(arr, quality=100, subsampling=0))
    assert (got == got[0, 0]).all()
    assert np.abs(got[0, 0].astype(int) - [200, 30, 90]).max() <= 2


def test_size_query_and_short_buffer():
    data = _encode(_image(33, 21, seed=1), quality=70)
    L = P.host_lib()
    buf = np.frombuffer(data, np.uint8)
    w, h, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ptr = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    assert L.hpm_host_decode_jpeg(ptr, buf.size, ctypes.byref(w), ctypes.byref(h), ctypes.byref(c), None, 0) == 0
    assert (w.value, h.value, c.value) == (33, 21, 3)
    short = np.zeros(33 * 21 * 3 - 1, np.uint8)
    rc = L.hpm_host_decode_jpeg(ptr, buf.size, ctypes.byref(w), ctypes.byref(h), ctypes.byref(c),
                                short.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), short.size)
    assert rc != 0 and b"smaller" in L.hpm_host_last_error()


def test_unsupported_and_corrupt_files_fail_with_a_reason():
    arr = _image(30, 20, seed=2)
    data = _encode(arr, quality=75)
    i = data.find(b"\xff\xc0")
    arithmetic = data[:i + 1] + b"\xc9" + data[i + 2:]  # SOF9: arithmetic-coded sequential
    with pytest.raises(P.HpmError, match="Huffman-coded"):
        P.decode_jpeg(arithmetic)
    buf = io.BytesIO()
    Image.fromarray(arr).convert("CMYK").save(buf, format="JPEG")
    with pytest.raises(P.HpmError, match="components"):
        P.decode_jpeg(buf.getvalue())
    with pytest.raises(P.HpmError, match="not a JPEG"):
        P.decode_jpeg(b"P6\n2 2\n255\n" + bytes(12))
    for cut in (2, 100):
        with pytest.raises(P.HpmError):
            P.decode_jpeg(data[:cut])
    # a header claiming 65535 x 65535 pixels is refused before anything is
    # allocated for it (the SOF0 segment: length, precision, height, width)
    huge = data[:i + 5] + b"\xff\xff\xff\xff" + data[i + 9:]
    with pytest.raises(P.HpmError, match="2\\^28"):
        P.decode_jpeg(huge)
    # a file that ends before its EOI marker decodes to the same pixels
    # (libjpeg warns and goes on; the marker carries no data)
    np.testing.assert_array_equal(P.decode_jpeg(data[:-2]), P.decode_jpeg(data))
