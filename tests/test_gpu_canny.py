"""The device edge map (hpm_canny: cv::Canny(image, low, high) of
ProcessProblem, main.cpp:280-289) against the host CannyEdges, byte for byte,
on grey and colour images, sizes that are not multiples of the 32 x 32
hysteresis tile, several threshold pairs, and a weak spiral edge fed from one
strong end, whose hysteresis has to cross many tiles in both directions."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


def _textured(w, h, channels, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    base = 128 + 60 * np.sin(xx / 7.0) * np.cos(yy / 11.0) + 40 * np.sign(np.sin((xx + 2 * yy) / 23.0))
    img = base[..., None] + rng.normal(0, 12, (h, w, channels))
    if channels == 3:
        img[..., 0] += 25 * np.sin(yy / 5.0)
        img[..., 2] -= 20 * np.cos(xx / 9.0)
    img = np.clip(img, 0, 255).astype(np.uint8)
    return img[..., 0] if channels == 1 else img


def _spiral(n=400):
    """A one-pixel-wide dark spiral on a flat background, brightness step 30
    (partly weak with low 50 / high 150 after Sobel), with a short step-120
    stretch at its outer end (strong)."""
    img = np.full((n, n), 140, np.uint8)
    c = n / 2
    t = np.linspace(0, 10 * 2 * np.pi, 60000)
    r = 4 + t * (c - 8) / t[-1]
    x = np.round(c + r * np.cos(t)).astype(int)
    y = np.round(c + r * np.sin(t)).astype(int)
    img[y, x] = 110
    img[y[-400:], x[-400:]] = 20
    return img


@pytest.mark.parametrize("w,h,channels", [(97, 61, 1), (320, 240, 3), (1000, 37, 3), (33, 500, 1)])
@pytest.mark.parametrize("low,high", [(50.0, 150.0), (10.0, 30.0), (100.0, 100.0)])
def test_canny_equals_host(w, h, channels, low, high):
    import hpm_mi355x as H
    from hpm_mi355x import pipeline as P
    img = _textured(w, h, channels, seed=w + h)
    print(f"case: hpm_canny {w}x{h}x{channels} thresholds {low}/{high}", flush=True)
    got = H.canny(img, low, high)
    want = P.canny(img, low, high)
    assert 0 < int((want > 0).sum()) < w * h
    np.testing.assert_array_equal(got, want)


def test_canny_hysteresis_crosses_tiles():
    import hpm_mi355x as H
    from hpm_mi355x import pipeline as P
    img = _spiral()
    print("case: hpm_canny 400x400 spiral", flush=True)
    got = H.canny(img, 50.0, 150.0)
    want = P.canny(img, 50.0, 150.0)
    np.testing.assert_array_equal(got, want)
    # the weak pixels that hysteresis adds lie in many tiles
    added = (got > 0) & ~(P.canny(img, 150.0, 150.0) > 0)
    ys, xs = np.nonzero(added)
    assert added.sum() > 3000
    assert len(set(zip((ys // 32).tolist(), (xs // 32).tolist()))) > 100


def test_canny_step_edge_closed_form():
    """A vertical step: the two columns beside the step have the same Sobel
    response; the suppression keeps the left one (> against the left
    neighbour, >= against the right), the last dark column.  Nothing else is
    an edge; a constant image has no edge."""
    import hpm_mi355x as H
    img = np.zeros((50, 70), np.uint8)
    img[:, 35:] = 200
    e = H.canny(img)
    assert set(np.nonzero(e.any(axis=0))[0].tolist()) == {34}
    assert e[:, 34].all()
    assert not H.canny(np.full((40, 40, 3), 77, np.uint8)).any()
