"""Oracle pins of the thumbnail pipeline (cv::cvtColor BGR2GRAY and
cv::resize INTER_LINEAR on 8-bit data; tracking_module.cc:82-93,
data/frame.cc:53-54, :136-138).  OpenCV is absent, so parity with it is
unpinned; the restatement of its published fixed-point formulas is pinned by
what they fix exactly: identity size, 2x downscale = rounded 2x2 mean,
constants, grey pixels, and the float bilinear within one level."""
import numpy as np

from oracle import or_py


def test_identity_and_constants():
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    np.testing.assert_array_equal(or_py.resize_linear(img, 53, 37), img)
    for c in (0, 1, 128, 254, 255):
        flat = np.full((41, 67, 3), c, np.uint8)
        for dW, dH in [(20, 11), (67, 41), (130, 90), (7, 5)]:
            assert np.all(or_py.resize_linear(flat, dW, dH) == c)


def test_half_size_is_rounded_box_mean():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (48, 96, 3), dtype=np.uint8).astype(np.int64)
    got = or_py.resize_linear(img.astype(np.uint8), 48, 24).astype(np.int64)
    s = img[0::2, 0::2] + img[0::2, 1::2] + img[1::2, 0::2] + img[1::2, 1::2]
    np.testing.assert_array_equal(got, (s + 2) // 4)


def _float_bilinear(img, dW, dH):
    H, W = img.shape

    def taps(n, d):
        f = (np.arange(d) + 0.5) * (n / d) - 0.5
        s = np.floor(f).astype(int)
        f = f - s
        f[s < 0] = 0.0
        s[s < 0] = 0
        last = s >= n - 1
        f[last] = 0.0
        s[last] = n - 1
        return s, np.minimum(s + 1, n - 1), f
    xs0, xs1, fx = taps(W, dW)
    ys0, ys1, fy = taps(H, dH)
    a = img.astype(np.float64)
    top = a[ys0][:, xs0] * (1 - fx) + a[ys0][:, xs1] * fx
    bot = a[ys1][:, xs0] * (1 - fx) + a[ys1][:, xs1] * fx
    return top * (1 - fy[:, None]) + bot * fy[:, None]


def test_resize_within_one_level_of_float_bilinear():
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (120, 240), dtype=np.uint8)
    for dW, dH in [(80, 40), (200, 100), (301, 151), (240, 120), (64, 30)]:
        got = or_py.resize_linear(img, dW, dH).astype(np.float64)
        assert np.max(np.abs(got - _float_bilinear(img, dW, dH))) <= 1.0, (dW, dH)
    ramp = np.tile(np.arange(240, dtype=np.uint8), (8, 1))
    got = or_py.resize_linear(ramp, 80, 8)[0].astype(np.float64)
    want = np.clip((np.arange(80) + 0.5) * 3 - 0.5, 0, 239)
    assert np.max(np.abs(got - want)) <= 1.0


def test_grey_conversion():
    rng = np.random.default_rng(3)
    bgr = rng.integers(0, 256, (30, 40, 3), dtype=np.uint8)
    g = or_py.to_gray(bgr, order=2).astype(np.float64)
    b, gg, r = (bgr[..., k].astype(np.float64) for k in range(3))
    assert np.max(np.abs(g - (0.299 * r + 0.587 * gg + 0.114 * b))) <= 0.51
    np.testing.assert_array_equal(or_py.to_gray(bgr[..., ::-1], order=1), or_py.to_gray(bgr, order=2))
    grey = np.repeat(rng.integers(0, 256, (10, 10, 1), dtype=np.uint8), 3, axis=2)
    np.testing.assert_array_equal(or_py.to_gray(grey), grey[..., 0])
    rgba = np.concatenate([bgr[..., ::-1], np.full((30, 40, 1), 7, np.uint8)], -1)
    np.testing.assert_array_equal(or_py.to_gray(rgba, order=1), or_py.to_gray(bgr, order=2))
