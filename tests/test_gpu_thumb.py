"""Thumbnails on the GPU (pmp_image_put_thumbnail, pmp_resize) against the
oracle's restatement of the 8-bit fixed-point cvtColor/INTER_LINEAR, bit for
bit, for downscales, upscales, odd sizes and 1/3/4 channels."""
import numpy as np
import pytest

from oracle import or_py

pytestmark = pytest.mark.gpu

SIZES = [((960, 1920), (480, 960)), ((2880, 5760), (960, 1920)), ((301, 517), (97, 203)), ((64, 128), (100, 211))]


@pytest.mark.parametrize("src,dst", SIZES)
def test_grey_thumbnail_bitwise(gpu_ctx, src, dst):
    rng = np.random.default_rng(sum(src))
    bgr = rng.integers(0, 256, (*src, 3), dtype=np.uint8)
    gpu_ctx.image_put_thumbnail(77, bgr, dst, color_order=2)
    got = gpu_ctx.image_get(77, dst)
    want = or_py.resize_linear(or_py.to_gray(bgr, order=2), dst[1], dst[0])
    np.testing.assert_array_equal(got, want)
    gpu_ctx.image_drop(77)


@pytest.mark.parametrize("channels", [1, 3, 4])
def test_resize_bitwise(gpu_ctx, channels):
    rng = np.random.default_rng(channels)
    shape = (481, 962) if channels == 1 else (481, 962, channels)
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    for dst in [(240, 480), (1000, 2000), (13, 17)]:
        np.testing.assert_array_equal(gpu_ctx.resize(img, dst), or_py.resize_linear(img, dst[1], dst[0]))
