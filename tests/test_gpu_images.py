"""Host uploads with a row pitch (pmp_image_put, pmp_image_put_thumbnail):
the rows go through the context's pinned staging buffer, and a pitched
source gives the same pooled image bit for bit as a contiguous one."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_u8p = ctypes.POINTER(ctypes.c_uint8)


def test_pitched_image_put(gpu_ctx):
    rng = np.random.default_rng(5)
    for H, W, pitch in ((960, 1920, 2048), (97, 131, 131), (5, 3, 64)):
        img = rng.integers(0, 256, (H, W), dtype=np.uint8)
        pad = rng.integers(0, 256, (H, pitch), dtype=np.uint8)
        pad[:, :W] = img
        print(f"[images] {W}x{H} pitch {pitch}", flush=True)
        rc = gpu_ctx.L.pmp_image_put(gpu_ctx.h, 7001, pad.ctypes.data_as(_u8p), W, H, pitch, 0)
        assert rc == 0, rc
        assert np.array_equal(gpu_ctx.image_get(7001, (H, W)), img)
        gpu_ctx.image_put(7002, img)
        assert np.array_equal(gpu_ctx.image_get(7002, (H, W)), img)


def test_pitched_thumbnail(gpu_ctx):
    rng = np.random.default_rng(6)
    H, W, C, pitch = 480, 960, 3, 960 * 3 + 128
    img = rng.integers(0, 256, (H, W, C), dtype=np.uint8)
    pad = rng.integers(0, 256, (H, pitch), dtype=np.uint8)
    pad[:, :W * C] = img.reshape(H, W * C)
    gpu_ctx.image_put_thumbnail(7003, img, (120, 240))
    want = gpu_ctx.image_get(7003, (120, 240))
    rc = gpu_ctx.L.pmp_image_put_thumbnail(gpu_ctx.h, 7004, pad.ctypes.data_as(_u8p), W, H, pitch, C, 2, 240, 120, 0)
    assert rc == 0, rc
    assert np.array_equal(gpu_ctx.image_get(7004, (120, 240)), want)
