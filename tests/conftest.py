import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libpmp.so")


_cache = {}


def cached(key, fn):
    if key not in _cache:
        _cache[key] = fn()
    return _cache[key]


@pytest.fixture(scope="session")
def kf_c0():
    """Config 0: 480x240 box room, reference + 2 neighbours (seed 0)."""
    from patchmatch_amd import synth
    return cached("c0", lambda: synth.config0(0))


@pytest.fixture(scope="session")
def kf_small6():
    """A reduced config-1 scene (spheres + tube, 6 neighbours) at 320x160, with
    config 0's texture scale (the 1920-pixel scale would alias at 320 pixels)."""
    from patchmatch_amd import synth
    return cached("s6", lambda: synth.config1(3, W=320, H=160, n_views=6, base_freq=synth.TEX_FREQ_C0))


@pytest.fixture(scope="session")
def gpu_ctx():
    from patchmatch_amd._lib import Context
    ctx = Context(0)
    yield ctx
    ctx.close()


def gt_planes(kf):
    """Planes (H, W, 4) through the ground-truth surface (normal from depth differences)."""
    from patchmatch_amd import synth
    B = synth.bearings(kf.W, kf.H)
    X = kf.gt_depth[..., None] * B
    n = np.cross(np.roll(X, -1, 1) - np.roll(X, 1, 1), np.roll(X, -1, 0) - np.roll(X, 1, 0))
    den = np.sum(n * X, -1, keepdims=True)
    p4 = np.zeros((kf.H, kf.W, 4), np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        p4[..., :3] = np.nan_to_num(n / den)
    return p4


def random_planes(kf, seed, jitter=0.2):
    """GT planes scaled by a random factor per pixel (varied, realistic hypotheses)."""
    rng = np.random.default_rng(seed)
    p4 = gt_planes(kf)
    p4[..., :3] *= rng.uniform(1 - jitter, 1 + jitter, (kf.H, kf.W, 1)).astype(np.float32)
    p4[..., :3] += rng.normal(0, 0.05, (kf.H, kf.W, 3)).astype(np.float32)
    return p4
