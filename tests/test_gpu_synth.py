"""The GPU ray caster of the benchmark inputs (libpmpsynth.so) against the
numpy renderer it restates (synth.render): same scene, same camera."""
import numpy as np
import pytest

from patchmatch_amd import synth

pytestmark = pytest.mark.gpu


def test_render_gpu_matches_numpy():
    scene = synth.make_scene(5, spheres=20, tube=True, cam_region=2.7)
    for k, pose in enumerate(synth.trajectory(5, 3, 0.4)):
        img_n, dep_n = synth.render(scene, pose, 160, 80)
        img_g, dep_g = synth.render_gpu(scene, pose, 160, 80)
        rel = np.abs(dep_g - dep_n) / dep_n
        assert np.quantile(rel, 0.999) < 1e-6, (k, rel.max())
        diff = np.abs(img_g - img_n)
        assert np.quantile(diff, 0.999) < 1e-3, (k, diff.max())
