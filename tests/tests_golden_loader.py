"""Loader of the committed fixture tests/golden/tiny_densify.npz."""
import os

import numpy as np

from patchmatch_amd.params import config0_params

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS = dict(iterations=2, seed=1234)


def load_fixture():
    fx = dict(np.load(os.path.join(HERE, "golden", "tiny_densify.npz")))
    return fx, config0_params(**PARAMS)
