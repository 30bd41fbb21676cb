"""The C-ABI library without a GPU: it loads, exports every entry point that
include/pmp/pmp.h declares, its parameter struct matches the Python and
oracle descriptions, and it refuses to run (no CPU fallback)."""
import ctypes
import os
import re

import pytest

from oracle import or_py
from patchmatch_amd import _lib
from patchmatch_amd.params import Params, default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "pmp", "pmp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(_lib.EXPORTS) | {"pmp_debug_copy"}


def test_param_struct_layout_and_defaults():
    L = _lib.load()
    p = Params()
    L.pmp_default_params(ctypes.byref(p))
    d = default_params()
    for f, _ in Params._fields_:
        assert getattr(p, f) == pytest.approx(getattr(d, f), rel=1e-7), f
    size, off_seed, off_min_depth, off_last = or_py.params_layout()
    assert ctypes.sizeof(Params) == size == 120
    assert (Params.seed.offset, Params.min_depth.offset, Params.median_rel_thr.offset) == (off_seed, off_min_depth,
                                                                                          off_last)


def test_no_gpu_means_loud_failure():
    L = _lib.load()
    h = ctypes.c_void_p()
    n = ctypes.c_int(0)
    hip = ctypes.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is visible")
    assert L.pmp_create(ctypes.byref(h), 0) == -3          # PMP_EDEVICE
    with pytest.raises(_lib.PmpError):
        _lib.Context(0)


def test_null_arguments_rejected():
    L = _lib.load()
    assert L.pmp_create(None, 0) == -1
    assert L.pmp_synchronize(None) == -1
    assert L.pmp_last_error(None) == b"null context"
