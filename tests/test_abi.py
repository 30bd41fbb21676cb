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


def test_batch_and_server_struct_layout(tmp_path):
    """The batch and server structs as a C compiler lays them out from the
    header equal the ctypes mirrors."""
    import subprocess
    fields = [("sizeof(pmp_keyframe)", ctypes.sizeof(_lib.Keyframe)),
              ("offsetof(pmp_keyframe, n_views)", _lib.Keyframe.n_views.offset),
              ("offsetof(pmp_keyframe, view_ids)", _lib.Keyframe.view_ids.offset),
              ("offsetof(pmp_keyframe, view_poses)", _lib.Keyframe.view_poses.offset),
              ("offsetof(pmp_keyframe, mask_id)", _lib.Keyframe.mask_id.offset),
              ("offsetof(pmp_keyframe, max_depth)", _lib.Keyframe.max_depth.offset),
              ("sizeof(pmp_batch_config)", ctypes.sizeof(_lib.BatchConfig)),
              ("sizeof(pmp_server_opts)", ctypes.sizeof(_lib.ServerOpts)),
              ("offsetof(pmp_server_opts, config)", _lib.ServerOpts.config.offset),
              ("offsetof(pmp_server_opts, tail_stages)", _lib.ServerOpts.tail_stages.offset),
              ("sizeof(pmp_request)", ctypes.sizeof(_lib.Request)),
              ("offsetof(pmp_request, ref_pixels)", _lib.Request.ref_pixels.offset),
              ("offsetof(pmp_request, mask_pixels)", _lib.Request.mask_pixels.offset),
              ("offsetof(pmp_request, landmark_xyz)", _lib.Request.landmark_xyz.offset),
              ("offsetof(pmp_request, prev_plane4)", _lib.Request.prev_plane4.offset),
              ("offsetof(pmp_request, prev_pose)", _lib.Request.prev_pose.offset),
              ("sizeof(pmp_request_info)", ctypes.sizeof(_lib.RequestInfo)),
              ("sizeof(pmp_server_info)", ctypes.sizeof(_lib.ServerInfo)),
              ("offsetof(pmp_server_info, in_flight)", _lib.ServerInfo.in_flight.offset)]
    src = tmp_path / "lay.c"
    body = "".join(f'printf("%zu\\n", (size_t)({e}));' for e, _ in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "pmp/pmp.h"\n'
                   f'int main(void){{{body}return 0;}}\n')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-std=c11", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [w for _, w in fields], list(zip([e for e, _ in fields], got, [w for _, w in fields]))


def test_roctx_toggle():
    """pmp_set_roctx needs no device: the ROCm marker library of this image
    opens, and tracing turns off again."""
    from patchmatch_amd import _lib
    assert _lib.set_roctx(True)
    assert _lib.set_roctx(False)
