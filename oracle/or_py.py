"""ctypes front-end of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module.  It is the checker the HIP
path is compared with, never part of the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from patchmatch_amd.params import Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class OrKeyframe(ctypes.Structure):
    _fields_ = [("W", ctypes.c_int32), ("H", ctypes.c_int32), ("N", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("ref", _u8p), ("views", ctypes.POINTER(_u8p)),
                ("ref_pose", _f64p), ("view_poses", _f64p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.or_init.argtypes = [P(OrKeyframe), P(Params), _f32p, _f32p, _f32p]
        L.or_score_field.argtypes = [P(OrKeyframe), P(Params), _f32p, _f32p, _f32p]
        L.or_score_field_f64.argtypes = [P(OrKeyframe), P(Params), _f32p, _f64p, _f64p]
        L.or_warp_taps.argtypes = [P(OrKeyframe), P(Params), _f32p, ctypes.c_int64, _i32p, _f32p, _f64p]
        L.or_score_at_coords_f64.argtypes = [P(OrKeyframe), P(Params), _f32p, ctypes.c_int64, _i32p, _f32p, _f64p,
                                             _f64p]
        L.or_propagate.argtypes = [P(OrKeyframe), P(Params), _f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_int, _i64p]
        L.or_propagate_rows.argtypes = [P(OrKeyframe), P(Params), _f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _i64p]
        L.or_finalize.argtypes = [P(OrKeyframe), P(Params), _u8p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                  _f32p, _f32p, _f32p, _i32p]
        L.or_densify.argtypes = [P(OrKeyframe), P(Params), _u8p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p, _f32p]
        L.or_clean.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(_f32p), _f64p, ctypes.c_float,
                               ctypes.c_int, _f32p]
        L.or_clean_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(_f32p), P(_f32p), _f64p, ctypes.c_float,
                                  ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32p]
        L.or_prune.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(_f32p), _f64p, _u8p, ctypes.c_float,
                               _f32p, ctypes.c_int64, _f64p, _u8p, _i32p, _i64p]
        L.or_eval_math.argtypes = [ctypes.c_int, ctypes.c_int64, _f32p, _f32p, _f32p]
        L.or_philox4x32_10.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        L.or_refine_normals.argtypes = [P(OrKeyframe), P(Params), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, _f32p]
        L.or_tables.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, _f32p]
        L.or_warp_point.argtypes = [P(OrKeyframe), P(Params), ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p]
        L.or_warp_point.restype = ctypes.c_int
        L.or_aggregate.argtypes = [P(Params), ctypes.c_int, _f32p]
        L.or_aggregate.restype = ctypes.c_float
        L.or_params_layout.argtypes = [_i64p]
        L.or_params_layout.restype = ctypes.c_int
        L.or_derive_poses.argtypes = [_f64p, ctypes.c_int, _f64p, _f32p, _f32p]
        L.or_to_gray.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _u8p]
        L.or_resize_linear.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _u8p,
                                       ctypes.c_int, ctypes.c_int]
        L.or_preinit_landmarks.argtypes = [P(OrKeyframe), P(Params), _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _u8p,
                                           ctypes.c_int64, _f64p, _f32p]
        L.or_preinit_warp.argtypes = [P(OrKeyframe), P(Params), _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _u8p,
                                      _f64p, _f32p, _f32p, _f32p]
        for name in ("or_init", "or_score_field", "or_score_field_f64", "or_propagate", "or_propagate_rows",
                     "or_finalize", "or_densify", "or_clean", "or_clean_ex", "or_prune", "or_eval_math", "or_refine_normals",
                     "or_tables", "or_derive_poses", "or_warp_taps", "or_score_at_coords_f64"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def set_threads(n: int):
    """OpenMP thread count of the oracle (libgomp reads it per parallel region)."""
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(int(n))
    except OSError:
        pass


class Keyframe:
    """Keeps the numpy buffers alive behind an OrKeyframe."""

    def __init__(self, ref, views, ref_pose, view_poses):
        self.ref = np.ascontiguousarray(ref, dtype=np.uint8)
        self.views = [np.ascontiguousarray(v, dtype=np.uint8) for v in views]
        self.ref_pose = np.ascontiguousarray(ref_pose, dtype=np.float64).reshape(12)
        self.view_poses = np.ascontiguousarray(view_poses, dtype=np.float64).reshape(-1, 12)
        self.H, self.W = self.ref.shape
        self.N = len(self.views)
        self._ptrs = (_u8p * self.N)(*[_p(v, _u8p) for v in self.views])
        self.c = OrKeyframe(self.W, self.H, self.N, 0, _p(self.ref, _u8p), self._ptrs,
                            _p(self.ref_pose, _f64p), _p(self.view_poses, _f64p))

    @classmethod
    def from_synth(cls, kf):
        return cls(kf.ref, kf.views, kf.ref_pose, kf.view_poses)


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed ({rc})")


def init(kf: Keyframe, p: Params):
    P = kf.W * kf.H
    plane = np.zeros((kf.H, kf.W, 4), np.float32)
    depth = np.zeros((kf.H, kf.W), np.float32)
    score = np.zeros((kf.H, kf.W), np.float32)
    _chk(lib().or_init(ctypes.byref(kf.c), ctypes.byref(p), _p(plane, _f32p), _p(depth, _f32p), _p(score, _f32p)),
         "init")
    return plane, depth, score


def score_field(kf: Keyframe, p: Params, plane4, views=False):
    plane4 = np.ascontiguousarray(plane4, dtype=np.float32)
    score = np.zeros((kf.H, kf.W), np.float32)
    vs = np.zeros((kf.H, kf.W, kf.N), np.float32) if views else None
    _chk(lib().or_score_field(ctypes.byref(kf.c), ctypes.byref(p), _p(plane4, _f32p), _p(score, _f32p),
                              _p(vs, _f32p) if views else None), "score_field")
    return (score, vs) if views else score


def score_field_f64(kf: Keyframe, p: Params, plane4, views=False):
    plane4 = np.ascontiguousarray(plane4, dtype=np.float32)
    score = np.zeros((kf.H, kf.W), np.float64)
    vs = np.zeros((kf.H, kf.W, kf.N), np.float64) if views else None
    _chk(lib().or_score_field_f64(ctypes.byref(kf.c), ctypes.byref(p), _p(plane4, _f32p), _p(score, _f64p),
                                  _p(vs, _f64p) if views else None), "score_field_f64")
    return (score, vs) if views else score


def _taps(p: Params):
    return len(range(-p.patch_half, p.patch_half + 1, p.patch_stride)) ** 2


def warp_taps(kf: Keyframe, p: Params, plane4, pix, exact=True):
    """FP32 sample coordinates (n, N, T, 2) of every tap and view at interior
    pixels pix (n, 2) = (i, j), and (exact) their FP64 counterparts."""
    plane4 = np.ascontiguousarray(plane4, dtype=np.float32)
    pix = np.ascontiguousarray(pix, dtype=np.int32).reshape(-1, 2)
    n, T = pix.shape[0], _taps(p)
    uv32 = np.zeros((n, kf.N, T, 2), np.float32)
    uv64 = np.zeros((n, kf.N, T, 2), np.float64) if exact else None
    _chk(lib().or_warp_taps(ctypes.byref(kf.c), ctypes.byref(p), _p(plane4, _f32p), n, _p(pix, _i32p),
                            _p(uv32, _f32p), _p(uv64, _f64p)), "warp_taps")
    return uv32, uv64


def score_at_coords_f64(kf: Keyframe, p: Params, plane4, pix, uv32):
    """FP64 score and per-view scores at pixels pix with the samples taken at
    the given coordinates uv32 (warp_taps' layout)."""
    plane4 = np.ascontiguousarray(plane4, dtype=np.float32)
    pix = np.ascontiguousarray(pix, dtype=np.int32).reshape(-1, 2)
    uv32 = np.ascontiguousarray(uv32, dtype=np.float32)
    n = pix.shape[0]
    assert uv32.shape == (n, kf.N, _taps(p), 2)
    s = np.zeros(n, np.float64)
    sv = np.zeros((n, kf.N), np.float64)
    _chk(lib().or_score_at_coords_f64(ctypes.byref(kf.c), ctypes.byref(p), _p(plane4, _f32p), n, _p(pix, _i32p),
                                      _p(uv32, _f32p), _p(s, _f64p), _p(sv, _f64p)), "score_at_coords_f64")
    return s, sv


def propagate(kf: Keyframe, p: Params, plane4, depth, score, it_begin, it_end):
    """In place on copies; returns (plane4, depth, score, n_evals)."""
    plane4 = np.array(plane4, dtype=np.float32, copy=True)
    depth = np.array(depth, dtype=np.float32, copy=True)
    score = np.array(score, dtype=np.float32, copy=True)
    ev = ctypes.c_int64(0)
    _chk(lib().or_propagate(ctypes.byref(kf.c), ctypes.byref(p), _p(plane4, _f32p), _p(depth, _f32p),
                            _p(score, _f32p), it_begin, it_end, ctypes.byref(ev)), "propagate")
    return plane4, depth, score, ev.value


def propagate_rows(kf: Keyframe, p: Params, plane4, depth, score, it, colour, r0, r1):
    ev = ctypes.c_int64(0)
    _chk(lib().or_propagate_rows(ctypes.byref(kf.c), ctypes.byref(p), _p(plane4, _f32p), _p(depth, _f32p),
                                 _p(score, _f32p), it, colour, r0, r1, ctypes.byref(ev)), "propagate_rows")
    return ev.value


def finalize(kf: Keyframe, p: Params, mask, init_maps, fin_maps):
    mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    ip, idp, isc = [np.ascontiguousarray(a, dtype=np.float32) for a in init_maps]
    fp, fdp, fsc = [np.ascontiguousarray(a, dtype=np.float32) for a in fin_maps]
    H, W = kf.H, kf.W
    od = np.zeros((H, W), np.float32)
    op = np.zeros((H, W, 3), np.float32)
    osc = np.zeros((H, W), np.float32)
    on = np.zeros((H, W), np.int32)
    _chk(lib().or_finalize(ctypes.byref(kf.c), ctypes.byref(p), None if mask is None else _p(mask, _u8p),
                           _p(ip, _f32p), _p(idp, _f32p), _p(isc, _f32p), _p(fp, _f32p), _p(fdp, _f32p),
                           _p(fsc, _f32p), _p(od, _f32p), _p(op, _f32p), _p(osc, _f32p), _p(on, _i32p)), "finalize")
    return od, op, osc, on


def densify(kf: Keyframe, p: Params, mask=None):
    """Full estimator; returns dict of outputs and raw maps."""
    H, W = kf.H, kf.W
    mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    out = dict(depth=np.zeros((H, W), np.float32), plane=np.zeros((H, W, 3), np.float32),
               score=np.zeros((H, W), np.float32), nghbr=np.zeros((H, W), np.int32),
               raw_plane=np.zeros((H, W, 4), np.float32), raw_depth=np.zeros((H, W), np.float32),
               raw_score=np.zeros((H, W), np.float32))
    _chk(lib().or_densify(ctypes.byref(kf.c), ctypes.byref(p), None if mask is None else _p(mask, _u8p),
                          _p(out["depth"], _f32p), _p(out["plane"], _f32p), _p(out["score"], _f32p),
                          _p(out["nghbr"], _i32p), _p(out["raw_plane"], _f32p), _p(out["raw_depth"], _f32p),
                          _p(out["raw_score"], _f32p)), "densify")
    return out


def clean(depths, poses, thr, min_consistent):
    depths = [np.ascontiguousarray(d, dtype=np.float32) for d in depths]
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 12)
    H, W = depths[0].shape
    arr = (_f32p * len(depths))(*[_p(d, _f32p) for d in depths])
    out = np.zeros((H, W), np.float32)
    _chk(lib().or_clean(W, H, len(depths), arr, _p(poses, _f64p), thr, min_consistent, _p(out, _f32p)), "clean")
    return out


def clean_ex(depths, planes, poses, thr, min_consistent, max_reproj_px=0.0, max_normal_deg=0.0):
    """or_clean_ex: the Cleaner with the optional round-trip and normal tests
    (thresholds as pmp_clean_ex: pixels and degrees, <= 0 switches off)."""
    depths = [np.ascontiguousarray(d, dtype=np.float32) for d in depths]
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 12)
    H, W = depths[0].shape
    arr = (_f32p * len(depths))(*[_p(d, _f32p) for d in depths])
    if planes is not None:
        planes = [np.ascontiguousarray(q, dtype=np.float32).reshape(H, W, 4) for q in planes]
        parr = (_f32p * len(planes))(*[_p(q, _f32p) for q in planes])
    else:
        parr = (_f32p * len(depths))()
    out = np.zeros((H, W), np.float32)
    _chk(lib().or_clean_ex(W, H, len(depths), arr, parr, _p(poses, _f64p), thr, min_consistent,
                           float(max_reproj_px), float(max_normal_deg), _p(out, _f32p)), "clean_ex")
    return out


def prune(depths, poses, bgr, thr):
    """Sequential Pruner.  depths[1:] are modified in place (as the reference does).
    Returns pruned depth, points (n,3), colours (n,3), keypoints (n,2)."""
    for d in depths:
        assert d.dtype == np.float32 and d.flags.c_contiguous
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 12)
    bgr = np.ascontiguousarray(bgr, dtype=np.uint8)
    H, W = depths[0].shape
    arr = (_f32p * len(depths))(*[_p(d, _f32p) for d in depths])
    pruned = np.zeros((H, W), np.float32)
    cap = H * W
    pts = np.zeros((cap, 3), np.float64)
    cols = np.zeros((cap, 3), np.uint8)
    kps = np.zeros((cap, 2), np.int32)
    cnt = ctypes.c_int64(0)
    _chk(lib().or_prune(W, H, len(depths), arr, _p(poses, _f64p), _p(bgr, _u8p), thr, _p(pruned, _f32p), cap,
                        _p(pts, _f64p), _p(cols, _u8p), _p(kps, _i32p), ctypes.byref(cnt)), "prune")
    n = cnt.value
    return pruned, pts[:n].copy(), cols[:n].copy(), kps[:n].copy()


MATH = {"rcp": 0, "rsqrt": 1, "sqrt": 2, "asin": 3, "acos": 4, "atan2": 5, "exp": 6, "log": 7,
        "sin2pi": 8, "cos2pi": 9, "lat": 10, "lon": 11}


def eval_math(name, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(a if b is None else b, dtype=np.float32)
    out = np.zeros_like(a)
    _chk(lib().or_eval_math(MATH[name], a.size, _p(a, _f32p), _p(b, _f32p), _p(out, _f32p)), "eval_math")
    return out


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def refine_normals(kf: Keyframe, p: Params, i, j, it, n=18):
    z = np.zeros(n, np.float32)
    _chk(lib().or_refine_normals(ctypes.byref(kf.c), ctypes.byref(p), i, j, it, n, _p(z, _f32p)), "refine_normals")
    return z


def warp_point(kf: Keyframe, p: Params, i, j, view, plane3):
    pl = np.ascontiguousarray(plane3, dtype=np.float32)
    uv = np.zeros(2, np.float32)
    _chk(lib().or_warp_point(ctypes.byref(kf.c), ctypes.byref(p), i, j, view, _p(pl, _f32p), _p(uv, _f32p)),
         "warp_point")
    return float(uv[0]), float(uv[1])


def aggregate(p: Params, scores):
    sv = np.ascontiguousarray(scores, dtype=np.float32)
    return float(lib().or_aggregate(ctypes.byref(p), sv.size, _p(sv, _f32p)))


def tables(W, H):
    rt = np.zeros((H, 2), np.float32)
    ct = np.zeros((W, 2), np.float32)
    lib().or_tables(W, H, _p(rt, _f32p), _p(ct, _f32p))
    return rt, ct


def params_layout():
    out = np.zeros(4, np.int64)
    lib().or_params_layout(_p(out, _i64p))
    return [int(x) for x in out]


def _maps6(work, init):
    out = [np.array(a, dtype=np.float32, copy=True) for a in (*work, *init)]
    return out


def preinit_landmarks(kf: Keyframe, p: Params, work, init, mask, xyz):
    """PreInitSparsePCL2 restated.  work/init: (plane4, depth, score) each.
    Returns (work', init', candidate scores, accepted count)."""
    m = _maps6(work, init)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    sc = np.zeros(len(xyz), np.float32)
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    n = lib().or_preinit_landmarks(ctypes.byref(kf.c), ctypes.byref(p), *[_p(a, _f32p) for a in m], _p(mk, _u8p),
                                   len(xyz), _p(xyz, _f64p), _p(sc, _f32p))
    _chk(min(n, 0), "preinit_landmarks")
    return tuple(m[:3]), tuple(m[3:]), sc, n


def preinit_warp(kf: Keyframe, p: Params, work, init, mask, prev_pose, prev_plane4, prev_depth):
    """PreInitResultWarp restated.  Returns (work', init', candidate scores, accepted count)."""
    m = _maps6(work, init)
    pp = np.ascontiguousarray(prev_pose, dtype=np.float64).reshape(12)
    pl = np.ascontiguousarray(prev_plane4, dtype=np.float32)
    pd = np.ascontiguousarray(prev_depth, dtype=np.float32)
    sc = np.zeros(pd.size, np.float32)
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    n = lib().or_preinit_warp(ctypes.byref(kf.c), ctypes.byref(p), *[_p(a, _f32p) for a in m], _p(mk, _u8p),
                              _p(pp, _f64p), _p(pl, _f32p), _p(pd, _f32p), _p(sc, _f32p))
    _chk(min(n, 0), "preinit_warp")
    return tuple(m[:3]), tuple(m[3:]), sc.reshape(pd.shape), n


def to_gray(img, order=2):
    """cvtColor to grey of an (H, W) or (H, W, C) uint8 image; order 0 grey, 1 RGB, 2 BGR."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    H, W = img.shape[:2]
    C = 1 if img.ndim == 2 else img.shape[2]
    out = np.zeros((H, W), np.uint8)
    _chk(lib().or_to_gray(_p(img, _u8p), W, H, W * C, C, order, _p(out, _u8p)), "to_gray")
    return out


def resize_linear(img, dW, dH):
    """cv::resize INTER_LINEAR of an 8-bit image with 1..4 channels."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    H, W = img.shape[:2]
    C = 1 if img.ndim == 2 else img.shape[2]
    out = np.zeros((dH, dW) if C == 1 and img.ndim == 2 else (dH, dW, C), np.uint8)
    _chk(lib().or_resize_linear(_p(img, _u8p), W, H, W * C, C, _p(out, _u8p), dW, dH), "resize_linear")
    return out
