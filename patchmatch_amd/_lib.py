"""ctypes binding of libpmp.so (the C-ABI of include/pmp/pmp.h).

There is no fallback: if the HIP library is not built, or no GPU is
visible, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import Params

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmp.so")

PMP_MAPS_WORK, PMP_MAPS_INIT, PMP_MAPS_FINAL = 0, 1, 2
STATUS = {0: "PMP_OK", -1: "PMP_EINVAL", -2: "PMP_ENOMEM", -3: "PMP_EDEVICE", -4: "PMP_ENOTFOUND", -5: "PMP_ESTATE",
          -6: "PMP_EBUSY", -7: "PMP_ECANCELED"}

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p


class Pose(ctypes.Structure):
    _fields_ = [("R", ctypes.c_double * 9), ("t", ctypes.c_double * 3)]

    @classmethod
    def from12(cls, v):
        v = np.asarray(v, dtype=np.float64).reshape(12)
        return cls((ctypes.c_double * 9)(*v[:9]), (ctypes.c_double * 3)(*v[9:]))


EXPORTS = [
    "pmp_create", "pmp_destroy", "pmp_last_error", "pmp_default_params", "pmp_synchronize",
    "pmp_image_put", "pmp_image_drop", "pmp_image_pool_stats", "pmp_set_views",
    "pmp_random_initialization", "pmp_score_plane_depth", "pmp_red_black_pass", "pmp_finalize",
    "pmp_densify", "pmp_set_maps", "pmp_get_maps", "pmp_get_normals", "pmp_eval_counts", "pmp_clean", "pmp_clean_ex",
    "pmp_pruner_add_view", "pmp_pruner_prune", "pmp_pruner_pop_head", "pmp_pruner_reset",
    "pmp_pruner_size", "pmp_pruner_get_map", "pmp_debug_eval_math", "pmp_debug_refine_normals",
    "pmp_set_timing", "pmp_get_timing", "pmp_set_mask", "pmp_debug_copy", "pmp_debug_epochs",
    "pmp_debug_epoch_report", "pmp_set_roctx",
    "pmp_batch_densify", "pmp_batch_get_maps", "pmp_batch_set_config", "pmp_batch_get_config",
    "pmp_batch_autotune", "pmp_server_create", "pmp_server_destroy", "pmp_server_last_error",
    "pmp_server_submit", "pmp_server_wait", "pmp_server_fetch", "pmp_server_stats", "pmp_server_pause",
    "pmp_server_cancel", "pmp_server_reset",
    "pmp_preinit_landmarks", "pmp_memorize_result", "pmp_preinit_warp", "pmp_forget_result",
    "pmp_write_dense_msgpack", "pmp_write_ply", "pmp_image_put_thumbnail", "pmp_resize", "pmp_image_get",
    "pmp_score_view", "pmp_depth_interval", "pmp_image_size", "pmp_clean_noalloc",
]

PMP_MAX_VIEWS = 10
PMP_MAX_BATCH = 128


class Keyframe(ctypes.Structure):
    """pmp_keyframe: one keyframe of a batch (pooled image ids + poses)."""
    _fields_ = [("ref_id", ctypes.c_uint64), ("ref_pose", Pose), ("n_views", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("view_ids", ctypes.c_uint64 * PMP_MAX_VIEWS),
                ("view_poses", Pose * PMP_MAX_VIEWS), ("mask_id", ctypes.c_int64),
                ("min_depth", ctypes.c_float), ("max_depth", ctypes.c_float)]

    @classmethod
    def make(cls, ref_id, ref_pose, view_ids, view_poses, mask_id=-1, min_depth=0.0, max_depth=0.0):
        k = cls()
        k.ref_id = ref_id
        k.ref_pose = Pose.from12(ref_pose)
        k.n_views = len(view_ids)
        for v, (vid, vp) in enumerate(zip(view_ids, view_poses)):
            k.view_ids[v] = vid
            k.view_poses[v] = Pose.from12(vp)
        k.mask_id = mask_id
        k.min_depth = min_depth
        k.max_depth = max_depth
        return k


class BatchConfig(ctypes.Structure):
    _fields_ = [("tile_rows", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("xcd_bands", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Timing(ctypes.Structure):
    _fields_ = [("init_ms", ctypes.c_double), ("propagate_ms", ctypes.c_double), ("finalize_ms", ctypes.c_double),
                ("propagate_launches", ctypes.c_int32), ("densify_calls", ctypes.c_int32)]

class ServerOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("max_slots", ctypes.c_int32), ("hbm_budget", ctypes.c_int64),
                ("edge_mode", ctypes.c_int32), ("inflight_kpx", ctypes.c_int32), ("config", BatchConfig),
                ("tail_stages", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("kf", Keyframe), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("ref_pixels", _u8p), ("view_pixels", _u8p * PMP_MAX_VIEWS), ("mask_pixels", _u8p),
                ("n_landmarks", ctypes.c_int64), ("landmark_xyz", _f64p),
                ("prev_plane4", _f32p), ("prev_depth", _f32p), ("prev_pose", ctypes.POINTER(Pose))]


class RequestInfo(ctypes.Structure):
    _fields_ = [("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("steps", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("n_views", ctypes.c_int32)]


class ServerInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("submitted", "completed", "failed", "steps", "pool_images",
                                              "pool_bytes", "maps_bytes", "uploads", "evictions")] + \
               [(n, ctypes.c_int32) for n in ("in_flight", "queued", "max_in_flight", "cancelled")]


class CameraInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("cols", ctypes.c_int32), ("rows", ctypes.c_int32),
                ("fps", ctypes.c_double)]


class DenseKeyframe(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("src_frm_id", ctypes.c_uint64), ("timestamp", ctypes.c_double),
                ("pose", Pose), ("n_points", ctypes.c_int64), ("points", _f64p), ("colors", _u8p),
                ("keypts", _i32p)]


_lib = None


class PmpError(RuntimeError):
    pass


def load():
    """Load libpmp.so and declare its signatures (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PmpError(f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    c_int, i32, i64, u64, f32 = ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float
    sig = {
        "pmp_create": ([P(_vp), i32], c_int),
        "pmp_destroy": ([_vp], None),
        "pmp_last_error": ([_vp], ctypes.c_char_p),
        "pmp_default_params": ([P(Params)], None),
        "pmp_synchronize": ([_vp], c_int),
        "pmp_image_put": ([_vp, u64, _u8p, i32, i32, i64, i32], c_int),
        "pmp_image_drop": ([_vp, u64], c_int),
        "pmp_image_pool_stats": ([_vp, _i64p, _i64p], c_int),
        "pmp_image_size": ([_vp, u64, _i32p, _i32p], c_int),
        "pmp_set_views": ([_vp, u64, P(Pose), i32, _u64p, P(Pose)], c_int),
        "pmp_random_initialization": ([_vp, P(Params)], c_int),
        "pmp_score_plane_depth": ([_vp, P(Params)], c_int),
        "pmp_red_black_pass": ([_vp, P(Params), i32, i32], c_int),
        "pmp_finalize": ([_vp, P(Params), _u8p], c_int),
        "pmp_densify": ([_vp, P(Params), _u8p], c_int),
        "pmp_set_maps": ([_vp, _f32p, _f32p, _f32p], c_int),
        "pmp_get_maps": ([_vp, i32, _f32p, _f32p, _f32p, _i32p], c_int),
        "pmp_get_normals": ([_vp, i32, _f32p], c_int),
        "pmp_eval_counts": ([_vp, _i64p, _i64p], c_int),
        "pmp_clean": ([_vp, i32, i32, i32, P(_f32p), P(Pose), f32, i32, _f32p], c_int),
        "pmp_clean_ex": ([_vp, i32, i32, i32, P(_f32p), P(_f32p), P(Pose), f32, i32, f32, f32, _f32p], c_int),
        "pmp_clean_noalloc": ([_vp, i32, i32, i32, P(_f32p), P(Pose), f32, i32, _f32p], c_int),
        "pmp_pruner_add_view": ([_vp, _f32p, P(Pose), _u8p, i32, i32], c_int),
        "pmp_pruner_prune": ([_vp, f32, _f32p, i64, _f64p, _u8p, _i32p, _i64p], c_int),
        "pmp_pruner_pop_head": ([_vp], c_int),
        "pmp_pruner_reset": ([_vp], c_int),
        "pmp_pruner_size": ([_vp], c_int),
        "pmp_pruner_get_map": ([_vp, i32, _f32p], c_int),
        "pmp_debug_eval_math": ([_vp, i32, i64, _f32p, _f32p, _f32p], c_int),
        "pmp_debug_refine_normals": ([_vp, P(Params), i32, i32, i32, i32, i32, _f32p], c_int),
        "pmp_set_timing": ([_vp, i32], c_int),
        "pmp_get_timing": ([_vp, P(Timing)], c_int),
        "pmp_set_mask": ([_vp, i64], c_int),
        "pmp_debug_copy": ([_vp, i64, i32], c_int),
        "pmp_debug_epochs": ([_vp, i32], c_int),
        "pmp_set_roctx": ([i32], c_int),
        "pmp_debug_epoch_report": ([_vp, _i32p, _i32p, _i32p, _i32p], c_int),
        "pmp_batch_densify": ([_vp, P(Params), i32, P(Keyframe)], c_int),
        "pmp_batch_get_maps": ([_vp, i32, _f32p, _f32p, _f32p, _i32p], c_int),
        "pmp_batch_set_config": ([_vp, P(BatchConfig)], c_int),
        "pmp_batch_get_config": ([_vp, P(BatchConfig)], c_int),
        "pmp_batch_autotune": ([_vp, P(Params), i32, P(Keyframe), P(BatchConfig)], c_int),
        "pmp_server_create": ([P(_vp), P(ServerOpts)], c_int),
        "pmp_server_destroy": ([_vp], None),
        "pmp_server_last_error": ([_vp], ctypes.c_char_p),
        "pmp_server_submit": ([_vp, P(Request), P(Params), _u64p], c_int),
        "pmp_server_wait": ([_vp, u64, i32], c_int),
        "pmp_server_fetch": ([_vp, u64, _f32p, _f32p, _f32p, _i32p, P(RequestInfo)], c_int),
        "pmp_server_stats": ([_vp, P(ServerInfo)], c_int),
        "pmp_server_pause": ([_vp, i32], c_int),
        "pmp_server_cancel": ([_vp, u64], c_int),
        "pmp_server_reset": ([_vp], c_int),
        "pmp_preinit_landmarks": ([_vp, P(Params), i64, _f64p, _u8p, _i64p], c_int),
        "pmp_memorize_result": ([_vp], c_int),
        "pmp_forget_result": ([_vp], c_int),
        "pmp_preinit_warp": ([_vp, P(Params), _u8p, _i64p], c_int),
        "pmp_write_dense_msgpack": ([ctypes.c_char_p, P(CameraInfo), i32, P(DenseKeyframe)], c_int),
        "pmp_write_ply": ([ctypes.c_char_p, i64, _f64p, _u8p], c_int),
        "pmp_image_put_thumbnail": ([_vp, u64, _u8p, i32, i32, i64, i32, i32, i32, i32, i32], c_int),
        "pmp_resize": ([_vp, _u8p, i32, i32, i64, i32, _u8p, i32, i32, i32], c_int),
        "pmp_image_get": ([_vp, u64, _u8p], c_int),
        "pmp_score_view": ([_f64p, i64, _u64p, _f64p, _f64p, i64, _u64p, _f64p], ctypes.c_double),
        "pmp_depth_interval": ([_f64p, i64, _f64p, _f64p, _f64p], c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def score_view(prev_pose, prev_ids, prev_xyz, cur_pose, cur_ids, cur_xyz) -> float:
    """Parallax score of the densify thread's keyframe gate (host code)."""
    L = load()
    a = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (prev_pose, prev_xyz, cur_pose, cur_xyz)]
    pi = np.ascontiguousarray(prev_ids, dtype=np.uint64)
    ci = np.ascontiguousarray(cur_ids, dtype=np.uint64)
    return L.pmp_score_view(_p(a[0], _f64p), len(pi), _p(pi, _u64p), _p(a[1], _f64p), _p(a[2], _f64p), len(ci),
                            _p(ci, _u64p), _p(a[3], _f64p))


def depth_interval(pose, xyz):
    """(min_depth, max_depth) of a keyframe from its landmarks (host code)."""
    L = load()
    pose = np.ascontiguousarray(pose, dtype=np.float64).reshape(12)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    lo, hi = ctypes.c_double(), ctypes.c_double()
    rc = L.pmp_depth_interval(_p(pose, _f64p), len(xyz), _p(xyz, _f64p), ctypes.byref(lo), ctypes.byref(hi))
    if rc != 0:
        raise PmpError(f"pmp_depth_interval: {STATUS.get(rc, rc)}")
    return lo.value, hi.value


def write_ply(path: str, xyz, rgb):
    """ASCII PLY of n points (world xyz doubles, R G B bytes); no GPU needed."""
    L = load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8).reshape(-1, 3)
    rc = L.pmp_write_ply(path.encode(), len(xyz), _p(xyz, _f64p), _p(rgb, _u8p))
    if rc != 0:
        raise PmpError(f"pmp_write_ply: {STATUS.get(rc, rc)}")


def write_dense_msgpack(path: str, camera: str, cols: int, rows: int, keyframes, fps: float = 30.0):
    """Dense map as MessagePack.  keyframes: dicts with id, pose (12), points (n,3),
    colors_bgr (n,3), keypts (n,2 row col), optional src_frm_id and timestamp."""
    L = load()
    keep = []
    arr = (DenseKeyframe * max(len(keyframes), 1))()
    for k, kf in enumerate(keyframes):
        pts = np.ascontiguousarray(kf["points"], dtype=np.float64).reshape(-1, 3)
        col = np.ascontiguousarray(kf["colors_bgr"], dtype=np.uint8).reshape(-1, 3)
        kps = np.ascontiguousarray(kf["keypts"], dtype=np.int32).reshape(-1, 2)
        keep += [pts, col, kps]
        arr[k] = DenseKeyframe(kf["id"], kf.get("src_frm_id", kf["id"]), kf.get("timestamp", 0.0),
                               Pose.from12(kf["pose"]), len(pts), _p(pts, _f64p), _p(col, _u8p), _p(kps, _i32p))
    cam = CameraInfo(camera.encode(), cols, rows, fps)
    rc = L.pmp_write_dense_msgpack(path.encode(), ctypes.byref(cam), len(keyframes), arr)
    if rc != 0:
        raise PmpError(f"pmp_write_dense_msgpack: {STATUS.get(rc, rc)}")



def set_roctx(on: bool) -> bool:
    """roctx ranges around the library's main calls (process-wide); False if
    the ROCm marker library could not be opened."""
    return load().pmp_set_roctx(1 if on else 0) == 0

class Context:
    """One pmp_ctx: device memory, a HIP stream and an image pool."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = _vp()
        rc = self.L.pmp_create(ctypes.byref(h), device)
        if rc != 0:
            raise PmpError(f"pmp_create(device={device}) failed: {STATUS.get(rc, rc)} (is a GPU visible?)")
        self.h = h
        self.W = self.H = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.pmp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.pmp_last_error(self.h).decode(errors="replace")
            raise PmpError(f"{what}: {STATUS.get(rc, rc)}: {msg}")

    # -- images / views
    def image_put(self, image_id: int, img):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        H, W = img.shape
        self._chk(self.L.pmp_image_put(self.h, image_id, _p(img, _u8p), W, H, W, 0), "pmp_image_put")

    def image_put_device(self, image_id: int, ptr: int, W: int, H: int, pitch: int):
        self._chk(self.L.pmp_image_put(self.h, image_id, ctypes.cast(ptr, _u8p), W, H, pitch, 1), "pmp_image_put")

    def image_put_thumbnail(self, image_id: int, img, out_shape, color_order=2):
        """Grey thumbnail (cvtColor + INTER_LINEAR resize) of an (H, W[, C]) uint8 image, pooled under image_id."""
        img = np.ascontiguousarray(img, dtype=np.uint8)
        H, W = img.shape[:2]
        C = 1 if img.ndim == 2 else img.shape[2]
        oH, oW = out_shape
        self._chk(self.L.pmp_image_put_thumbnail(self.h, image_id, _p(img, _u8p), W, H, W * C, C, color_order, oW,
                                                 oH, 0), "pmp_image_put_thumbnail")

    def resize(self, img, out_shape):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        H, W = img.shape[:2]
        C = 1 if img.ndim == 2 else img.shape[2]
        oH, oW = out_shape
        out = np.zeros((oH, oW) if img.ndim == 2 else (oH, oW, C), np.uint8)
        self._chk(self.L.pmp_resize(self.h, _p(img, _u8p), W, H, W * C, C, _p(out, _u8p), oW, oH, 0), "pmp_resize")
        return out

    def image_size(self, image_id: int):
        """(height, width) of a pooled image."""
        w, h = ctypes.c_int32(), ctypes.c_int32()
        self._chk(self.L.pmp_image_size(self.h, image_id, ctypes.byref(w), ctypes.byref(h)), "pmp_image_size")
        return h.value, w.value

    def image_get(self, image_id: int, shape=None):
        """Copy a pooled grey image back to the host (shape, if given, must be its (H, W))."""
        hw = self.image_size(image_id)
        if shape is not None and tuple(shape) != hw:
            raise ValueError(f"image {image_id} is {hw[0]}x{hw[1]} (H x W), not {tuple(shape)}")
        out = np.zeros(hw, np.uint8)
        self._chk(self.L.pmp_image_get(self.h, image_id, _p(out, _u8p)), "pmp_image_get")
        return out

    def _mask(self, mask):
        """A host mask as contiguous uint8 of the current keyframe's (H, W), or None."""
        if mask is None:
            return None
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.shape != (self.H, self.W):
            raise ValueError(f"mask shape {m.shape} differs from the keyframe's {(self.H, self.W)}")
        return m

    def image_drop(self, image_id: int):
        self._chk(self.L.pmp_image_drop(self.h, image_id), "pmp_image_drop")

    def pool_stats(self):
        c, b = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self.L.pmp_image_pool_stats(self.h, ctypes.byref(c), ctypes.byref(b)), "pmp_image_pool_stats")
        return c.value, b.value

    def set_views(self, ref_id, ref_pose, view_ids, view_poses):
        n = len(view_ids)
        ids = (ctypes.c_uint64 * n)(*view_ids)
        poses = (Pose * n)(*[Pose.from12(p) for p in view_poses])
        self._chk(self.L.pmp_set_views(self.h, ref_id, ctypes.byref(Pose.from12(ref_pose)), n, ids, poses),
                  "pmp_set_views")
        self.H, self.W = self.image_size(ref_id)

    def load_keyframe(self, ref, views, ref_pose, view_poses, base_id=0):
        """Upload reference + neighbours (ids base_id, base_id+1, ...) and select them."""
        self.image_put(base_id, ref)
        for k, v in enumerate(views):
            self.image_put(base_id + 1 + k, v)
        self.set_views(base_id, ref_pose, [base_id + 1 + k for k in range(len(views))], view_poses)

    # -- stages
    def random_initialization(self, p: Params):
        self._chk(self.L.pmp_random_initialization(self.h, ctypes.byref(p)), "pmp_random_initialization")

    def score_plane_depth(self, p: Params):
        self._chk(self.L.pmp_score_plane_depth(self.h, ctypes.byref(p)), "pmp_score_plane_depth")

    def red_black_pass(self, p: Params, it_begin: int, it_end: int):
        self._chk(self.L.pmp_red_black_pass(self.h, ctypes.byref(p), it_begin, it_end), "pmp_red_black_pass")

    def finalize(self, p: Params, mask=None):
        m = self._mask(mask)
        self._chk(self.L.pmp_finalize(self.h, ctypes.byref(p), _p(m, _u8p)), "pmp_finalize")

    def densify(self, p: Params, mask=None):
        m = self._mask(mask)
        self._chk(self.L.pmp_densify(self.h, ctypes.byref(p), _p(m, _u8p)), "pmp_densify")

    def preinit_landmarks(self, p: Params, xyz, mask=None) -> int:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        m = self._mask(mask)
        acc = ctypes.c_int64()
        self._chk(self.L.pmp_preinit_landmarks(self.h, ctypes.byref(p), len(xyz), _p(xyz, _f64p), _p(m, _u8p),
                                               ctypes.byref(acc)), "pmp_preinit_landmarks")
        return acc.value

    def memorize_result(self):
        self._chk(self.L.pmp_memorize_result(self.h), "pmp_memorize_result")

    def forget_result(self):
        self._chk(self.L.pmp_forget_result(self.h), "pmp_forget_result")

    def preinit_warp(self, p: Params, mask=None) -> int:
        m = self._mask(mask)
        acc = ctypes.c_int64()
        self._chk(self.L.pmp_preinit_warp(self.h, ctypes.byref(p), _p(m, _u8p), ctypes.byref(acc)),
                  "pmp_preinit_warp")
        return acc.value

    def set_mask(self, mask_id: int):
        self._chk(self.L.pmp_set_mask(self.h, mask_id), "pmp_set_mask")

    def set_timing(self, enable: bool):
        self._chk(self.L.pmp_set_timing(self.h, 1 if enable else 0), "pmp_set_timing")

    def get_timing(self) -> dict:
        t = Timing()
        self._chk(self.L.pmp_get_timing(self.h, ctypes.byref(t)), "pmp_get_timing")
        return {f: getattr(t, f) for f, _ in Timing._fields_}

    def synchronize(self):
        self._chk(self.L.pmp_synchronize(self.h), "pmp_synchronize")

    def set_maps(self, plane4=None, depth=None, score=None):
        a = None if plane4 is None else np.ascontiguousarray(plane4, dtype=np.float32)
        b = None if depth is None else np.ascontiguousarray(depth, dtype=np.float32)
        c = None if score is None else np.ascontiguousarray(score, dtype=np.float32)
        self._chk(self.L.pmp_set_maps(self.h, _p(a, _f32p), _p(b, _f32p), _p(c, _f32p)), "pmp_set_maps")

    def get_maps(self, which=PMP_MAPS_WORK):
        H, W = self.H, self.W
        plane = np.zeros((H, W, 4), np.float32)
        depth = np.zeros((H, W), np.float32)
        score = np.zeros((H, W), np.float32)
        nghbr = np.zeros((H, W), np.int32) if which == PMP_MAPS_FINAL else None
        self._chk(self.L.pmp_get_maps(self.h, which, _p(plane, _f32p), _p(depth, _f32p), _p(score, _f32p),
                                      _p(nghbr, _i32p)), "pmp_get_maps")
        return (plane, depth, score, nghbr) if which == PMP_MAPS_FINAL else (plane, depth, score)

    def get_normals(self, which=PMP_MAPS_FINAL):
        """Unit camera-facing normals -p/|p| of a map set, (H, W, 3)."""
        out = np.zeros((self.H, self.W, 3), np.float32)
        self._chk(self.L.pmp_get_normals(self.h, which, _p(out, _f32p)), "pmp_get_normals")
        return out

    def eval_counts(self):
        e, c = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self.L.pmp_eval_counts(self.h, ctypes.byref(e), ctypes.byref(c)), "pmp_eval_counts")
        return e.value, c.value

    # -- many keyframes per launch
    def batch_densify(self, p: Params, kfs):
        arr = (Keyframe * len(kfs))(*kfs)
        self._chk(self.L.pmp_batch_densify(self.h, ctypes.byref(p), len(kfs), arr), "pmp_batch_densify")

    def batch_get_maps(self, k: int, shape):
        H, W = shape
        plane = np.zeros((H, W, 4), np.float32)
        depth = np.zeros((H, W), np.float32)
        score = np.zeros((H, W), np.float32)
        nghbr = np.zeros((H, W), np.int32)
        self._chk(self.L.pmp_batch_get_maps(self.h, k, _p(plane, _f32p), _p(depth, _f32p), _p(score, _f32p),
                                            _p(nghbr, _i32p)), "pmp_batch_get_maps")
        return plane, depth, score, nghbr

    def batch_set_config(self, tile_rows=4, blocks_per_cu=0, use_graph=1, xcd_bands=1):
        c = BatchConfig(tile_rows, blocks_per_cu, use_graph, xcd_bands)
        self._chk(self.L.pmp_batch_set_config(self.h, ctypes.byref(c)), "pmp_batch_set_config")

    def batch_get_config(self) -> dict:
        c = BatchConfig()
        self._chk(self.L.pmp_batch_get_config(self.h, ctypes.byref(c)), "pmp_batch_get_config")
        return c.as_dict()

    def batch_autotune(self, p: Params, kfs) -> dict:
        arr = (Keyframe * len(kfs))(*kfs)
        c = BatchConfig()
        self._chk(self.L.pmp_batch_autotune(self.h, ctypes.byref(p), len(kfs), arr, ctypes.byref(c)),
                  "pmp_batch_autotune")
        return c.as_dict()

    # -- cleaner / pruner
    def clean(self, depths, poses, thr, min_consistent):
        depths = [np.ascontiguousarray(d, dtype=np.float32) for d in depths]
        H, W = depths[0].shape
        arr = (_f32p * len(depths))(*[_p(d, _f32p) for d in depths])
        ps = (Pose * len(depths))(*[Pose.from12(p) for p in poses])
        out = np.zeros((H, W), np.float32)
        self._chk(self.L.pmp_clean(self.h, W, H, len(depths), arr, ps, thr, min_consistent, _p(out, _f32p)),
                  "pmp_clean")
        return out

    def clean_noalloc(self, depths, poses, thr, min_consistent):
        """pmp_clean_noalloc: clean() with no device allocation (a 16 MiB device array per GPU)."""
        depths = [np.ascontiguousarray(d, dtype=np.float32) for d in depths]
        H, W = depths[0].shape
        arr = (_f32p * len(depths))(*[_p(d, _f32p) for d in depths])
        ps = (Pose * len(depths))(*[Pose.from12(p) for p in poses])
        out = np.zeros((H, W), np.float32)
        self._chk(self.L.pmp_clean_noalloc(self.h, W, H, len(depths), arr, ps, thr, min_consistent, _p(out, _f32p)),
                  "pmp_clean_noalloc")
        return out

    def clean_ex(self, depths, planes, poses, thr, min_consistent, max_reproj_px=0.0, max_normal_deg=0.0):
        """pmp_clean_ex: clean() with the optional round-trip (pixels) and
        normal-angle (degrees) tests; planes (H, W, 4) are needed for the latter."""
        depths = [np.ascontiguousarray(d, dtype=np.float32) for d in depths]
        H, W = depths[0].shape
        arr = (_f32p * len(depths))(*[_p(d, _f32p) for d in depths])
        parr = None
        if planes is not None:
            planes = [np.ascontiguousarray(q, dtype=np.float32).reshape(H, W, 4) for q in planes]
            parr = (_f32p * len(planes))(*[_p(q, _f32p) for q in planes])
        ps = (Pose * len(depths))(*[Pose.from12(p) for p in poses])
        out = np.zeros((H, W), np.float32)
        self._chk(self.L.pmp_clean_ex(self.h, W, H, len(depths), arr, parr, ps, thr, min_consistent,
                                      float(max_reproj_px), float(max_normal_deg), _p(out, _f32p)), "pmp_clean_ex")
        return out

    def pruner_add_view(self, depth, pose, bgr):
        d = np.ascontiguousarray(depth, dtype=np.float32)
        b = np.ascontiguousarray(bgr, dtype=np.uint8)
        H, W = d.shape
        self._chk(self.L.pmp_pruner_add_view(self.h, _p(d, _f32p), ctypes.byref(Pose.from12(pose)), _p(b, _u8p), W, H),
                  "pmp_pruner_add_view")

    def pruner_prune(self, thr, shape):
        H, W = shape
        cap = H * W
        pruned = np.zeros((H, W), np.float32)
        pts = np.zeros((cap, 3), np.float64)
        cols = np.zeros((cap, 3), np.uint8)
        kps = np.zeros((cap, 2), np.int32)
        cnt = ctypes.c_int64()
        self._chk(self.L.pmp_pruner_prune(self.h, thr, _p(pruned, _f32p), cap, _p(pts, _f64p), _p(cols, _u8p),
                                          _p(kps, _i32p), ctypes.byref(cnt)), "pmp_pruner_prune")
        n = cnt.value
        return pruned, pts[:n], cols[:n], kps[:n]

    def pruner_pop_head(self):
        self._chk(self.L.pmp_pruner_pop_head(self.h), "pmp_pruner_pop_head")

    def pruner_reset(self):
        self._chk(self.L.pmp_pruner_reset(self.h), "pmp_pruner_reset")

    def pruner_size(self):
        return self.L.pmp_pruner_size(self.h)

    def pruner_get_map(self, index, shape):
        out = np.zeros(shape, np.float32)
        self._chk(self.L.pmp_pruner_get_map(self.h, index, _p(out, _f32p)), "pmp_pruner_get_map")
        return out

    # -- diagnostics
    MATH = {"rcp": 0, "rsqrt": 1, "sqrt": 2, "asin": 3, "acos": 4, "atan2": 5, "exp": 6, "log": 7,
            "sin2pi": 8, "cos2pi": 9, "lat": 10, "lon": 11}

    def eval_math(self, name, a, b=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
        out = np.zeros_like(a)
        self._chk(self.L.pmp_debug_eval_math(self.h, self.MATH[name], a.size, _p(a, _f32p), _p(bb, _f32p),
                                             _p(out, _f32p)), "pmp_debug_eval_math")
        return out

    def debug_epochs(self, mode: int):
        """Pass epochs of the red-black passes: 0 off, 1 on, 2 on with tile 0
        started one pass ahead (one violation forced from the host)."""
        self._chk(self.L.pmp_debug_epochs(self.h, int(mode)), "pmp_debug_epochs")

    def debug_epoch_report(self) -> dict:
        v = [ctypes.c_int32(0) for _ in range(4)]
        self._chk(self.L.pmp_debug_epoch_report(self.h, *[ctypes.byref(x) for x in v]), "pmp_debug_epoch_report")
        return dict(zip(("tiles", "violations", "min_epoch", "max_epoch"), (x.value for x in v)))

    def debug_copy(self, nbytes: int, repeats: int = 1):
        self._chk(self.L.pmp_debug_copy(self.h, nbytes, repeats), "pmp_debug_copy")

    def refine_normals(self, p: Params, W, i, j, it, n=18):
        out = np.zeros(n, np.float32)
        self._chk(self.L.pmp_debug_refine_normals(self.h, ctypes.byref(p), W, i, j, it, n, _p(out, _f32p)),
                  "pmp_debug_refine_normals")
        return out


class Server:
    """pmp_server: densification requests served with continuous batching."""

    def __init__(self, device=0, max_slots=16, hbm_budget=0, edge_mode=0, tile_rows=0, blocks_per_cu=0,
                 xcd_bands=1, inflight_kpx=0, tail_stages=0):
        """tile_rows 0: the library's default launch shape."""
        self.L = load()
        o = ServerOpts(device, max_slots, hbm_budget, edge_mode, inflight_kpx,
                       BatchConfig(tile_rows, blocks_per_cu, 1, xcd_bands), tail_stages)
        h = _vp()
        rc = self.L.pmp_server_create(ctypes.byref(h), ctypes.byref(o))
        if rc != 0:
            raise PmpError(f"pmp_server_create failed: {STATUS.get(rc, rc)} (is a GPU visible?)")
        self.h = h
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.pmp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.pmp_server_last_error(self.h).decode(errors="replace")
            raise PmpError(f"{what}: {STATUS.get(rc, rc)}: {msg}")

    def submit(self, p: Params, ref_id, ref_pose, view_ids, view_poses, shape, ref=None, views=None, mask_id=-1,
               mask=None, min_depth=0.0, max_depth=0.0, landmarks=None, prev=None) -> int:
        """Queue one keyframe; arrays are needed only for ids the pool does not hold.
        landmarks: optional (n, 3) world positions that seed the maps after init.
        prev: optional warm start (plane (H, W, 4), depth (H, W), pose (12,)): the
        previous keyframe's final maps and pose, warped in after init."""
        H, W = shape
        r = Request()
        r.kf = Keyframe.make(ref_id, ref_pose, view_ids, view_poses, mask_id, min_depth, max_depth)
        r.width, r.height = W, H
        keep = []

        def px(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.uint8)
            if a.shape != (H, W):
                raise ValueError(f"image shape {a.shape} differs from the request's {(H, W)}")
            keep.append(a)
            return a.ctypes.data_as(_u8p)
        r.ref_pixels = px(ref)
        for v in range(len(view_ids)):
            r.view_pixels[v] = px(None if views is None else views[v])
        r.mask_pixels = px(mask)
        if landmarks is not None and len(landmarks):
            lm = np.ascontiguousarray(landmarks, dtype=np.float64).reshape(-1, 3)
            keep.append(lm)
            r.n_landmarks = lm.shape[0]
            r.landmark_xyz = lm.ctypes.data_as(_f64p)
        if prev is not None:
            pl = np.ascontiguousarray(prev[0], dtype=np.float32)
            dp = np.ascontiguousarray(prev[1], dtype=np.float32)
            if pl.shape != (H, W, 4) or dp.shape != (H, W):
                raise ValueError(f"previous maps {pl.shape}, {dp.shape} differ from the request's {(H, W)}")
            pose = Pose.from12(prev[2])
            keep += [pl, dp, pose]
            r.prev_plane4 = pl.ctypes.data_as(_f32p)
            r.prev_depth = dp.ctypes.data_as(_f32p)
            r.prev_pose = ctypes.pointer(pose)
        t = ctypes.c_uint64()
        self._chk(self.L.pmp_server_submit(self.h, ctypes.byref(r), ctypes.byref(p), ctypes.byref(t)),
                  "pmp_server_submit")
        self.shapes[t.value] = (H, W)
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> bool:
        rc = self.L.pmp_server_wait(self.h, ticket, timeout_ms)
        if rc == -6:
            return False
        self._chk(rc, "pmp_server_wait")
        return True

    def fetch(self, ticket: int):
        """(plane, depth, score, nghbr, info) of a finished request; forgets the ticket."""
        H, W = self.shapes.pop(ticket)
        plane = np.zeros((H, W, 4), np.float32)
        depth = np.zeros((H, W), np.float32)
        score = np.zeros((H, W), np.float32)
        nghbr = np.zeros((H, W), np.int32)
        info = RequestInfo()
        self._chk(self.L.pmp_server_fetch(self.h, ticket, _p(plane, _f32p), _p(depth, _f32p), _p(score, _f32p),
                                          _p(nghbr, _i32p), ctypes.byref(info)), "pmp_server_fetch")
        return plane, depth, score, nghbr, {f: getattr(info, f) for f, _ in RequestInfo._fields_}

    def pause(self, paused: bool = True):
        self._chk(self.L.pmp_server_pause(self.h, 1 if paused else 0), "pmp_server_pause")

    def cancel(self, ticket: int):
        """End one request (fetch then raises PMP_ECANCELED unless it had finished)."""
        self._chk(self.L.pmp_server_cancel(self.h, ticket), "pmp_server_cancel")

    def reset(self):
        """Cancel every queued and in-flight request and empty the image pool (blocks)."""
        self._chk(self.L.pmp_server_reset(self.h), "pmp_server_reset")

    def stats(self) -> dict:
        s = ServerInfo()
        self._chk(self.L.pmp_server_stats(self.h, ctypes.byref(s)), "pmp_server_stats")
        return {f: getattr(s, f) for f, _ in ServerInfo._fields_}
