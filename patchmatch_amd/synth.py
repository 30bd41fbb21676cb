"""Seeded synthetic equirectangular scenes with ground-truth depth.

Scenes are ray cast in double precision: a box room (8 x 6 x 3 m) whose
surfaces carry 4-octave value noise, optionally 20 random spheres and an open
cylindrical tube segment of 1.2 m diameter.  Cameras are equirectangular with
the reference's bearing convention (``PatchMatchCuda.cu:191-198``):
b = (cos(lat) sin(lon), -sin(lat), cos(lat) cos(lon)), lon = 2 pi (j/W - 1/2),
lat = -pi (i/H - 1/2); so camera y points down and z forward.  Poses are
world->camera (R, t) as ``data/keyframe.cc:235-243`` returns them.

Depth is the distance along the unit bearing, which is what
1/|b.p| measures for a plane p in the densifier.
"""
from __future__ import annotations

import dataclasses
import numpy as np

ROOM_LO = np.array([-4.0, -1.5, -3.0])
ROOM_HI = np.array([4.0, 1.5, 3.0])

# Lattice frequency (cells per metre) of the texture's coarsest octave.  The
# finest octave has 8 x as many cells, and each config's scene sizes them to
# about 3 pixels at 3 m in its working panorama (one pixel there spans
# 3 m x 2 pi / W): 1 cell per metre for the 480-pixel box room (config 0),
# 4 for the 1920-pixel scene of configs 1-4.  Coarser texture starves the
# 11 x 11 patch (at 1 cell per metre and 1920 pixels the finest cells span
# 13 pixels, and the reference algorithm puts 15 % of the pixels within 1 %
# of the true depth); finer texture aliases.
TEX_FREQ_C0 = 1.0
TEX_FREQ_C1 = 4.0


def _split_seeds(seed: int, n: int) -> list[int]:
    """SplitMix64 stream of n 64-bit seeds from one master seed."""
    out = []
    x = seed & 0xFFFFFFFFFFFFFFFF
    for _ in range(n):
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        out.append(z ^ (z >> 31))
    return out


class ValueNoise3:
    """3-D value noise: random lattice values, quintic-faded trilinear blend."""

    def __init__(self, seed: int, size: int = 256):
        rng = np.random.default_rng(seed)
        self.perm = rng.permutation(size).astype(np.int64)
        self.perm = np.concatenate([self.perm, self.perm])
        self.vals = rng.uniform(-1.0, 1.0, size)
        self.mask = size - 1

    def _hash(self, ix, iy, iz):
        p = self.perm
        m = self.mask
        return p[(p[(p[ix & m] + iy) & m] + iz) & m]

    def __call__(self, x, y, z):
        ix, iy, iz = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64), np.floor(z).astype(np.int64)
        fx, fy, fz = x - ix, y - iy, z - iz
        fade = lambda t: t * t * t * (t * (t * 6.0 - 15.0) + 10.0)
        ux, uy, uz = fade(fx), fade(fy), fade(fz)
        v = self.vals
        out = 0.0
        for dx in (0, 1):
            wx = ux if dx else 1.0 - ux
            for dy in (0, 1):
                wy = uy if dy else 1.0 - uy
                for dz in (0, 1):
                    wz = uz if dz else 1.0 - uz
                    out = out + wx * wy * wz * v[self._hash(ix + dx, iy + dy, iz + dz)]
        return out


@dataclasses.dataclass
class Scene:
    spheres: np.ndarray          # (S, 4): cx, cy, cz, r
    tube: tuple | None           # (y0, z0, radius, x0, x1) axis along x
    noise: ValueNoise3
    base_freq: float = 1.0

    def texture(self, P, obj):
        """4-octave value noise at world points P (n, 3); obj shifts the lattice."""
        off = (obj.astype(np.float64) * 17.31)[:, None]
        Q = P * self.base_freq + off
        val = np.zeros(P.shape[0])
        amp, f = 1.0, 1.0
        for _ in range(4):
            val += amp * self.noise(Q[:, 0] * f, Q[:, 1] * f, Q[:, 2] * f)
            amp *= 0.5
            f *= 2.0
        return np.clip(128.0 + 110.0 * val, 0.0, 255.0)

    def cast(self, C, D):
        """Nearest hit of rays C + t D (D unit, (n,3)); returns t and object ids."""
        n = D.shape[0]
        with np.errstate(divide="ignore", invalid="ignore"):
            tb = np.where(D > 0, (ROOM_HI - C) / D, (ROOM_LO - C) / D)
        tb = np.where(np.isfinite(tb) & (tb > 0), tb, np.inf)
        face = np.argmin(tb, axis=1)
        t = tb[np.arange(n), face]
        obj = face * 2 + (D[np.arange(n), face] > 0)          # 0..5 walls
        for s, (cx, cy, cz, r) in enumerate(self.spheres):
            oc = C - np.array([cx, cy, cz])
            # elementwise products, not BLAS: render() runs on several
            # threads at once, and the BLAS of this image is not safe for
            # concurrent calls (its results then vary from run to run)
            if oc.ndim == 1:
                b = D[:, 0] * oc[0] + D[:, 1] * oc[1] + D[:, 2] * oc[2]
                c = float(oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2]) - r * r
            else:
                b = np.einsum("ij,ij->i", D, oc)
                c = np.einsum("ij,ij->i", oc, oc) - r * r
            disc = b * b - c
            ok = disc > 0
            sq = np.sqrt(np.where(ok, disc, 0.0))
            t1 = -b - sq
            t2 = -b + sq
            ts = np.where(t1 > 1e-9, t1, np.where(t2 > 1e-9, t2, np.inf))
            ts = np.where(ok, ts, np.inf)
            hit = ts < t
            t = np.where(hit, ts, t)
            obj = np.where(hit, 6 + s, obj)
        if self.tube is not None:
            y0, z0, R, x0, x1 = self.tube
            oy = C[..., 1] - y0
            oz = C[..., 2] - z0
            a = D[:, 1] ** 2 + D[:, 2] ** 2
            b = D[:, 1] * oy + D[:, 2] * oz
            c = oy * oy + oz * oz - R * R
            disc = b * b - a * c
            ok = (disc > 0) & (a > 1e-12)
            sq = np.sqrt(np.where(ok, disc, 0.0))
            aa = np.where(a > 1e-12, a, 1.0)
            best = np.full(n, np.inf)
            for root in ((-b - sq) / aa, (-b + sq) / aa):
                xh = C[..., 0] + root * D[:, 0]
                valid = ok & (root > 1e-9) & (xh >= x0) & (xh <= x1)
                best = np.where(valid & (root < best), root, best)
            hit = best < t
            t = np.where(hit, best, t)
            obj = np.where(hit, 6 + len(self.spheres), obj)
        return t, obj


def bearings(W: int, H: int) -> np.ndarray:
    """Unit bearings (H, W, 3) in the camera frame, double precision."""
    j = np.arange(W, dtype=np.float64)
    i = np.arange(H, dtype=np.float64)
    lon = 2.0 * np.pi * (j / W - 0.5)
    lat = -np.pi * (i / H - 0.5)
    cl, sl = np.cos(lat)[:, None], np.sin(lat)[:, None]
    return np.stack([cl * np.sin(lon)[None, :], np.broadcast_to(-sl, (H, W)), cl * np.cos(lon)[None, :]], -1)


def yaw_pose(yaw: float, centre) -> np.ndarray:
    """12-vector world->camera pose (R row-major, t) of a camera at centre with yaw about y."""
    c, s = np.cos(yaw), np.sin(yaw)
    R_wc = np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])
    R_cw = R_wc.T
    t = -R_cw @ np.asarray(centre, dtype=np.float64)
    return np.concatenate([R_cw.ravel(), t])


def render(scene: Scene, pose: np.ndarray, W: int, H: int):
    """Noise-free intensity (float, H x W) and depth (H x W) of one camera."""
    R = pose[:9].reshape(3, 3)
    t = pose[9:]
    # elementwise, not BLAS (see Scene.cast): C = -R^T t, rows of D = R^T b
    C = -(R[0] * t[0] + R[1] * t[1] + R[2] * t[2])
    B = bearings(W, H).reshape(-1, 3)
    D = B[:, :1] * R[0] + B[:, 1:2] * R[1] + B[:, 2:3] * R[2]
    depth, obj = scene.cast(C, D)
    P = C + depth[:, None] * D
    img = scene.texture(P, obj)
    return img.reshape(H, W), depth.reshape(H, W)


def render_planes(scene: Scene, pose: np.ndarray, W: int, H: int):
    """Ground-truth plane map (H, W, 4) float32 of one camera: p = n / (n . X)
    in the camera frame (X the surface point, n its normal), so p . (d b) = 1
    as in the densifier's plane maps; the fourth component is 0."""
    R = pose[:9].reshape(3, 3)
    t = pose[9:]
    C = -R.T @ t
    B = bearings(W, H).reshape(-1, 3)
    D = B @ R
    depth, obj = scene.cast(C, D)
    P = C + depth[:, None] * D
    n = np.zeros_like(P)
    wall = obj < 6
    face = np.where(wall, obj // 2, 0)
    n[wall, face[wall]] = 1.0
    ns = len(scene.spheres)
    for s_, (cx, cy, cz, r) in enumerate(scene.spheres):
        m = obj == 6 + s_
        n[m] = (P[m] - np.array([cx, cy, cz])) / r
    if scene.tube is not None:
        y0, z0 = scene.tube[0], scene.tube[1]
        m = obj == 6 + ns
        n[m, 1] = P[m, 1] - y0
        n[m, 2] = P[m, 2] - z0
    nc = n @ R.T                                   # rows: R n
    Xc = depth[:, None] * B
    out = np.zeros((H * W, 4))
    out[:, :3] = nc / np.einsum("ij,ij->i", nc, Xc)[:, None]
    return out.reshape(H, W, 4).astype(np.float32)


def tilt_planes(pl: np.ndarray, rng: np.random.Generator, frac: float, angle: float) -> np.ndarray:
    """Plane map with a random fraction of pixels tilted by `angle` radians:
    each chosen plane vector is rotated about a random axis perpendicular to
    it, keeping its length (the plane through the same depth tilts)."""
    out = pl.copy()
    m = rng.random(pl.shape[:2]) < frac
    v = pl[m, :3].astype(np.float64)
    r = rng.normal(size=v.shape)
    u = v / np.linalg.norm(v, axis=1, keepdims=True)
    w = r - np.sum(r * u, 1, keepdims=True) * u
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    nv = np.linalg.norm(v, axis=1, keepdims=True)
    out[m, :3] = (nv * (np.cos(angle) * u + np.sin(angle) * w)).astype(np.float32)
    return out


_synth_lib = None


def render_gpu(scene: Scene, pose: np.ndarray, W: int, H: int):
    """render() on the GPU (libpmpsynth.so, built from csrc/synth_render.hip):
    the same cast and texture in double precision; float32 intensity and depth.
    For benchmark inputs at sizes the numpy path cannot render in seconds."""
    import ctypes
    import os
    global _synth_lib
    if _synth_lib is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpmpsynth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        dp, ip, fp = (ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32),
                      ctypes.POINTER(ctypes.c_float))
        L.pmps_render.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, ip, dp, ctypes.c_int,
                                  ctypes.c_double, fp, fp]
        L.pmps_render.restype = ctypes.c_int
        _synth_lib = L
    dp = ctypes.POINTER(ctypes.c_double)
    pose = np.ascontiguousarray(pose, dtype=np.float64)
    sph = np.ascontiguousarray(scene.spheres, dtype=np.float64).reshape(-1, 4)
    tube = None if scene.tube is None else np.ascontiguousarray(scene.tube, dtype=np.float64)
    perm = np.ascontiguousarray(scene.noise.perm, dtype=np.int32)
    vals = np.ascontiguousarray(scene.noise.vals, dtype=np.float64)
    img = np.zeros((H, W), np.float32)
    dep = np.zeros((H, W), np.float32)
    rc = _synth_lib.pmps_render(pose.ctypes.data_as(dp), W, H, sph.ctypes.data_as(dp), sph.shape[0],
                                None if tube is None else tube.ctypes.data_as(dp),
                                perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), vals.ctypes.data_as(dp),
                                int(scene.noise.mask), float(scene.base_freq),
                                img.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                dep.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if rc != 0:
        raise RuntimeError(f"pmps_render failed with HIP error {rc}")
    return img, dep


def photometric(img: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """gain U(0.8,1.2), bias U(-10,10), N(0, 2^2) noise, clamp, round -> uint8."""
    gain = rng.uniform(0.8, 1.2)
    bias = rng.uniform(-10.0, 10.0)
    out = img * gain + bias + rng.normal(0.0, 2.0, img.shape)
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def make_scene(seed: int, spheres: int = 0, tube: bool = False, cam_region: float = 0.8,
               base_freq: float = TEX_FREQ_C0) -> Scene:
    s_scene, s_tex = _split_seeds(seed, 2)
    rng = np.random.default_rng(s_scene)
    sph = []
    while len(sph) < spheres:
        r = rng.uniform(0.15, 0.5)
        c = rng.uniform(ROOM_LO + r + 0.05, ROOM_HI - r - 0.05)
        if np.hypot(c[0], c[2]) < r + cam_region + 0.6:
            continue
        if tube and abs(c[2] - 2.0) < r + 0.7 and -1.6 < c[0] < 0.6:
            continue
        sph.append([c[0], c[1], c[2], r])
    tb = (0.5, 2.0, 0.6, -1.5, 0.5) if tube else None
    return Scene(np.array(sph, dtype=np.float64).reshape(-1, 4), tb, ValueNoise3(s_tex), base_freq)


@dataclasses.dataclass
class Keyframe:
    """One reference keyframe and its neighbourhood, ready for the densifier."""
    W: int
    H: int
    ref: np.ndarray               # (H, W) uint8
    views: list                   # N x (H, W) uint8
    ref_pose: np.ndarray          # (12,)
    view_poses: np.ndarray        # (N, 12)
    mask: np.ndarray              # (H, W) uint8, 0 = ignored
    gt_depth: np.ndarray          # (H, W) float64, reference camera
    bgr: np.ndarray               # (H, W, 3) uint8 colour thumbnail (grey replicated)


def make_keyframe(seed: int, W: int, H: int, n_views: int, baseline=(0.1, 0.3),
                  spheres: int = 0, tube: bool = False, mask_frac: float = 0.15, workers: int = 8,
                  base_freq: float = TEX_FREQ_C0) -> Keyframe:
    """Reference camera near the room centre plus n_views neighbours at random
    baselines/yaws (seeded)."""
    s_scene, s_pose, s_photo = _split_seeds(seed ^ 0x5EED, 3)
    scene = make_scene(s_scene, spheres=spheres, tube=tube, base_freq=base_freq)
    rng = np.random.default_rng(s_pose)
    c0 = np.array([rng.uniform(-0.2, 0.2), rng.uniform(-0.05, 0.05), rng.uniform(-0.2, 0.2)])
    ref_pose = yaw_pose(rng.uniform(-np.pi, np.pi), c0)
    poses = []
    for _ in range(n_views):
        az = rng.uniform(-np.pi, np.pi)
        el = rng.uniform(-0.2, 0.2)
        L = rng.uniform(*baseline)
        d = np.array([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)])
        poses.append(yaw_pose(rng.uniform(-np.pi, np.pi), c0 + L * d))
    prng = np.random.default_rng(s_photo)
    allposes = [ref_pose] + poses
    if W * H * len(allposes) >= 4_000_000 and workers > 1:
        # noise-free renders on worker threads (numpy releases the GIL in the
        # large array operations; no child process outlives the call);
        # photometric noise afterwards, in order, so results do not depend on
        # the worker count
        import concurrent.futures as cf
        with cf.ThreadPoolExecutor(max_workers=min(workers, len(allposes))) as ex:
            rendered = list(ex.map(render, [scene] * len(allposes), allposes, [W] * len(allposes),
                                   [H] * len(allposes)))
    else:
        rendered = [render(scene, p, W, H) for p in allposes]
    img, gt = rendered[0]
    ref = photometric(img, prng)
    views = [photometric(r[0], prng) for r in rendered[1:]]
    mask = np.ones((H, W), np.uint8)
    mask[int(np.floor((1.0 - mask_frac) * H)):, :] = 0
    bgr = np.repeat(ref[:, :, None], 3, axis=2)
    return Keyframe(W, H, ref, views, ref_pose, np.array(poses), mask, gt, bgr)


def config0(seed: int = 0) -> Keyframe:
    """480x240, reference + 2 neighbours, box room, baselines 0.1-0.3 m."""
    return make_keyframe(seed, 480, 240, 2, baseline=(0.1, 0.3))


def config1(seed: int = 0, W: int = 1920, H: int = 960, n_views: int = 6, workers: int = 8,
            base_freq: float = TEX_FREQ_C1) -> Keyframe:
    """1920x960, reference + 6 neighbours, room + 20 spheres + tube, baselines 0.05-0.5 m."""
    return make_keyframe(seed, W, H, n_views, baseline=(0.05, 0.5), spheres=20, tube=True, workers=workers,
                         base_freq=base_freq)


def sequence_gpu(seed: int, n: int, W: int, H: int, spacing: float = 0.2, mask_frac: float = 0.15,
                 with_gt: bool = True, base_freq: float = TEX_FREQ_C1):
    """n keyframes along the seeded trajectory through the config-1 scene
    (room + 20 spheres + tube, spheres kept clear of the path), rendered at
    W x H on the GPU: dict with poses (n, 12), images (n x (H, W) uint8), gt
    (n x (H, W) float32 depth, or None entries without with_gt) and the mask."""
    s_scene, s_photo = _split_seeds(seed ^ 0x5E0, 2)
    scene = make_scene(s_scene, spheres=20, tube=True, cam_region=2.7, base_freq=base_freq)
    poses = trajectory(seed, n, spacing)
    prng = np.random.default_rng(s_photo)
    images, gts = [], []
    for k in range(n):
        img, dep = render_gpu(scene, poses[k], W, H)
        images.append(photometric(img.astype(np.float64), prng))
        gts.append(dep if with_gt else None)
    mask = np.ones((H, W), np.uint8)
    mask[int(np.floor((1.0 - mask_frac) * H)):, :] = 0
    return {"poses": poses, "images": images, "gt": gts, "mask": mask}


def trajectory(seed: int, n: int, spacing: float = 0.2):
    """n camera poses on a seeded smooth closed spline through the room (yaw follows the path)."""
    rng = np.random.default_rng(_split_seeds(seed ^ 0x7A7, 1)[0])
    k = 6
    ang = np.linspace(0, 2 * np.pi, k, endpoint=False)
    rad = rng.uniform(0.8, 1.6, k)
    ctrl = np.stack([rad * np.cos(ang) * 1.6, rng.uniform(-0.1, 0.1, k), rad * np.sin(ang)], 1)
    # Catmull-Rom, densely sampled, then re-sampled by arc length
    pts = []
    for a in range(k):
        p0, p1, p2, p3 = ctrl[(a - 1) % k], ctrl[a], ctrl[(a + 1) % k], ctrl[(a + 2) % k]
        for s in np.linspace(0, 1, 200, endpoint=False):
            pts.append(0.5 * ((2 * p1) + (-p0 + p2) * s + (2 * p0 - 5 * p1 + 4 * p2 - p3) * s * s
                              + (-p0 + 3 * p1 - 3 * p2 + p3) * s ** 3))
    pts = np.array(pts)
    seg = np.linalg.norm(np.diff(pts, axis=0, append=pts[:1]), axis=1)
    arc = np.concatenate([[0], np.cumsum(seg)])[:-1]
    out = []
    for m in range(n):
        s = (m * spacing) % arc[-1]
        q = int(np.searchsorted(arc, s, side="right") - 1)
        nxt = pts[(q + 1) % len(pts)]
        yaw = np.arctan2(nxt[0] - pts[q][0], nxt[2] - pts[q][2])
        out.append(yaw_pose(yaw, pts[q]))
    return np.array(out)
