"""GPU (torch) renderer of the seeded synthetic scenes of synth.py, for
benchmark inputs at sizes the numpy renderer cannot produce in seconds
(5760x2880 takes ~30 s per image on one CPU core).

Same scene description (synth.Scene: room, spheres, tube, value-noise
texture), same camera convention and double precision; the photometric
jitter uses torch's generator, so images are not bit-identical to synth.py's
(nothing compares them: they are inputs, and the ground-truth depth comes
from the same cast).
"""
from __future__ import annotations

import numpy as np
import torch

from . import synth


class TorchScene:
    def __init__(self, scene: synth.Scene, device):
        self.dev = device
        f64 = torch.float64
        self.lo = torch.tensor(synth.ROOM_LO, dtype=f64, device=device)
        self.hi = torch.tensor(synth.ROOM_HI, dtype=f64, device=device)
        self.spheres = torch.tensor(scene.spheres, dtype=f64, device=device).reshape(-1, 4)
        self.tube = scene.tube
        self.perm = torch.tensor(scene.noise.perm, dtype=torch.int64, device=device)
        self.vals = torch.tensor(scene.noise.vals, dtype=f64, device=device)
        self.mask = scene.noise.mask
        self.base_freq = scene.base_freq

    def _noise(self, x, y, z):
        ix, iy, iz = torch.floor(x), torch.floor(y), torch.floor(z)
        fx, fy, fz = x - ix, y - iy, z - iz
        ix, iy, iz = ix.long(), iy.long(), iz.long()
        fade = lambda t: t * t * t * (t * (t * 6.0 - 15.0) + 10.0)
        ux, uy, uz = fade(fx), fade(fy), fade(fz)
        p, m = self.perm, self.mask
        out = torch.zeros_like(x)
        for dx in (0, 1):
            wx = ux if dx else 1.0 - ux
            hx = p[(ix + dx) & m]
            for dy in (0, 1):
                wy = uy if dy else 1.0 - uy
                hy = p[(hx + iy + dy) & m]
                for dz in (0, 1):
                    wz = uz if dz else 1.0 - uz
                    out = out + wx * wy * wz * self.vals[p[(hy + iz + dz) & m]]
        return out

    def texture(self, P, obj):
        Q = P * self.base_freq + (obj.to(torch.float64) * 17.31)[:, None]
        val = torch.zeros(P.shape[0], dtype=torch.float64, device=self.dev)
        amp, f = 1.0, 1.0
        for _ in range(4):
            val = val + amp * self._noise(Q[:, 0] * f, Q[:, 1] * f, Q[:, 2] * f)
            amp *= 0.5
            f *= 2.0
        return torch.clamp(128.0 + 110.0 * val, 0.0, 255.0)

    def cast(self, C, D):
        n = D.shape[0]
        inf = torch.tensor(float("inf"), dtype=torch.float64, device=self.dev)
        tb = torch.where(D > 0, (self.hi - C) / D, (self.lo - C) / D)
        tb = torch.where(torch.isfinite(tb) & (tb > 0), tb, inf)
        t, face = torch.min(tb, dim=1)
        obj = face * 2 + (D.gather(1, face[:, None])[:, 0] > 0).long()
        for s in range(self.spheres.shape[0]):
            cx, cy, cz, r = [float(v) for v in self.spheres[s]]
            oc = C - torch.tensor([cx, cy, cz], dtype=torch.float64, device=self.dev)
            b = D @ oc
            c = float(oc @ oc) - r * r
            disc = b * b - c
            ok = disc > 0
            sq = torch.sqrt(torch.where(ok, disc, torch.zeros_like(disc)))
            t1, t2 = -b - sq, -b + sq
            ts = torch.where(t1 > 1e-9, t1, torch.where(t2 > 1e-9, t2, inf))
            ts = torch.where(ok, ts, inf)
            hit = ts < t
            t = torch.where(hit, ts, t)
            obj = torch.where(hit, torch.full_like(obj, 6 + s), obj)
        if self.tube is not None:
            y0, z0, R, x0, x1 = self.tube
            oy, oz = float(C[1]) - y0, float(C[2]) - z0
            a = D[:, 1] ** 2 + D[:, 2] ** 2
            b = D[:, 1] * oy + D[:, 2] * oz
            c = oy * oy + oz * oz - R * R
            disc = b * b - a * c
            ok = (disc > 0) & (a > 1e-12)
            sq = torch.sqrt(torch.where(ok, disc, torch.zeros_like(disc)))
            aa = torch.where(a > 1e-12, a, torch.ones_like(a))
            best = torch.full((n,), float("inf"), dtype=torch.float64, device=self.dev)
            for root in ((-b - sq) / aa, (-b + sq) / aa):
                xh = float(C[0]) + root * D[:, 0]
                valid = ok & (root > 1e-9) & (xh >= x0) & (xh <= x1)
                best = torch.where(valid & (root < best), root, best)
            hit = best < t
            t = torch.where(hit, best, t)
            obj = torch.where(hit, torch.full_like(obj, 6 + self.spheres.shape[0]), obj)
        return t, obj

    def render(self, pose, W, H):
        """Noise-free intensity and depth (H x W, float64 tensors) of one camera."""
        pose = np.asarray(pose, dtype=np.float64)
        R = torch.tensor(pose[:9].reshape(3, 3), dtype=torch.float64, device=self.dev)
        t = torch.tensor(pose[9:], dtype=torch.float64, device=self.dev)
        C = -R.T @ t
        j = torch.arange(W, dtype=torch.float64, device=self.dev)
        i = torch.arange(H, dtype=torch.float64, device=self.dev)
        lon = 2.0 * np.pi * (j / W - 0.5)
        lat = -np.pi * (i / H - 0.5)
        cl, sl = torch.cos(lat)[:, None], torch.sin(lat)[:, None]
        B = torch.stack([cl * torch.sin(lon)[None, :], (-sl).expand(H, W), cl * torch.cos(lon)[None, :]], -1)
        D = B.reshape(-1, 3) @ R
        depth, obj = self.cast(C, D)
        P = C + depth[:, None] * D
        img = self.texture(P, obj)
        return img.reshape(H, W), depth.reshape(H, W)


def photometric(img, gen):
    """gain U(0.8,1.2), bias U(-10,10), N(0, 2^2) noise, clamp, round -> uint8 (torch)."""
    dev = img.device
    u = torch.rand(2, generator=gen, device=dev, dtype=torch.float64)
    gain = 0.8 + 0.4 * float(u[0])
    bias = -10.0 + 20.0 * float(u[1])
    noise = torch.randn(img.shape, generator=gen, device=dev, dtype=torch.float64) * 2.0
    return torch.clamp(torch.round(img * gain + bias + noise), 0, 255).to(torch.uint8)


def sequence(seed: int, n: int, W: int, H: int, spacing: float = 0.2, device="cuda", mask_frac: float = 0.15):
    """n keyframes along the seeded trajectory through the config-1 scene
    (room + 20 spheres + tube), rendered at W x H: dict with 'poses' (n, 12),
    'images' (list of (H, W) uint8 numpy), 'gt' (list of (H, W) float32
    depth), 'mask' (H, W) uint8."""
    s_scene, s_photo = synth._split_seeds(seed ^ 0x5E0, 2)
    scene = synth.make_scene(s_scene, spheres=20, tube=True, cam_region=2.7)
    ts = TorchScene(scene, device)
    poses = synth.trajectory(seed, n, spacing)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(s_photo & 0x7FFFFFFFFFFFFFFF))
    images, gts = [], []
    for k in range(n):
        img, dep = ts.render(poses[k], W, H)
        images.append(photometric(img, gen).cpu().numpy())
        gts.append(dep.to(torch.float32).cpu().numpy())
    mask = np.ones((H, W), np.uint8)
    mask[int(np.floor((1.0 - mask_frac) * H)):, :] = 0
    return {"poses": poses, "images": images, "gt": gts, "mask": mask}
