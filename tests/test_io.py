"""Output formats (pmp_write_dense_msgpack, pmp_write_ply; host code, no
GPU): the documents the reference's map_database_io writes
(io/map_database_io.cc:22-81) and its export_dense_msg_to_ply.py reads."""
import msgpack
import numpy as np
import pytest

from patchmatch_amd import _lib


def _rot(yaw, pitch):
    cy, sy, cp, sp = np.cos(yaw), np.sin(yaw), np.cos(pitch), np.sin(pitch)
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])
    return Ry @ Rx


def _quat_to_rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _keyframes(rng):
    kfs = []
    for k, (yaw, pitch) in enumerate([(0.3, 0.1), (2.9, -0.4), (-2.0, 1.2), (np.pi, 0.0)]):
        n = int(rng.integers(0, 40)) if k != 1 else 0
        R = _rot(yaw, pitch)
        kfs.append(dict(id=[7, 12, 100, 3][k], src_frm_id=50 + k, timestamp=0.5 * k,
                        pose=np.concatenate([R.ravel(), rng.normal(size=3)]), points=rng.normal(size=(n, 3)),
                        colors_bgr=rng.integers(0, 256, (n, 3)), keypts=rng.integers(0, 960, (n, 2))))
    return kfs


def test_dense_msgpack_document(tmp_path):
    rng = np.random.default_rng(1)
    kfs = _keyframes(rng)
    path = str(tmp_path / "map.msg_dense")
    _lib.write_dense_msgpack(path, "equirect_cam", 1920, 960, kfs, fps=10.0)
    doc = msgpack.unpackb(open(path, "rb").read(), raw=False)
    assert list(doc) == ["cameras", "keyframes", "landmarks"]
    cam = doc["cameras"]["equirect_cam"]
    assert (cam["model_type"], cam["cols"], cam["rows"], cam["fps"]) == ("Equirectangular", 1920, 960, 10.0)
    assert list(cam) == sorted(cam)
    # landmarks numbered over the keyframes in the given order, keys in byte order
    total = sum(len(kf["points"]) for kf in kfs)
    lms = doc["landmarks"]
    assert list(lms) == sorted(str(i) for i in range(total))
    lid = 0
    for kf in kfs:
        entry = doc["keyframes"][str(kf["id"])]
        assert list(entry) == sorted(entry)
        assert entry["lm_ids"] == list(range(lid, lid + len(kf["points"])))
        assert entry["keypts"] == [{"pt": [int(c), int(r)]} for r, c in kf["keypts"]]   # [x, y] = [col, row]
        R = kf["pose"][:9].reshape(3, 3)
        np.testing.assert_allclose(_quat_to_rot(entry["rot_cw"]), R, atol=1e-12)
        np.testing.assert_allclose(np.linalg.norm(entry["rot_cw"]), 1.0, atol=1e-12)
        assert entry["trans_cw"] == list(kf["pose"][9:])
        assert (entry["src_frm_id"], entry["ts"]) == (kf["src_frm_id"], kf["timestamp"])
        for q in range(len(kf["points"])):
            lm = lms[str(lid + q)]
            assert list(lm) == ["1st_keyfrm", "color", "n_fnd", "n_vis", "pos_w", "ref_keyfrm"]
            assert lm["pos_w"] == list(kf["points"][q])
            b, g, r = kf["colors_bgr"][q]
            assert lm["color"] == [float(r), float(g), float(b)]
            assert (lm["1st_keyfrm"], lm["ref_keyfrm"], lm["n_vis"], lm["n_fnd"]) == (kf["id"], kf["id"], 1, 1)
        lid += len(kf["points"])
    assert list(doc["keyframes"]) == sorted(str(kf["id"]) for kf in kfs)


def test_ply_matches_the_msgpack_export(tmp_path):
    """The PLY written directly equals what the reference's exporter derives
    from the MessagePack file (landmarks' pos_w and color, in file order)."""
    rng = np.random.default_rng(2)
    kfs = _keyframes(rng)
    path = str(tmp_path / "map.msg_dense")
    _lib.write_dense_msgpack(path, "cam", 960, 480, kfs)
    doc = msgpack.unpackb(open(path, "rb").read(), raw=False)
    pts = np.array([lm["pos_w"] for lm in doc["landmarks"].values()])
    rgb = np.array([lm["color"] for lm in doc["landmarks"].values()], dtype=np.uint8)
    ply = str(tmp_path / "map.ply")
    _lib.write_ply(ply, pts, rgb)
    lines = open(ply).read().splitlines()
    assert lines[:10] == ["ply", "format ascii 1.0", f"element vertex {len(pts)}", "property float x",
                          "property float y", "property float z", "property uchar red", "property uchar green",
                          "property uchar blue", "end_header"]
    body = np.array([[float(v) for v in ln.split()] for ln in lines[10:]]).reshape(-1, 6)
    np.testing.assert_allclose(body[:, :3], pts, atol=5e-7)
    np.testing.assert_array_equal(body[:, 3:], rgb)


def test_writers_reject_bad_arguments(tmp_path):
    with pytest.raises(_lib.PmpError):
        _lib.write_ply(str(tmp_path / "no_such_dir" / "x.ply"), np.zeros((1, 3)), np.zeros((1, 3)))
    L = _lib.load()
    assert L.pmp_write_ply(None, 0, None, None) == -1
    assert L.pmp_write_dense_msgpack(None, None, 0, None) == -1
