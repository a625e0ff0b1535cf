"""COLMAP sparse model -> the dense folder the pipeline reads (images/,
cams/%08d_cam.txt, pair.txt), after colmap2mvsnet_acm.py:305-451.

    python -m hpm_mi355x.colmap --dense_folder <colmap workspace> --save_folder <out>
        [--max_d 192] [--interval_scale 1] [--model_ext .txt|.bin]

The workspace holds images/ and sparse/{cameras,images,points3D}{.txt,.bin}
(COLMAP's documented text and binary model formats).  What the converter
writes, per registered image in increasing image-id order (output index i):

* cams/%08d_cam.txt: the world-to-camera extrinsic (R from the image's unit
  quaternion, t), the intrinsic K = [fx 0 cx; 0 fy cy; 0 0 1] from the first
  parameters of its camera model (distortion is not applied), and the depth
  line `depth_min interval depth_num depth_max`: the 1st / 99th percentile of
  the camera-frame z of the image's observed 3D points, times 0.75 / 1.25;
  depth_num is --max_d, or with --max_d 0 the number of inverse-depth steps of
  one pixel at depth_min (colmap2mvsnet_acm.py:367-397);
* pair.txt: for every image the min(20, n - 1) others of highest score, where
  the score of a pair is the number of 3D points both observe, set to 0 when
  the 75th percentile of their triangulation angles is below 1 degree
  (colmap2mvsnet_acm.py:281-303, 428-435);
* images/%08d.<ext>: the image itself.  JPEG, PGM and PPM files are copied
  unchanged (the reference copies JPEG files and re-encodes any other format
  as JPEG with OpenCV); other formats are decoded with Pillow and written
  losslessly as PGM / PPM.
"""
from __future__ import annotations

import argparse
import os
import shutil
import struct
from dataclasses import dataclass, field

import numpy as np

# COLMAP camera models: id -> (name, number of parameters, parameter names
# up to the principal point)
_MODELS = {
    0: ("SIMPLE_PINHOLE", 3, ("f", "cx", "cy")),
    1: ("PINHOLE", 4, ("fx", "fy", "cx", "cy")),
    2: ("SIMPLE_RADIAL", 4, ("f", "cx", "cy")),
    3: ("RADIAL", 5, ("f", "cx", "cy")),
    4: ("OPENCV", 8, ("fx", "fy", "cx", "cy")),
    5: ("OPENCV_FISHEYE", 8, ("fx", "fy", "cx", "cy")),
    6: ("FULL_OPENCV", 12, ("fx", "fy", "cx", "cy")),
    7: ("FOV", 5, ("fx", "fy", "cx", "cy")),
    8: ("SIMPLE_RADIAL_FISHEYE", 4, ("f", "cx", "cy")),
    9: ("RADIAL_FISHEYE", 5, ("f", "cx", "cy")),
    10: ("THIN_PRISM_FISHEYE", 12, ("fx", "fy", "cx", "cy")),
}
_MODEL_BY_NAME = {v[0]: (k, v) for k, v in _MODELS.items()}


@dataclass
class ColmapCamera:
    id: int
    model: str
    width: int
    height: int
    params: np.ndarray


@dataclass
class ColmapImage:
    id: int
    qvec: np.ndarray  # w, x, y, z
    tvec: np.ndarray
    camera_id: int
    name: str
    xys: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    point3d_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))


@dataclass
class ColmapPoint:
    id: int
    xyz: np.ndarray


# ------------------------------------------------------------------ reading
def _data_lines(path):
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if line and not line.startswith("#"):
                yield line


def _read_cameras_txt(path):
    cams = {}
    for line in _data_lines(path):
        e = line.split()
        cams[int(e[0])] = ColmapCamera(int(e[0]), e[1], int(e[2]), int(e[3]), np.array([float(x) for x in e[4:]]))
    return cams


def _read_images_txt(path):
    # one line of pose, then one line of 2D points (x y point3D_id)*, which
    # may be empty; comment lines are skipped, empty lines are not
    imgs = {}
    with open(path) as fh:
        lines = [ln.rstrip("\n") for ln in fh if not ln.startswith("#")]
    k = 0
    while k < len(lines):
        if not lines[k].strip():
            k += 1
            continue
        e = lines[k].split()
        pts = lines[k + 1].split() if k + 1 < len(lines) else []
        k += 2
        img = ColmapImage(int(e[0]), np.array([float(x) for x in e[1:5]]), np.array([float(x) for x in e[5:8]]),
                          int(e[8]), e[9])
        img.xys = np.array([[float(pts[i]), float(pts[i + 1])] for i in range(0, len(pts), 3)]).reshape(-1, 2)
        img.point3d_ids = np.array([int(pts[i]) for i in range(2, len(pts), 3)], np.int64)
        imgs[img.id] = img
    return imgs


def _read_points_txt(path):
    pts = {}
    for line in _data_lines(path):
        e = line.split()
        pts[int(e[0])] = ColmapPoint(int(e[0]), np.array([float(x) for x in e[1:4]]))
    return pts


def _unpack(fh, fmt):
    size = struct.calcsize("<" + fmt)
    data = fh.read(size)
    if len(data) != size:
        raise ValueError("truncated COLMAP binary model")
    return struct.unpack("<" + fmt, data)


def _read_cameras_bin(path):
    cams = {}
    with open(path, "rb") as fh:
        (n,) = _unpack(fh, "Q")
        for _ in range(n):
            cid, mid, w, h = _unpack(fh, "iiQQ")
            name, npar, _ = _MODELS[mid]
            cams[cid] = ColmapCamera(cid, name, w, h, np.array(_unpack(fh, "d" * npar)))
    return cams


def _read_images_bin(path):
    imgs = {}
    with open(path, "rb") as fh:
        (n,) = _unpack(fh, "Q")
        for _ in range(n):
            v = _unpack(fh, "idddddddi")
            name = bytearray()
            while True:
                c = fh.read(1)
                if not c:
                    raise ValueError("truncated COLMAP binary model")
                if c == b"\x00":
                    break
                name += c
            (m,) = _unpack(fh, "Q")
            xyid = _unpack(fh, "ddq" * m) if m else ()
            img = ColmapImage(v[0], np.array(v[1:5]), np.array(v[5:8]), v[8], name.decode("utf-8"))
            img.xys = np.array([[xyid[3 * i], xyid[3 * i + 1]] for i in range(m)]).reshape(-1, 2)
            img.point3d_ids = np.array([xyid[3 * i + 2] for i in range(m)], np.int64)
            imgs[img.id] = img
    return imgs


def _read_points_bin(path):
    pts = {}
    with open(path, "rb") as fh:
        (n,) = _unpack(fh, "Q")
        for _ in range(n):
            v = _unpack(fh, "QdddBBBd")
            (tl,) = _unpack(fh, "Q")
            fh.read(8 * tl)  # the track (image id, point2D index) pairs
            pts[v[0]] = ColmapPoint(v[0], np.array(v[1:4]))
    return pts


def read_model(sparse_dir, ext=".txt"):
    """(cameras, images, points3D) dicts keyed by COLMAP id."""
    if ext == ".txt":
        return (_read_cameras_txt(os.path.join(sparse_dir, "cameras.txt")),
                _read_images_txt(os.path.join(sparse_dir, "images.txt")),
                _read_points_txt(os.path.join(sparse_dir, "points3D.txt")))
    if ext == ".bin":
        return (_read_cameras_bin(os.path.join(sparse_dir, "cameras.bin")),
                _read_images_bin(os.path.join(sparse_dir, "images.bin")),
                _read_points_bin(os.path.join(sparse_dir, "points3D.bin")))
    raise ValueError("model extension must be .txt or .bin")


# ------------------------------------------------------------------ geometry
def quaternion_to_rotation(q):
    """Rotation matrix of the unit quaternion (w, x, y, z)."""
    w, x, y, z = (float(v) for v in q)
    return np.array([[1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * w * z, 2 * z * x + 2 * w * y],
                     [2 * x * y + 2 * w * z, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * w * x],
                     [2 * z * x - 2 * w * y, 2 * y * z + 2 * w * x, 1 - 2 * x * x - 2 * y * y]])


def intrinsic_matrix(cam: ColmapCamera):
    names = _MODEL_BY_NAME[cam.model][1][2]
    p = dict(zip(names, cam.params))
    fx = p.get("fx", p.get("f"))
    fy = p.get("fy", p.get("f"))
    return np.array([[fx, 0.0, p["cx"]], [0.0, fy, p["cy"]], [0.0, 0.0, 1.0]])


def extrinsic_matrix(img: ColmapImage):
    E = np.zeros((4, 4))
    E[:3, :3] = quaternion_to_rotation(img.qvec)
    E[:3, 3] = img.tvec
    E[3, 3] = 1.0
    return E


def depth_range(img, E, K, points, max_d=192, interval_scale=1.0):
    """(depth_min, interval, depth_num, depth_max) of colmap2mvsnet_acm.py:367-397."""
    zs = []
    for pid in img.point3d_ids:
        if pid == -1:
            continue
        X = np.append(points[int(pid)].xyz, 1.0)
        zs.append(float((E @ X)[2]))
    if not zs:
        raise ValueError(f"image {img.id} observes no 3D point: no depth range")
    zs.sort()
    dmin = zs[int(len(zs) * 0.01)] * 0.75
    dmax = zs[int(len(zs) * 0.99)] * 1.25
    if max_d == 0:  # inverse-depth steps of one pixel at depth_min
        Kinv = np.linalg.inv(K)
        R, t = E[:3, :3], E[:3, 3]
        Rinv = np.linalg.inv(R)
        P1 = Rinv @ (Kinv @ np.array([K[0, 2], K[1, 2], 1.0]) * dmin - t)
        P2 = Rinv @ (Kinv @ np.array([K[0, 2] + 1, K[1, 2], 1.0]) * dmin - t)
        num = (1 / dmin - 1 / dmax) / (1 / dmin - 1 / (dmin + np.linalg.norm(P2 - P1)))
    else:
        num = max_d
    return dmin, (dmax - dmin) / (num - 1) / interval_scale, num, dmax


def pair_score(img_i, img_j, E_i, E_j, points):
    """Shared observed points; 0 when the 75th percentile of their
    triangulation angles is below 1 degree (colmap2mvsnet_acm.py:281-303).
    Every occurrence in image i of a point that image j also observes counts."""
    in_j = set(int(p) for p in img_j.point3d_ids)
    C_i = -E_i[:3, :3].T @ E_i[:3, 3]
    C_j = -E_j[:3, :3].T @ E_j[:3, 3]
    angles = []
    for pid in img_i.point3d_ids:
        pid = int(pid)
        if pid == -1 or pid not in in_j:
            continue
        X = points[pid].xyz
        a, b = C_i - X, C_j - X
        angles.append((180 / np.pi) * np.arccos(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b)))
    score = float(len(angles))
    if angles and sorted(angles)[int(len(angles) * 0.75)] < 1:
        score = 0.0
    return score


def select_views(scores):
    """Per image the min(20, n - 1) highest-scoring images (the ordering of
    numpy's default argsort, reversed, as colmap2mvsnet_acm.py:431-434)."""
    n = len(scores)
    k = min(20, n - 1)
    return [[(int(j), scores[i, j]) for j in np.argsort(scores[i])[::-1][:k]] for i in range(n)]


# ------------------------------------------------------------------ writing
def convert(dense_folder, save_folder, max_d=192, interval_scale=1.0, model_ext=".txt"):
    """Writes save_folder/{cams, images, pair.txt}; returns the number of images."""
    cameras, images, points = read_model(os.path.join(dense_folder, "sparse"), model_ext)
    order = sorted(images)  # output index i <- the i-th smallest image id
    imgs = [images[k] for k in order]
    n = len(imgs)
    E = [extrinsic_matrix(im) for im in imgs]
    K = [intrinsic_matrix(cameras[im.camera_id]) for im in imgs]
    ranges = [depth_range(im, e, k, points, max_d, interval_scale) for im, e, k in zip(imgs, E, K)]
    scores = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            scores[i, j] = scores[j, i] = pair_score(imgs[i], imgs[j], E[i], E[j], points)
    views = select_views(scores)

    cam_dir = os.path.join(save_folder, "cams")
    img_dir = os.path.join(save_folder, "images")
    for d in (cam_dir, img_dir):
        if os.path.exists(d):
            shutil.rmtree(d)
        os.makedirs(d)
    for i in range(n):
        with open(os.path.join(cam_dir, f"{i:08d}_cam.txt"), "w") as fh:
            fh.write("extrinsic\n")
            for r in range(4):
                fh.write("".join(repr(float(E[i][r, c])) + " " for c in range(4)) + "\n")
            fh.write("\nintrinsic\n")
            for r in range(3):
                fh.write("".join(repr(float(K[i][r, c])) + " " for c in range(3)) + "\n")
            fh.write("\n%f %f %f %f\n" % ranges[i])
    with open(os.path.join(save_folder, "pair.txt"), "w") as fh:
        fh.write(f"{n}\n")
        for i, sel in enumerate(views):
            fh.write(f"{i}\n{len(sel)} " + "".join(f"{j} {int(s)} " for j, s in sel) + "\n")
    for i, im in enumerate(imgs):
        _copy_image(os.path.join(dense_folder, "images", im.name), img_dir, i)
    return n


def _copy_image(src, img_dir, i):
    """images/%08d.<ext> of output index i.  JPEG files are copied unchanged,
    as colmap2mvsnet_acm.py:450-451 does, and so are binary PGM / PPM files;
    the reference re-encodes any other format as JPEG (:447-449), here it is
    decoded with Pillow, when that is importable, and written losslessly as
    PGM (grey) or PPM (colour)."""
    ext = os.path.splitext(src)[1].lower()
    if ext in (".jpg", ".jpeg"):
        shutil.copyfile(src, os.path.join(img_dir, f"{i:08d}.jpg"))
        return
    if ext in (".pgm", ".ppm"):
        shutil.copyfile(src, os.path.join(img_dir, f"{i:08d}{ext}"))
        return
    try:
        from PIL import Image
    except ImportError as e:
        raise ValueError(f"{src}: only JPEG, PGM and PPM images can be copied without Pillow") from e
    with Image.open(src) as img:
        grey = img.mode in ("L", "I;16", "I", "1")
        img.convert("L" if grey else "RGB").save(os.path.join(img_dir, f"{i:08d}.{'pgm' if grey else 'ppm'}"))


def main(argv=None):
    ap = argparse.ArgumentParser(description="COLMAP sparse model -> images/, cams/, pair.txt")
    ap.add_argument("--dense_folder", required=True)
    ap.add_argument("--save_folder", required=True)
    ap.add_argument("--max_d", type=int, default=192)
    ap.add_argument("--interval_scale", type=float, default=1.0)
    ap.add_argument("--model_ext", default=".txt", choices=[".txt", ".bin"])
    a = ap.parse_args(argv)
    os.makedirs(a.save_folder, exist_ok=True)
    n = convert(a.dense_folder, a.save_folder, a.max_d, a.interval_scale, a.model_ext)
    print(f"converted {n} images into {a.save_folder}")


if __name__ == "__main__":
    main()
