"""COLMAP sparse model -> images/ cams/ pair.txt (hpm_mi355x.colmap, after
colmap2mvsnet_acm.py:305-451), on the CPU.

A synthetic COLMAP workspace is written from a ray-cast scene: one PINHOLE
camera, each image's pose as a unit quaternion, 3D points sampled on the
visible surfaces and observed by every image that sees them unoccluded.
Checks:
  * quaternion -> rotation on closed forms (identity, 90 deg about z, and the
    round trip of every scene rotation);
  * cams files: the scene's R, t and K, and the depth line recomputed here
    from the observed points' camera-frame z (1st / 99th percentile x 0.75 /
    1.25, 192 steps);
  * pair.txt: every pair score equals the count of shared points recomputed
    here, each list holds the min(20, n - 1) best in non-increasing order;
  * a pair whose shared points are seen under less than 1 degree scores 0;
  * the text and binary forms of the same model convert to identical files;
  * --max_d 0: the inverse-depth step count of one pixel at depth_min.
"""
import os
import struct

import numpy as np
import pytest

from hpm_mi355x import colmap, scenes


def _rot_to_quat(R):
    """Unit quaternion (w, x, y, z) of a rotation matrix (Shepperd's method)."""
    tr = np.trace(R)
    if tr > 0:
        s = 2.0 * np.sqrt(tr + 1.0)
        q = [0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s]
    else:
        i = int(np.argmax(np.diag(R)))
        j, k = (i + 1) % 3, (i + 2) % 3
        s = 2.0 * np.sqrt(1.0 + R[i, i] - R[j, j] - R[k, k])
        q = [0.0] * 4
        q[0] = (R[k, j] - R[j, k]) / s
        q[1 + i] = 0.25 * s
        q[1 + j] = (R[j, i] + R[i, j]) / s
        q[1 + k] = (R[k, i] + R[i, k]) / s
    q = np.array(q)
    return q if q[0] >= 0 else -q


def _observations(sc, n_points=400, seed=0):
    """World points on visible surfaces and, per image, the ids it observes."""
    rng = np.random.default_rng(seed)
    pts = []
    for v, cam in enumerate(sc.cameras):
        ys = rng.integers(0, cam.height, n_points // len(sc.cameras))
        xs = rng.integers(0, cam.width, n_points // len(sc.cameras))
        for x, y in zip(xs, ys):
            d = float(sc.depths[v][y, x])
            Xc = np.array([(x - cam.K[0, 2]) / cam.K[0, 0] * d, (y - cam.K[1, 2]) / cam.K[1, 1] * d, d])
            pts.append(cam.R.T @ (Xc - cam.t))
    obs = []
    for v, cam in enumerate(sc.cameras):
        ids, xys = [], []
        for pid, X in enumerate(pts):
            Xc = cam.R @ X + cam.t
            u, w = cam.K[0, 0] * Xc[0] / Xc[2] + cam.K[0, 2], cam.K[1, 1] * Xc[1] / Xc[2] + cam.K[1, 2]
            iu, iv = int(round(u)), int(round(w))
            if 0 <= iu < cam.width and 0 <= iv < cam.height and abs(sc.depths[v][iv, iu] - Xc[2]) < 0.01 * Xc[2]:
                ids.append(pid + 1)  # COLMAP point ids start at 1
                xys.append((u, w))
        obs.append((ids, xys))
    return pts, obs


def _write_text_model(root, sc, pts, obs, image_ids):
    sp = os.path.join(root, "sparse")
    os.makedirs(sp, exist_ok=True)
    os.makedirs(os.path.join(root, "images"), exist_ok=True)
    c0 = sc.cameras[0]
    with open(os.path.join(sp, "cameras.txt"), "w") as fh:
        fh.write("# Camera list\n")
        fh.write(f"1 PINHOLE {c0.width} {c0.height} " + " ".join(repr(float(x)) for x in
                                                           (c0.K[0, 0], c0.K[1, 1], c0.K[0, 2], c0.K[1, 2])) + "\n")
    with open(os.path.join(sp, "images.txt"), "w") as fh:
        fh.write("# Image list\n")
        for v, cam in enumerate(sc.cameras):
            q = _rot_to_quat(cam.R)
            fh.write(f"{image_ids[v]} " + " ".join(repr(float(x)) for x in q) + " " +
                     " ".join(repr(float(x)) for x in cam.t) + f" 1 view{v}.pgm\n")
            ids, xys = obs[v]
            fh.write(" ".join(f"{float(u)!r} {float(w)!r} {pid}" for pid, (u, w) in zip(ids, xys)) + "\n")
            with open(os.path.join(root, "images", f"view{v}.pgm"), "wb") as im:
                im.write(f"P5\n{cam.width} {cam.height}\n255\n".encode() + sc.images[v].tobytes())
    with open(os.path.join(sp, "points3D.txt"), "w") as fh:
        fh.write("# 3D point list\n")
        for pid, X in enumerate(pts):
            fh.write(f"{pid + 1} {float(X[0])!r} {float(X[1])!r} {float(X[2])!r} 128 128 128 0.5\n")


def _write_binary_model(root, sc, pts, obs, image_ids):
    sp = os.path.join(root, "sparse")
    c0 = sc.cameras[0]
    with open(os.path.join(sp, "cameras.bin"), "wb") as fh:
        fh.write(struct.pack("<Q", 1))
        fh.write(struct.pack("<iiQQ", 1, 1, c0.width, c0.height))
        fh.write(struct.pack("<4d", c0.K[0, 0], c0.K[1, 1], c0.K[0, 2], c0.K[1, 2]))
    with open(os.path.join(sp, "images.bin"), "wb") as fh:
        fh.write(struct.pack("<Q", len(sc.cameras)))
        for v, cam in enumerate(sc.cameras):
            q = _rot_to_quat(cam.R)
            fh.write(struct.pack("<i7di", image_ids[v], *q, *cam.t, 1))
            fh.write(f"view{v}.pgm".encode() + b"\x00")
            ids, xys = obs[v]
            fh.write(struct.pack("<Q", len(ids)))
            for pid, (u, w) in zip(ids, xys):
                fh.write(struct.pack("<ddq", u, w, pid))
    with open(os.path.join(sp, "points3D.bin"), "wb") as fh:
        fh.write(struct.pack("<Q", len(pts)))
        for pid, X in enumerate(pts):
            fh.write(struct.pack("<QdddBBBd", pid + 1, *X, 128, 128, 128, 0.5))
            fh.write(struct.pack("<Q", 0))


@pytest.fixture(scope="module")
def workspace(tmp_path_factory):
    sc = scenes.make_plane_scene(seed=8, n_cams=6, width=96, height=72, f=78.75, n_src=5)
    pts, obs = _observations(sc)
    root = str(tmp_path_factory.mktemp("colmap"))
    image_ids = [7, 3, 11, 5, 9, 20]  # unordered COLMAP ids: output order is by id
    _write_text_model(root, sc, pts, obs, image_ids)
    return sc, pts, obs, root, image_ids


def test_quaternion_closed_forms():
    np.testing.assert_allclose(colmap.quaternion_to_rotation([1, 0, 0, 0]), np.eye(3))
    c = np.sqrt(0.5)
    np.testing.assert_allclose(colmap.quaternion_to_rotation([c, 0, 0, c]),
                               [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-15)
    R = scenes.look_at_camera((1.0, -0.5, -3.0), (0, 0, 0), 100, 0, 0, 1, 1).R
    np.testing.assert_allclose(colmap.quaternion_to_rotation(_rot_to_quat(R)), R, atol=1e-14)


def _read_cam(path):
    tok = open(path).read().split()
    E = np.array([float(x) for x in tok[1:17]]).reshape(4, 4)
    K = np.array([float(x) for x in tok[18:27]]).reshape(3, 3)
    return E, K, [float(x) for x in tok[27:31]]


def test_cams_and_depth_ranges(workspace, tmp_path):
    sc, pts, obs, root, image_ids = workspace
    n = colmap.convert(root, str(tmp_path))
    assert n == 6
    order = np.argsort(image_ids)  # output index i <- the i-th smallest id
    for i, v in enumerate(order):
        E, K, depth = _read_cam(os.path.join(tmp_path, "cams", f"{i:08d}_cam.txt"))
        cam = sc.cameras[v]
        np.testing.assert_allclose(E[:3, :3], cam.R, atol=1e-12)
        np.testing.assert_allclose(E[:3, 3], cam.t, atol=1e-12)
        np.testing.assert_allclose(K, cam.K, atol=1e-12)
        zs = sorted((cam.R @ pts[pid - 1] + cam.t)[2] for pid in obs[v][0])
        dmin, dmax = zs[int(len(zs) * 0.01)] * 0.75, zs[int(len(zs) * 0.99)] * 1.25
        np.testing.assert_allclose(depth, [dmin, (dmax - dmin) / 191, 192, dmax], rtol=1e-6, atol=1e-6)
        assert os.path.exists(os.path.join(tmp_path, "images", f"{i:08d}.pgm"))


def test_pair_scores_and_order(workspace, tmp_path):
    sc, pts, obs, root, image_ids = workspace
    colmap.convert(root, str(tmp_path))
    order = list(np.argsort(image_ids))
    tok = open(os.path.join(tmp_path, "pair.txt")).read().split()
    assert int(tok[0]) == 6
    k = 1
    for i in range(6):
        assert int(tok[k]) == i
        m = int(tok[k + 1])
        assert m == 5
        entries = [(int(tok[k + 2 + 2 * e]), int(tok[k + 3 + 2 * e])) for e in range(m)]
        k += 2 + 2 * m
        scores = [s for _, s in entries]
        assert scores == sorted(scores, reverse=True)
        for j, s in entries:
            if j == i:
                continue
            shared = set(obs[order[i]][0]) & set(obs[order[j]][0])
            assert s == len(shared), (i, j)
        assert max(scores) > 10


def test_small_triangulation_angle_scores_zero(tmp_path):
    """Two cameras 1 cm apart looking at points 3 m away: every shared point
    is seen under about 0.2 deg, so their score is 0; a third camera 1 m away
    keeps its counts."""
    root = str(tmp_path / "ws")
    os.makedirs(os.path.join(root, "sparse"))
    os.makedirs(os.path.join(root, "images"))
    pts = [np.array([x, y, 3.0]) for x in (-0.2, 0.0, 0.2) for y in (-0.2, 0.2)]
    centers = [np.zeros(3), np.array([0.01, 0, 0]), np.array([1.0, 0, 0])]
    with open(os.path.join(root, "sparse", "cameras.txt"), "w") as fh:
        fh.write("1 SIMPLE_PINHOLE 64 48 50 32 24\n")
    with open(os.path.join(root, "sparse", "images.txt"), "w") as fh:
        for k, C in enumerate(centers):
            t = -C
            fh.write(f"{k + 1} 1 0 0 0 {t[0]} {t[1]} {t[2]} 1 im{k}.pgm\n")
            fh.write(" ".join(f"0 0 {p + 1}" for p in range(len(pts))) + "\n")
            with open(os.path.join(root, "images", f"im{k}.pgm"), "wb") as im:
                im.write(b"P5\n64 48\n255\n" + bytes(64 * 48))
    with open(os.path.join(root, "sparse", "points3D.txt"), "w") as fh:
        for p, X in enumerate(pts):
            fh.write(f"{p + 1} {X[0]} {X[1]} {X[2]} 0 0 0 0\n")
    out = str(tmp_path / "out")
    os.makedirs(out)
    colmap.convert(root, out)
    tok = open(os.path.join(out, "pair.txt")).read().split()
    lists = {}
    k = 1
    for _ in range(3):
        i, m = int(tok[k]), int(tok[k + 1])
        lists[i] = {int(tok[k + 2 + 2 * e]): int(tok[k + 3 + 2 * e]) for e in range(m)}
        k += 2 + 2 * m
    assert lists[0][1] == 0
    # image 1's zero-score neighbour ties with its own zero score: the
    # reversed argsort keeps itself (index 1) and drops image 0, as the
    # reference's view selection does
    assert set(lists[1]) == {2, 1} and lists[1][1] == 0
    assert lists[0][2] == len(pts) and lists[2][0] == len(pts) and lists[2][1] == len(pts)


def test_binary_model_converts_identically(workspace, tmp_path):
    sc, pts, obs, root, image_ids = workspace
    _write_binary_model(root, sc, pts, obs, image_ids)
    a, b = tmp_path / "txt", tmp_path / "bin"
    os.makedirs(a)
    os.makedirs(b)
    colmap.convert(root, str(a), model_ext=".txt")
    colmap.convert(root, str(b), model_ext=".bin")
    for rel in ["pair.txt"] + [f"cams/{i:08d}_cam.txt" for i in range(6)]:
        assert open(a / rel).read() == open(b / rel).read(), rel


def test_inverse_depth_steps(workspace, tmp_path):
    """--max_d 0: depth_num = (1/dmin - 1/dmax) / (1/dmin - 1/(dmin + s)),
    s the world distance between the rays of the principal point and its
    right neighbour at depth dmin, i.e. dmin / fx for a pinhole."""
    sc, pts, obs, root, image_ids = workspace
    colmap.convert(root, str(tmp_path), max_d=0)
    E, K, (dmin, interval, num, dmax) = _read_cam(os.path.join(tmp_path, "cams", "00000000_cam.txt"))
    s = dmin / K[0, 0]
    expect = (1 / dmin - 1 / dmax) / (1 / dmin - 1 / (dmin + s))
    assert abs(num - expect) <= 1e-5 * expect + 1e-6
    assert abs(interval - (dmax - dmin) / (num - 1)) <= 1e-5 * interval


def test_image_formats_are_copied_or_converted(tmp_path):
    """JPEG, PGM and PPM files are copied byte for byte (JPEG as %08d.jpg, as
    colmap2mvsnet_acm.py:450-451 copies them); a PNG is decoded and written
    losslessly as PPM."""
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(0)
    rgb = rng.integers(0, 255, (12, 16, 3), dtype=np.uint8)
    src = tmp_path / "src"
    out = tmp_path / "out"
    src.mkdir()
    out.mkdir()
    Image.fromarray(rgb).save(src / "a.JPG", format="JPEG", quality=90)
    Image.fromarray(rgb).save(src / "b.png")
    (src / "c.pgm").write_bytes(b"P5\n2 1\n255\n\x01\x02")
    for i, name in enumerate(["a.JPG", "b.png", "c.pgm"]):
        colmap._copy_image(str(src / name), str(out), i)
    assert (out / "00000000.jpg").read_bytes() == (src / "a.JPG").read_bytes()
    ppm = (out / "00000001.ppm").read_bytes()
    assert ppm.startswith(b"P6") and ppm.endswith(rgb.tobytes())
    assert (out / "00000002.pgm").read_bytes() == (src / "c.pgm").read_bytes()
