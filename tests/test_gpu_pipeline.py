"""The dense-folder pipeline through libhpm_host.so on an MI355X: the
single-level schedule of main.cpp:1249-1291 (photometric pass, confidence,
planar-prior pass, three geometric passes with mandatory consistency, fusion)
on a synthetic dense folder (images/*.pgm, cams/*_cam.txt, pair.txt).

Checks: every pass writes the reference's files (.dmb with the reference
layout, both PLY files); the final geometric depth maps agree with the
ray-cast ground truth; fusion through the pipeline equals the oracle's
consistency check on the maps it read (same points, same order).

The pyramid schedule (main.cpp:1183-1352) runs on a 320x240 folder with a
size bound of 100: levels of 80x60, 160x120 and 320x240, so JBU, the
hierarchy pass and prior passes from 2 and 1 levels coarser all run; the
finest level's depth maps must reach the same ground-truth accuracy.
"""
import os

import numpy as np
import pytest

import oracle_py as O
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


@pytest.fixture(scope="module")
def dense_folder(tmp_path_factory):
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=4, n_cams=5, width=160, height=120, f=131.25, n_src=4)
    folder = str(tmp_path_factory.mktemp("dense"))
    scenes.write_dense_folder(sc, folder)
    return sc, folder


def test_pipeline_runs_and_fuses(dense_folder):
    from hpm_mi355x import pipeline as P
    sc, folder = dense_folder
    print("case: run_pipeline 5 views 160x120 fp32", flush=True)
    n = P.run_pipeline(folder, precision="fp32", seed=3)
    assert n > 5000
    out = os.path.join(folder, "HPM_MVS_plusplus")
    for i in range(5):
        rf = os.path.join(out, f"2333_{i:08d}")
        for f in ("depths.dmb", "depths_geom.dmb", "normals.dmb", "costs.dmb", "confidence.dmb", "texture0.dmb"):
            assert os.path.exists(os.path.join(rf, f)), f
        d = P.read_dmb(os.path.join(rf, "depths_geom.dmb"))
        nm = P.read_dmb(os.path.join(rf, "normals.dmb"))
        assert d.shape == (120, 160) and nm.shape == (120, 160, 3)
        gt = sc.depths[i]
        rel = np.abs(d - gt) / gt
        print(f"view {i}: depth within 1% of GT {np.mean(rel <= 0.01):.3f}")
        assert np.mean(rel <= 0.01) > 0.75
    with open(os.path.join(out, "HPM_MVS_plusplus.ply"), "rb") as fh:
        head = fh.read(400).split(b"end_header\n")[0].decode()
    assert f"element vertex {n}" in head
    with open(os.path.join(out, "fused.ply"), "rb") as fh:
        data = fh.read()
    hdr, body = data.split(b"end_header\n", 1)
    k = int([l for l in hdr.decode().splitlines() if l.startswith("element vertex")][0].split()[-1])
    assert len(body) == k * 27 and k <= n


def test_pipeline_fusion_equals_oracle(dense_folder):
    """Fusion of the pipeline's final maps against the oracle's RunFusion."""
    from hpm_mi355x import pipeline as P
    import hpm_mi355x as H
    sc, folder = dense_folder
    out = os.path.join(folder, "HPM_MVS_plusplus")
    depths = [np.ascontiguousarray(P.read_dmb(os.path.join(out, f"2333_{i:08d}", "depths_geom.dmb"))) for i in range(5)]
    normals = [np.ascontiguousarray(P.read_dmb(os.path.join(out, f"2333_{i:08d}", "normals.dmb"))) for i in range(5)]
    sources = [[j for j, s in sc.pairs[i] if s > 0] for i in range(5)]
    cams, hc = [], []
    for i, c in enumerate(sc.cameras):  # the camera as the file holds it (decimal text -> float32)
        tok = open(os.path.join(folder, "cams", f"{i:08d}_cam.txt")).read().split()
        ext = np.array(tok[1:17], np.float32).reshape(4, 4)
        K = np.array(tok[18:27], np.float32).reshape(3, 3)
        dmin, dmax = np.float32(tok[27]), np.float32(tok[30])
        cams.append(O.camera(K, ext[:3, :3], ext[:3, 3], c.width, c.height, dmin, dmax))
        hc.append(H.make_camera(K, ext[:3, :3], ext[:3, 3], c.width, c.height, dmin, dmax))
    rp, rx = O.consistency(0, cams, depths, normals, sources)
    n = P.fusion(folder, geom=True)
    assert n == len(rp)
    gp, gx = H.consistency_check(H.CHECK_FUSION, hc, depths, normals, sources)
    np.testing.assert_array_equal(gp, rp)
    np.testing.assert_array_equal(gx, rx)
    # the PLY files hold those points in that order (ExportPointCloud /
    # ExportFusedPointCloud, HPM.cpp:410-498): position, normalised normal,
    # and the colour of the fused pixel (a grey image repeats its value)
    grey = [np.fromfile(os.path.join(folder, "images", f"{i:08d}.pgm"), np.uint8)[-160 * 120:] for i in range(5)]
    col = np.array([grey[v][p] for v, p in rp], np.uint8)
    with open(os.path.join(out, "HPM_MVS_plusplus.ply"), "rb") as fh:
        body = fh.read().split(b"end_header\n", 1)[1]
    rec = np.frombuffer(body, np.dtype([("x", "<f4", 3), ("c", "u1", 3)]))
    assert len(rec) == n
    np.testing.assert_array_equal(rec["x"], rx[:, :3])
    np.testing.assert_array_equal(rec["c"], np.repeat(col[:, None], 3, axis=1))
    with open(os.path.join(out, "fused.ply"), "rb") as fh:
        body = fh.read().split(b"end_header\n", 1)[1]
    rec = np.frombuffer(body, np.dtype([("x", "<f4", 3), ("n", "<f4", 3), ("c", "u1", 3)]))
    nrm = rx[:, 3:].astype(np.float64)
    keep = np.isfinite(rx[:, :3]).all(1) & (np.sum(nrm * nrm, 1) > 1e-20)
    assert len(rec) == int(keep.sum())
    np.testing.assert_array_equal(rec["x"], rx[keep, :3])
    np.testing.assert_allclose(rec["n"], (nrm[keep] / np.linalg.norm(nrm[keep], axis=1, keepdims=True)), atol=1e-6)
    np.testing.assert_array_equal(rec["c"], np.repeat(col[keep, None], 3, axis=1))


def test_sky_mask_fusion_equals_oracle(dense_folder):
    """--mask fusion (RunFusion_Sky_Strict) over the pipeline's final maps:
    masks/%08d.ppm paint a block of view 0 and a block of view 3 in the sky
    colour (R 55, G 235, B 234) and single channels of it elsewhere; the kept
    points are the oracle's strict-threshold pass filtered by the mask test,
    in order, and the PLY holds exactly them."""
    from hpm_mi355x import pipeline as P
    import hpm_mi355x as H
    sc, folder = dense_folder
    out = os.path.join(folder, "HPM_MVS_plusplus")
    if not os.path.exists(os.path.join(out, "2333_00000004", "depths_geom.dmb")):  # run alone
        P.run_pipeline(folder, precision="fp32", seed=3)
    os.makedirs(os.path.join(folder, "masks"), exist_ok=True)
    masks = []
    for i in range(5):
        m = np.full((120, 160, 3), 128, np.uint8)  # R G B
        if i == 0:
            m[20:60, 30:90] = (55, 235, 234)
        if i == 3:
            m[60:110, 100:150, 0] = 55     # R alone at its sky value also rejects
            m[0:30, 0:40, 1] = 235         # and G alone
        masks.append(m)
        with open(os.path.join(folder, "masks", f"{i:08d}.ppm"), "wb") as fh:
            fh.write(b"P6\n160 120\n255\n" + m.tobytes())
    n = P.fusion_mask(folder, geom=True)
    depths = [np.ascontiguousarray(P.read_dmb(os.path.join(out, f"2333_{i:08d}", "depths_geom.dmb"))) for i in range(5)]
    normals = [np.ascontiguousarray(P.read_dmb(os.path.join(out, f"2333_{i:08d}", "normals.dmb"))) for i in range(5)]
    sources = [[j for j, s in sc.pairs[i] if s > 0] for i in range(5)]
    cams = []
    for i, c in enumerate(sc.cameras):
        tok = open(os.path.join(folder, "cams", f"{i:08d}_cam.txt")).read().split()
        ext = np.array(tok[1:17], np.float32).reshape(4, 4)
        K = np.array(tok[18:27], np.float32).reshape(3, 3)
        cams.append(O.camera(K, ext[:3, :3], ext[:3, 3], c.width, c.height, np.float32(tok[27]), np.float32(tok[30])))
    rp, _ = O.consistency(2, cams, depths, normals, sources)
    keep = []
    for v, p in rp:
        r, g, b = masks[v].reshape(-1, 3)[p]
        keep.append(b != 234 and g != 235 and r != 55)
    keep = np.array(keep)
    print(f"strict pass {len(rp)} points, kept {keep.sum()}")
    assert 0 < keep.sum() < len(rp)
    assert n == keep.sum()
    with open(os.path.join(out, "HPM_MVS_plusplus_mask.ply"), "rb") as fh:
        head = fh.read(400).split(b"end_header\n")[0].decode()
    assert f"element vertex {n}" in head


def test_pyramid_pipeline_three_levels(tmp_path):
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=4, n_cams=5, width=320, height=240, f=262.5, n_src=4)
    folder = str(tmp_path)
    scenes.write_dense_folder(sc, folder)
    print("case: run_pyramid_pipeline 5 views 320x240 bound 100 fp16", flush=True)
    n = P.run_pyramid_pipeline(folder, initial_image_scale=1, size_bound=100, precision="fp16", seed=9)
    assert n > 4 * 5000
    out = os.path.join(folder, "HPM_MVS_plusplus")
    for i in range(5):
        rf = os.path.join(out, f"2333_{i:08d}")
        # the photometric / hierarchy pass of each level writes its texture map at the level's size
        for level, shape in ((2, (60, 80)), (1, (120, 160)), (0, (240, 320))):
            assert P.read_dmb(os.path.join(rf, f"texture{level}.dmb")).shape == shape
        d = P.read_dmb(os.path.join(rf, "depths_geom.dmb"))
        assert d.shape == (240, 320)
        gt = sc.depths[i]
        rel = np.abs(d - gt) / gt
        print(f"view {i}: depth within 1% of GT {np.mean(rel <= 0.01):.3f}")
        assert np.mean(rel <= 0.01) > 0.75
    with open(os.path.join(out, "HPM_MVS_plusplus.ply"), "rb") as fh:
        head = fh.read(400).split(b"end_header\n")[0].decode()
    assert f"element vertex {n}" in head


def test_pipeline_files_do_not_depend_on_host_threads(tmp_path):
    """The order-free passes run their views on several host threads; every
    file the pipeline writes (.dmb maps, both PLY files) is byte-identical to
    a run in view order."""
    import filecmp
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=6, n_cams=5, width=160, height=120, f=131.25, n_src=4)
    runs = {}
    for threads in (1, 4):
        folder = str(tmp_path / f"t{threads}")
        scenes.write_dense_folder(sc, folder)
        P.set_threads(threads)
        print(f"case: pyramid pipeline, 5 views 160x120, bound 100, {threads} host threads", flush=True)
        runs[threads] = (folder, P.run_pyramid_pipeline(folder, initial_image_scale=1, size_bound=100,
                                                        precision="fp16", seed=4))
    P.set_threads(0)
    assert runs[1][1] == runs[4][1]
    names = []
    for root, _, files in os.walk(os.path.join(runs[1][0], "HPM_MVS_plusplus")):
        for f in files:
            names.append(os.path.relpath(os.path.join(root, f), runs[1][0]))
    assert len(names) > 5 * 6
    for rel in names:
        assert filecmp.cmp(os.path.join(runs[1][0], rel), os.path.join(runs[4][0], rel), shallow=False), rel



def test_pipeline_fused_launch_for_large_views_identical(tmp_path):
    """Reference views from hpm_host_set_fused_min_pixels on run as a one-view
    fused launch with 16-column tiles (by default from 4 Mpx).  With the bound
    at 1 pixel every pass of a 3-level pyramid (photometric, prior, geometric,
    hierarchy, mandatory) takes that path, and every file is byte-identical
    to a run through the contexts' own launches."""
    import filecmp
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=6, n_cams=5, width=320, height=240, f=262.5, n_src=4)
    runs = {}
    try:
        for bound in (1, 1 << 40):
            folder = str(tmp_path / f"b{bound}")
            scenes.write_dense_folder(sc, folder)
            P.set_fused_min_pixels(bound)
            print(f"case: pyramid pipeline, 5 views 320x240, bound 100, fused launch from {bound} px", flush=True)
            runs[bound] = (folder, P.run_pyramid_pipeline(folder, initial_image_scale=1, size_bound=100,
                                                          precision="fp16", seed=4))
    finally:
        P.set_fused_min_pixels()
    a, b = runs[1], runs[1 << 40]
    assert a[1] == b[1] > 4 * 5000
    names = []
    for root, _, files in os.walk(os.path.join(a[0], "HPM_MVS_plusplus")):
        for f in files:
            names.append(os.path.relpath(os.path.join(root, f), a[0]))
    assert len(names) > 5 * 6
    for rel in names:
        assert filecmp.cmp(os.path.join(a[0], rel), os.path.join(b[0], rel), shallow=False), rel


def test_pipeline_reads_jpeg_images(tmp_path):
    """images/%08d.jpg (what the reference reads, main.cpp:125): a colour JPEG
    folder and a folder holding the decoded pixels as PPM give byte-identical
    output files, and the depth maps from the lossy images stay close to the
    ground-truth accuracy of the lossless grey images."""
    import filecmp
    import io
    Image = pytest.importorskip("PIL.Image")
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=6, n_cams=5, width=160, height=120, f=131.25, n_src=4)
    runs = {}
    for kind in ("pgm", "jpg", "ppm"):
        folder = str(tmp_path / kind)
        scenes.write_dense_folder(sc, folder)
        for i, im in enumerate(sc.images if kind != "pgm" else []):
            g = np.asarray(im, np.int32)
            rgb = np.stack([np.clip(g + 12, 0, 255), g, np.clip(g - 9, 0, 255)], -1).astype(np.uint8)
            buf = io.BytesIO()
            Image.fromarray(rgb).save(buf, format="JPEG", quality=95)
            os.remove(os.path.join(folder, "images", f"{i:08d}.pgm"))
            if kind == "jpg":
                with open(os.path.join(folder, "images", f"{i:08d}.jpg"), "wb") as fh:
                    fh.write(buf.getvalue())
            else:
                dec = np.asarray(Image.open(io.BytesIO(buf.getvalue())).convert("RGB"))
                with open(os.path.join(folder, "images", f"{i:08d}.ppm"), "wb") as fh:
                    fh.write(f"P6\n{dec.shape[1]} {dec.shape[0]}\n255\n".encode() + dec.tobytes())
        print(f"case: pyramid pipeline, 5 views 160x120 from {kind} images", flush=True)
        runs[kind] = (folder, P.run_pyramid_pipeline(folder, initial_image_scale=1, size_bound=100,
                                                     precision="fp16", seed=4))
    assert runs["jpg"][1] == runs["ppm"][1] > 4 * 2000
    names = []
    for root, _, files in os.walk(os.path.join(runs["jpg"][0], "HPM_MVS_plusplus")):
        for f in files:
            names.append(os.path.relpath(os.path.join(root, f), runs["jpg"][0]))
    assert len(names) > 5 * 6
    for rel in names:
        assert filecmp.cmp(os.path.join(runs["jpg"][0], rel), os.path.join(runs["ppm"][0], rel), shallow=False), rel
    acc = {}
    for kind in ("pgm", "jpg"):
        acc[kind] = []
        for i in range(5):
            d = P.read_dmb(os.path.join(runs[kind][0], "HPM_MVS_plusplus", f"2333_{i:08d}", "depths_geom.dmb"))
            acc[kind].append(float(np.mean(np.abs(d - sc.depths[i]) / sc.depths[i] <= 0.01)))
        print(f"{kind}: depth within 1% of GT per view", " ".join(f"{a:.3f}" for a in acc[kind]))
    assert np.mean(acc["jpg"]) > np.mean(acc["pgm"]) - 0.05


@pytest.mark.parametrize("schedule", ["pyramid", "single"])
def test_confidence_overlap_gives_the_same_files(tmp_path, schedule):
    """The passes after each confidence evaluation run alongside it, each
    view waiting for its own confidence map; every file equals that of
    running them one after the other."""
    import filecmp
    from hpm_mi355x import pipeline as P
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=8, n_cams=6, width=320, height=240, f=262.5, n_src=4)
    runs = {}
    try:
        for overlap in (True, False):
            folder = str(tmp_path / f"overlap{int(overlap)}")
            scenes.write_dense_folder(sc, folder)
            P.set_confidence_overlap(overlap)
            print(f"case: {schedule} pipeline 6 views 320x240, confidence overlap {overlap}", flush=True)
            if schedule == "pyramid":
                n = P.run_pyramid_pipeline(folder, initial_image_scale=1, size_bound=100, precision="fp16", seed=2)
            else:
                n = P.run_pipeline(folder, precision="fp16", seed=2)
            runs[overlap] = (folder, n)
    finally:
        P.set_confidence_overlap(True)
    assert runs[True][1] == runs[False][1] > 1000
    names = []
    for root, _, files in os.walk(os.path.join(runs[True][0], "HPM_MVS_plusplus")):
        for f in files:
            names.append(os.path.relpath(os.path.join(root, f), runs[True][0]))
    assert len(names) > 6 * 6
    for rel in names:
        assert filecmp.cmp(os.path.join(runs[True][0], rel), os.path.join(runs[False][0], rel), shallow=False), rel
