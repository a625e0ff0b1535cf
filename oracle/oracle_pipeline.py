"""The single-level HPM-MVS++ pipeline on the CPU oracle -- TEST
INFRASTRUCTURE ONLY (tests/ use it as the checker of the library's
pipeline; the product never imports it).

Restates, over a dense folder (images/%08d.pgm, cams/%08d_cam.txt,
pair.txt):
  * GenerateSampleList (main.cpp:171-200): pair.txt, sources with score <= 0
    dropped; ReadCamera (HPM.cpp:159-184); InuputInitialization's grey images
    and depth range (HPM.cpp:644-691);
  * the single-level schedule (main.cpp:1249-1291): photometric pass;
    ConfidenceEvaluation; prior pass at distance 0; three geometric passes,
    the second and third with multi-geometry and mandatory consistency, each
    after a ConfidenceEvaluation of the previous maps, the last two in view
    order (a view reads the depths_geom.dmb its sources wrote earlier in the
    same pass); RunFusion;
  * ProcessProblem per view and pass (main.cpp:237-634): the reloads of the
    geometric pass (HPM.cpp:786-847) and of the prior pass
    (CudaHypothesesReload, HPM.cpp:942-962), the distance-0 planar prior
    (support points, raster, planes, depth test: orc_prior_level), and the
    files each pass writes.
RunPatchMatch, ConfidenceEvaluation and RunFusion are the oracle's own
(oracle_pm.cpp, oracle_host.cpp).

Two inputs come from the library instead, both unpinned against OpenCV: the
edge map (cv::Canny's stand-in) and the Delaunay triangles of the support
points (cv::Subdiv2D's stand-in, whose triangle order decides which triangle
owns a shared raster pixel).  They are passed in as functions.

Seeds: the reference seeds curand from clock64(); the library's counter-based
generator takes seed + i (photometric), seed + 1000 + i (prior) and
seed + 2000 (k + 2) + i (geometric pass k) for view i, and so does this.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

import oracle_py as O

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def _f32(tok: str) -> float:
    """The float `>>` reads from a token (strtof: correctly rounded)."""
    return float(_libc.strtof(tok.encode(), None))


def read_pair(path):
    """GenerateSampleList (main.cpp:171-200): [(ref, [src, ...]), ...]."""
    toks = open(path).read().split()
    n, k = int(toks[0]), 1
    out = []
    for _ in range(n):
        ref = int(toks[k]); m = int(toks[k + 1]); k += 2
        srcs = []
        for _ in range(m):
            sid, score = int(toks[k]), float(toks[k + 1]); k += 2
            if score > 0:
                srcs.append(sid)
        out.append((ref, srcs))
    return out


def read_cam(path, width, height) -> O.orc_camera:
    """ReadCamera (HPM.cpp:159-184): 3 rows of [R | t] after "extrinsic" (the
    4th row skipped), K after "intrinsic", then depth_min interval num
    depth_max; width and height from the image."""
    toks = open(path).read().split()
    i = toks.index("extrinsic") + 1
    R, t = [], []
    for r in range(3):
        row = [_f32(x) for x in toks[i + 4 * r:i + 4 * r + 4]]
        R += row[:3]
        t.append(row[3])
    j = toks.index("intrinsic") + 1
    K = [_f32(x) for x in toks[j:j + 9]]
    dmin, dmax = _f32(toks[j + 9]), _f32(toks[j + 12])
    return O.camera(np.array(K).reshape(3, 3), np.array(R).reshape(3, 3), t, width, height, dmin, dmax)


def read_pgm(path):
    with open(path, "rb") as fh:
        data = fh.read()
    parts, pos = [], 0
    while len(parts) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            pos = data.index(b"\n", pos)
            continue
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        parts.append(data[pos:end]); pos = end
    assert parts[0] == b"P5" and int(parts[3]) == 255
    w, h = int(parts[1]), int(parts[2])
    return np.frombuffer(data[pos + 1:pos + 1 + w * h], np.uint8).reshape(h, w)


class Folder:
    def __init__(self, folder):
        self.problems = read_pair(os.path.join(folder, "pair.txt"))
        self.gray8, self.cams = {}, {}
        for ref, srcs in self.problems:
            for v in [ref] + srcs:
                if v in self.gray8:
                    continue
                g = read_pgm(os.path.join(folder, "images", f"{v:08d}.pgm"))
                self.gray8[v] = g
                self.cams[v] = read_cam(os.path.join(folder, "cams", f"{v:08d}_cam.txt"), g.shape[1], g.shape[0])

    def gray(self, v):  # cvtColor + convertTo(CV_32F): a grey image as float
        return self.gray8[v].astype(np.float32)


def _planes_from_maps(depth, normal):
    """(world normal, depth) per pixel, as CudaHypothesesReload and the
    geometric reload assemble them."""
    p = np.zeros(depth.shape + (4,), np.float32)
    p[..., :3] = normal
    p[..., 3] = depth
    return p


def run_pipeline(folder, seed, canny_fn, delaunay_fn):
    """Returns (files, fused): files[(view, name)] = map as the library's
    pipeline writes it (name: depths, depths_geom, normals, costs,
    confidence, texture0), fused = (image index, pixel) of every fused point
    in order, with its world position (RunFusion's PointCloud)."""
    F = Folder(folder)
    n = len(F.problems)
    files = {}
    edges = {ref: canny_fn(F.gray8[ref]).astype(np.uint32) for ref, _ in F.problems}
    index = {ref: i for i, (ref, _) in enumerate(F.problems)}

    def process(i, geom=False, prior=False, mand=False, multi=False, pass_seed=0):
        ref, srcs = F.problems[i]
        ids = [ref] + srcs
        cams = [F.cams[v] for v in ids]
        imgs = [F.gray(v) for v in ids]
        prm = O.params(len(cams), cams[0], geom_consistency=int(geom), prior_consistency=int(prior),
                       mand_consistency=int(mand))
        kw = dict(canny=edges[ref])
        if geom:
            suffix = "depths_geom" if multi else "depths"
            kw["depth_images"] = [None] + [files[(index[v], suffix)] for v in srcs]
            kw["planes"] = _planes_from_maps(files[(i, suffix)], files[(i, "normals")])
            kw["costs"] = files[(i, "costs")].copy()
        if mand or prior:
            kw["conf"] = files[(i, "confidence")].copy()
        if prior:
            depth, costs = files[(i, "depths")], files[(i, "costs")]
            conf = files[(i, "confidence")]
            h, w = depth.shape
            xy = np.zeros((2 * ((w + 4) // 5) * ((h + 4) // 5), 2), np.int32)
            cnt = ctypes.c_int(0)
            O.lib().orc_support_points(O._p(np.ascontiguousarray(costs)), O._p(np.ascontiguousarray(conf)), w, h,
                                       O._p(xy), len(xy), ctypes.byref(cnt))
            tris = np.ascontiguousarray(delaunay_fn(w, h, xy[:cnt.value]), np.int32)
            prior_planes = np.zeros((h, w, 4), np.float32)
            masks = np.zeros((h, w), np.uint32)
            O.lib().orc_prior_level(ctypes.byref(cams[0]), O._p(np.ascontiguousarray(depth)), O._p(tris), len(tris),
                                    O._p(prior_planes), O._p(masks))
            kw["prior"], kw["masks"] = prior_planes, masks
            kw["planes"] = _planes_from_maps(depth, files[(i, "normals")])
            kw["costs"] = costs.copy()
        r = O.run_patchmatch(cams, imgs, prm, pass_seed, **kw)
        out = "depths_geom" if geom else "depths"
        files[(i, out)] = np.ascontiguousarray(r["planes"][..., 3])
        files[(i, "normals")] = np.ascontiguousarray(r["planes"][..., :3])
        files[(i, "costs")] = np.ascontiguousarray(r["costs"])
        if not mand and not prior:
            files[(i, "texture0")] = np.ascontiguousarray(r["texture"])

    def confidence(geom):
        name = "depths_geom" if geom else "depths"
        cams = [F.cams[ref] for ref, _ in F.problems]
        maps = O.consistency(1, cams, [files[(i, name)] for i in range(n)], [files[(i, "normals")] for i in range(n)],
                             [[index[s] for s in srcs] for _, srcs in F.problems])
        for i in range(n):
            files[(i, "confidence")] = maps[i]

    for i in range(n):
        process(i, pass_seed=seed + i)
    confidence(False)
    for i in range(n):
        process(i, prior=True, pass_seed=seed + 1000 + i)
    for it in range(3):  # main.cpp:1269-1290
        if it > 0:
            confidence(True)
        for i in range(n):
            process(i, geom=True, mand=it > 0, multi=it > 0, pass_seed=seed + 2000 * (it + 2) + i)
    cams = [F.cams[ref] for ref, _ in F.problems]
    pairs, pts = O.consistency(0, cams, [files[(i, "depths_geom")] for i in range(n)],
                               [files[(i, "normals")] for i in range(n)],
                               [[index[s] for s in srcs] for _, srcs in F.problems])
    return files, (pairs, pts)
