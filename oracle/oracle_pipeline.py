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

run_pyramid_pipeline restates the multi-scale schedule (main.cpp:1183-1352)
the same way: ComputeMultiScaleSettings (main.cpp:202-235), the level images
of PrepareScaleImages (main.cpp:107-155) and their cameras
(HPM.cpp:654-676), the photometric pass at the coarsest level and, at the
finer ones, JointBilateralUpsampling of the coarser depths_geom
(main.cpp:636-654, RunJBU HPM.cpp:1260-1320) followed by a hierarchy pass
(HPM.cpp:849-905); then the prior passes of every distance from the
coarsest scale down (main.cpp:1254-1262), each after a ConfidenceEvaluation,
with the coarse planar prior of distances 1 and 2 (main.cpp:443-575); then
the three geometric passes; finally RunFusion.  The image resizing follows
the library's cv::resize stand-in (unpinned against OpenCV, as the edge map
and the triangulation); its seeds are seed + 10007 p + i for the p-th pass
over the views.
"""
from __future__ import annotations

import math

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


# ------------------------------------------------------------ pyramid schedule

def resize_linear(src, dw, dh, round_u8=False):
    """cv::resize(INTER_LINEAR)'s stand-in in the library (half-pixel centres,
    clamped borders; positions in double, weights and interpolation in float;
    8-bit results rounded to nearest, floor(v + 0.5))."""
    sh, sw = src.shape[:2]

    def axis(n, sn):
        s = sn / n
        f = (np.arange(n, dtype=np.float64) + 0.5) * s - 0.5
        i = np.floor(f).astype(np.int64)
        w = (f - i).astype(np.float32)
        lo = i < 0
        i[lo], w[lo] = 0, 0
        hi = i >= sn - 1
        i[hi], w[hi] = sn - 1, 0
        return i, w

    xi, ax = axis(dw, sw)
    yi, ay = axis(dh, sh)
    xb, yb = np.minimum(xi + 1, sw - 1), np.minimum(yi + 1, sh - 1)
    f = src.astype(np.float32)
    a, b = f[yi][:, xi], f[yi][:, xb]
    c, d = f[yb][:, xi], f[yb][:, xb]
    top = a + ax[None, :] * (b - a)
    bot = c + ax[None, :] * (d - c)
    v = (top + ay[:, None] * (bot - top)).astype(np.float32)
    if round_u8:
        v = np.clip(np.floor(v + np.float32(0.5)), 0, 255).astype(np.uint8)
    return np.ascontiguousarray(v)  # fancy indexing may leave the result in column order


def _std_round(x):
    """std::round of a non-negative float (halves away from zero)."""
    return int(math.floor(float(x) + 0.5))


def _level_camera(cam, w, h, ow, oh):
    """The camera of a view prepared at w x h from ow x oh: K scaled by the
    float ratios (HPM.cpp:668-676)."""
    c = O.orc_camera.from_buffer_copy(cam)
    sx = np.float32(w) / np.float32(ow)
    sy = np.float32(h) / np.float32(oh)
    c.K[0] = float(np.float32(c.K[0]) * sx)
    c.K[2] = float(np.float32(c.K[2]) * sx)
    c.K[4] = float(np.float32(c.K[4]) * sy)
    c.K[5] = float(np.float32(c.K[5]) * sy)
    c.width, c.height = int(w), int(h)
    return c


def multiscale_settings(max_sides, initial_image_scale, size_bound):
    """ComputeMultiScaleSettings (main.cpp:202-235): per problem the initial
    max side round(side / scale) (integer, halves up) and the number of
    halvings that bring it to size_bound or below; returns (initial, k)."""
    initial, k = [], []
    for side in max_sides:
        m = max(1, (side + initial_image_scale // 2) // initial_image_scale)
        n, cur = 0, m
        while cur > size_bound:
            cur //= 2
            n += 1
        initial.append(m)
        k.append(n)
    return initial, k


def level_size(w, h, target_max_side):
    """PrepareScaleImages' size for a w x h image at a level of max side
    target_max_side (main.cpp:125-133): unchanged when it fits, otherwise
    scaled by the float min(target / w, target / h), rounded."""
    if max(w, h) <= target_max_side:
        return w, h
    factor = min(np.float32(target_max_side) / np.float32(w), np.float32(target_max_side) / np.float32(h))
    return max(1, _std_round(np.float32(w) * factor)), max(1, _std_round(np.float32(h) * factor))


def prior_distances(max_hpm_scale, level, num_down_after):
    """The hpm_scale_distance of every prior pass of a level, coarsest prior
    first (main.cpp:1254-1262): for hpm_scale = max .. level, distance =
    hpm_scale - num_downscale - 1 with num_downscale already decremented."""
    return [[hs - nd - 1 for nd in num_down_after] for hs in range(max_hpm_scale, level - 1, -1)]


def run_pyramid_pipeline(folder, initial_image_scale, size_bound, seed, canny_fn, delaunay_fn):
    """Returns (files, fused) as run_pipeline, for the multi-scale schedule:
    files hold the maps the last pass of each kind left (the finest level's,
    and texture<k> of every level k)."""
    F = Folder(folder)
    n = len(F.problems)
    index = {ref: i for i, (ref, _) in enumerate(F.problems)}
    files = {}
    initial_max, num_down = multiscale_settings([max(F.gray8[ref].shape) for ref, _ in F.problems],
                                                initial_image_scale, size_bound)
    max_down = max(num_down)
    cur_max = [0] * n
    ids = sorted({v for ref, srcs in F.problems for v in [ref] + srcs})
    passes = [0]

    for level in range(max_down, -1, -1):
        for i in range(n):  # main.cpp:1239-1244
            if num_down[i] >= 0:
                cur_max[i] = int(initial_max[i] / math.pow(2.0, num_down[i]))
                num_down[i] -= 1
        # PrepareScaleImages (main.cpp:107-155): every image takes the size of
        # the problem whose reference it is
        gray8, cams = {}, {}
        for v in ids:
            if v not in index:
                raise ValueError(f"no pyramid size for source image {v}")
            target = cur_max[index[v]]
            img = F.gray8[v]
            oh, ow = img.shape
            nw, nh = level_size(ow, oh, target)
            if (nw, nh) != (ow, oh):
                img = resize_linear(img, nw, nh, round_u8=True)
            gray8[v] = img
            cams[v] = _level_camera(F.cams[v], img.shape[1], img.shape[0], ow, oh)
        edges = {ref: canny_fn(gray8[ref]).astype(np.uint32) for ref, _ in F.problems}

        def process(i, mode, distance=0, multi=False, mand=False):
            ref, srcs = F.problems[i]
            vids = [ref] + srcs
            vc = [cams[v] for v in vids]
            imgs = [gray8[v].astype(np.float32) for v in vids]
            h, w = gray8[ref].shape
            pass_seed = seed + 10007 * passes[0] + i
            geom = mode == "geom"
            prm = O.params(len(vc), vc[0], geom_consistency=int(geom), prior_consistency=int(mode == "prior"),
                           mand_consistency=int(mand), hierarchy=int(mode == "hierarchy"))
            kw = dict(canny=edges[ref])
            if mode == "hierarchy":  # CudaSpaceInitialization, HPM.cpp:849-905
                fine = files[(i, "depths")]
                cn, cc = files[(i, "normals")], files[(i, "costs")]
                ch, cw = cc.shape
                upsample = cw != h or ch != w  # the coarse width is compared with the fine rows (HPM.cpp:869)
                coarse = np.zeros((ch, cw, 4), np.float32)
                coarse[..., :3] = cn
                if upsample:
                    coarse[..., 3] = cc
                else:
                    flat = np.zeros(ch * cw, np.float32)
                    m = min(flat.size, fine.size)
                    flat[:m] = fine.reshape(-1)[:m]
                    coarse[..., 3] = flat.reshape(ch, cw)
                prm.upsample, prm.scaled_cols, prm.scaled_rows = int(upsample), float(cw), float(ch)
                planes0 = np.zeros((h, w, 4), np.float32)
                planes0[..., 3] = fine
                kw["planes"], kw["scaled"] = planes0, coarse
            if geom:
                suffix = "depths_geom" if multi else "depths"
                kw["depth_images"] = [None] + [files[(index[v], suffix)] for v in srcs]
                kw["planes"] = _planes_from_maps(files[(i, suffix)], files[(i, "normals")])
                kw["costs"] = files[(i, "costs")].copy()
            if mand or mode == "prior":
                kw["conf"] = files[(i, "confidence")].copy()
            if mode == "prior":
                if not 0 <= distance <= 2:  # nothing runs; the loaded maps go back unchanged
                    return
                depth, costs, conf = files[(i, "depths")], files[(i, "costs")], files[(i, "confidence")]
                kw["prior"], kw["masks"] = _prior(vc[0], gray8[ref], depth, costs, conf, distance, delaunay_fn)
                kw["planes"] = _planes_from_maps(depth, files[(i, "normals")])
                kw["costs"] = costs.copy()
            r = O.run_patchmatch(vc, imgs, prm, pass_seed, **kw)
            files[(i, "depths_geom" if geom else "depths")] = np.ascontiguousarray(r["planes"][..., 3])
            files[(i, "normals")] = np.ascontiguousarray(r["planes"][..., :3])
            files[(i, "costs")] = np.ascontiguousarray(r["costs"])
            if not mand and mode != "prior":  # ProcessProblem writes the texture map outside those passes
                files[(i, f"texture{level}")] = np.ascontiguousarray(r["texture"])

        def each_view(*args, **kw):
            for i in range(n):
                process(i, *args, **kw)
            passes[0] += 1

        def confidence(geom):
            name = "depths_geom" if geom else "depths"
            vc = [cams[ref] for ref, _ in F.problems]
            maps = O.consistency(1, vc, [files[(i, name)] for i in range(n)], [files[(i, "normals")] for i in range(n)],
                                 [[index[s] for s in srcs] for _, srcs in F.problems])
            for i in range(n):
                files[(i, "confidence")] = maps[i]

        if level != max_down:  # JointBilateralUpsampling, then the hierarchy pass
            for i, (ref, _) in enumerate(F.problems):
                coarse = files[(i, "depths_geom")]
                fine = gray8[ref].astype(np.float32)
                scale = max(fine.shape[0] // coarse.shape[0], fine.shape[1] // coarse.shape[1])
                if scale == 1:  # RunJBU returns without writing (HPM.cpp:1272-1275)
                    continue
                out = O.jbu_depth(fine, coarse)
                bad = np.argwhere(out != out)  # the NaN patch of HPM.cpp:1297-1300
                for (y, x) in bad:
                    out[y, x] = coarse[min(y // 2, coarse.shape[0] - 1), min(x // 2, coarse.shape[1] - 1)]
                files[(i, "depths")] = out
            each_view("hierarchy")
        else:
            each_view("photometric")
        for dist in prior_distances(max_down, level, num_down):  # prior passes, coarsest prior first
            confidence(False)
            for i in range(n):
                process(i, "prior", distance=dist[i])
            passes[0] += 1
        for it in range(3):  # main.cpp:1269-1290
            if it > 0:
                confidence(True)
            each_view("geom", multi=it > 0, mand=it > 0)
    vc = [cams[ref] for ref, _ in F.problems]
    pairs, pts = O.consistency(0, vc, [files[(i, "depths_geom")] for i in range(n)],
                               [files[(i, "normals")] for i in range(n)],
                               [[index[s] for s in srcs] for _, srcs in F.problems])
    return files, (pairs, pts)


def _prior(cam, gray8, depth, costs, conf, distance, delaunay_fn):
    """The prior planes and masks of a prior pass at `distance`
    (main.cpp:324-575): at distance 0 on the level itself, at 1 and 2 on the
    maps resized by 1 / (2 distance), upsampled by the depth+normal JBU and
    reloaded per pixel."""
    h, w = depth.shape
    if distance == 0:
        xy = np.zeros((2 * ((w + 4) // 5) * ((h + 4) // 5), 2), np.int32)
        cnt = ctypes.c_int(0)
        O.lib().orc_support_points(O._p(np.ascontiguousarray(costs)), O._p(np.ascontiguousarray(conf)), w, h,
                                   O._p(xy), len(xy), ctypes.byref(cnt))
        tris = np.ascontiguousarray(delaunay_fn(w, h, xy[:cnt.value]), np.int32)
        prior = np.zeros((h, w, 4), np.float32)
        masks = np.zeros((h, w), np.uint32)
        O.lib().orc_prior_level(ctypes.byref(cam), O._p(np.ascontiguousarray(depth)), O._p(tris), len(tris),
                                O._p(prior), O._p(masks))
        return prior, masks
    factor = np.float32(1.0 / (distance * 2))
    hw, hh = _std_round(np.float32(w) * factor), _std_round(np.float32(h) * factor)
    depth_ds = resize_linear(depth, hw, hh)
    costs_ds = resize_linear(costs, hw, hh)
    conf_ds = resize_linear(conf, hw, hh)
    # GetSupportPoints_Classify_Check scans (int)(width * factor) columns (HPM.cpp:1030-1031)
    sw, sh = min(int(np.float32(w) * factor), hw), min(int(np.float32(h) * factor), hh)
    xy = np.zeros((2 * ((sw + 4) // 5) * ((sh + 4) // 5) + 2, 2), np.int32)
    cnt = ctypes.c_int(0)
    O.lib().orc_support_points_scan(O._p(costs_ds), O._p(conf_ds), hw, sw, sh, O._p(xy), len(xy), ctypes.byref(cnt))
    tris = np.ascontiguousarray(delaunay_fn(hw, hh, xy[:cnt.value]), np.int32)
    prior_d = np.zeros((hh, hw), np.float32)
    prior_n = np.zeros((hh, hw, 4), np.float32)
    O.lib().orc_prior_coarse(ctypes.byref(cam), O._p(depth_ds), hw, hh, O._p(tris), len(tris), float(factor),
                             O._p(prior_d), O._p(prior_n))
    # JointBilateralUpsampling_prior (HPM.cpp:1323-1392) upsamples nothing at an integer scale of 1
    up_d = np.zeros((h, w), np.float32)
    up_n = np.zeros((h, w, 4), np.float32)
    if max(h // hh, w // hw) != 1:
        up_d, up_n = O.jbu_prior(gray8.astype(np.float32), prior_d, prior_n)
    prior = np.zeros((h, w, 4), np.float32)
    masks = np.zeros((h, w), np.uint32)
    O.lib().orc_prior_reload(ctypes.byref(cam), O._p(np.ascontiguousarray(up_d, np.float32)),
                             O._p(np.ascontiguousarray(up_n, np.float32)), O._p(prior), O._p(masks))
    return prior, masks
