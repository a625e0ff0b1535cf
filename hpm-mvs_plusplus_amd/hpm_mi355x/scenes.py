"""Seeded synthetic multi-view scenes: textured planes viewed by pinhole cameras.

The reference ships no data (it reads COLMAP-converted `images/`, `cams/`,
`pair.txt` folders, main.cpp:123-124,173 and HPM.cpp:159-184).  The workloads
here are ray-cast procedurally so that ground-truth depth and normals are known:

* planes with offsets U[lo, hi] m along the scene axis and normals within a
  maximum angle of it, each textured with seeded 4-octave value noise;
* a fronto-parallel background plane at 5 m so that every ray hits a surface;
* cameras on a horizontal arc (or two rings) aimed at the origin;
* 8-bit grey images with additive Gaussian noise.

Camera convention is the reference's: x_cam = R X + t, K the intrinsics
(main.h:39-47); world y points down, the scene axis is +z.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Camera:
    K: np.ndarray  # 3x3
    R: np.ndarray  # 3x3 world->camera
    t: np.ndarray  # 3
    width: int
    height: int
    depth_min: float = 1.0
    depth_max: float = 5.0

    @property
    def center(self) -> np.ndarray:
        return -self.R.T @ self.t


@dataclass
class Plane:
    normal: np.ndarray  # unit, world frame, facing the cameras (n_z < 0)
    point: np.ndarray   # a point on the plane
    tex_seed: int
    e1: np.ndarray = field(default=None)
    e2: np.ndarray = field(default=None)
    half_extent: float = math.inf  # a square patch of this half side around `point` (inf: the whole plane)

    def __post_init__(self):
        a = np.array([1.0, 0.0, 0.0]) if abs(self.normal[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e1 = a - self.normal * (a @ self.normal)
        self.e1 = e1 / np.linalg.norm(e1)
        self.e2 = np.cross(self.normal, self.e1)


    def intersect(self, C, rays):
        """Ray parameter t of C + t * ray (inf where parallel)."""
        denom = rays @ self.normal
        num = (self.point - C) @ self.normal
        with np.errstate(divide="ignore", invalid="ignore"):
            t = num / denom
        if math.isfinite(self.half_extent):
            with np.errstate(invalid="ignore"):
                d = C[None, None, :] + t[..., None] * rays - self.point
                inside = (np.abs(d @ self.e1) <= self.half_extent) & (np.abs(d @ self.e2) <= self.half_extent)
            t = np.where(inside, t, np.inf)
        return t

    def surface(self, X):
        """Texture coordinates (m) and world normals at the points X (n x 3)."""
        d = X - self.point
        return d @ self.e1, d @ self.e2, np.broadcast_to(self.normal, X.shape)


@dataclass
class Sphere:
    center: np.ndarray
    radius: float
    tex_seed: int

    def intersect(self, C, rays):
        """Nearest ray parameter t >= 0 of C + t * ray on the sphere (inf on a miss)."""
        oc = C - self.center
        a = np.einsum("...i,...i->...", rays, rays)
        b = 2.0 * (rays @ oc)
        c = oc @ oc - self.radius ** 2
        disc = b * b - 4 * a * c
        with np.errstate(invalid="ignore", divide="ignore"):
            t = (-b - np.sqrt(disc)) / (2 * a)
        return np.where(disc >= 0, t, np.inf)

    def surface(self, X):
        """Texture coordinates (arc lengths in m) and outward normals."""
        d = (X - self.center) / self.radius
        az = np.arctan2(d[:, 0], -d[:, 2])
        el = np.arcsin(np.clip(d[:, 1], -1.0, 1.0))
        return self.radius * az, self.radius * el, d


@dataclass
class Scene:
    planes: list
    cameras: list
    images: list          # uint8 HxW per camera
    depths: list          # float32 ground-truth depth per camera
    normals: list         # float32 HxWx3 world normals per camera
    pairs: list           # per camera: list of (src_id, score)


def look_at_camera(center, target, f, cx, cy, width, height, down=(0.0, 1.0, 0.0)) -> Camera:
    center = np.asarray(center, dtype=np.float64)
    fwd = np.asarray(target, dtype=np.float64) - center
    fwd /= np.linalg.norm(fwd)
    d = np.asarray(down, dtype=np.float64)
    right = np.cross(d, fwd)
    right /= np.linalg.norm(right)
    down_v = np.cross(fwd, right)
    R = np.stack([right, down_v, fwd])
    t = -R @ center
    K = np.array([[f, 0.0, cx], [0.0, f, cy], [0.0, 0.0, 1.0]])
    return Camera(K=K, R=R, t=t, width=width, height=height)


def arc_cameras(n, radius, arc_deg, f, cx, cy, width, height):
    cams = []
    for k in range(n):
        th = math.radians(-arc_deg / 2 + arc_deg * k / max(n - 1, 1))
        c = radius * np.array([math.sin(th), 0.0, -math.cos(th)])
        cams.append(look_at_camera(c, (0.0, 0.0, 0.0), f, cx, cy, width, height))
    return cams


def ring_cameras(n, radius, elevations_deg, f, cx, cy, width, height, span_deg=60.0):
    cams = []
    per = [n // len(elevations_deg) + (1 if i < n % len(elevations_deg) else 0) for i in range(len(elevations_deg))]
    for el, m in zip(elevations_deg, per):
        e = math.radians(el)
        for k in range(m):
            az = math.radians(-span_deg / 2 + span_deg * k / max(m - 1, 1))
            c = radius * np.array([math.cos(e) * math.sin(az), -math.sin(e), -math.cos(e) * math.cos(az)])
            cams.append(look_at_camera(c, (0.0, 0.0, 0.0), f, cx, cy, width, height))
    return cams


def random_planes(rng, n, axis_origin, offset_lo, offset_hi, max_tilt_deg):
    planes = []
    for _ in range(n):
        off = rng.uniform(offset_lo, offset_hi)
        # normal within max_tilt of -z (facing the cameras), uniform in solid angle
        cos_max = math.cos(math.radians(max_tilt_deg))
        cz = rng.uniform(cos_max, 1.0)
        phi = rng.uniform(0, 2 * math.pi)
        sz = math.sqrt(max(0.0, 1 - cz * cz))
        n_ = np.array([sz * math.cos(phi), sz * math.sin(phi), -cz])
        p = np.asarray(axis_origin, dtype=np.float64) + np.array([0.0, 0.0, off])
        planes.append(Plane(normal=n_, point=p, tex_seed=int(rng.integers(1 << 30))))
    return planes


def value_noise(u, v, seed, octaves=4, base_freq=6.0):
    """4-octave value noise in [0, 1]: smoothstep-interpolated lattice values."""
    tab_rng = np.random.default_rng(seed)
    out = np.zeros_like(u, dtype=np.float64)
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        table = tab_rng.random((256, 256))
        f = base_freq * (2 ** o)
        x = u * f
        y = v * f
        x0 = np.floor(x)
        y0 = np.floor(y)
        fx = x - x0
        fy = y - y0
        sx = fx * fx * (3 - 2 * fx)
        sy = fy * fy * (3 - 2 * fy)
        ix = x0.astype(np.int64) & 255
        iy = y0.astype(np.int64) & 255
        ix1 = (ix + 1) & 255
        iy1 = (iy + 1) & 255
        a = table[iy, ix]
        b = table[iy, ix1]
        c = table[iy1, ix]
        d = table[iy1, ix1]
        val = (a * (1 - sx) + b * sx) * (1 - sy) + (c * (1 - sx) + d * sx) * sy
        out += amp * val
        norm += amp
        amp *= 0.5
    return out / norm


def render(cam: Camera, planes, noise_sigma, rng, rows_per_chunk=256):
    """Ray-cast the nearest surface (planes and spheres) per pixel: uint8
    image, GT depth, GT world normal."""
    H, W = cam.height, cam.width
    img = np.zeros((H, W), np.float64)
    depth = np.zeros((H, W), np.float32)
    normal = np.zeros((H, W, 3), np.float32)
    Kinv = np.linalg.inv(cam.K)
    C = cam.center
    all_planes = list(planes)
    xs = np.arange(W, dtype=np.float64)
    for r0 in range(0, H, rows_per_chunk):
        r1 = min(H, r0 + rows_per_chunk)
        ys = np.arange(r0, r1, dtype=np.float64)
        u, v = np.meshgrid(xs, ys)
        pix = np.stack([u, v, np.ones_like(u)], axis=-1)
        rays_cam = pix @ Kinv.T                 # z = 1 in camera frame
        rays_w = rays_cam @ cam.R               # R^T applied to each row vector
        best_t = np.full(u.shape, np.inf)
        best_k = np.full(u.shape, -1, np.int64)
        for k, pl in enumerate(all_planes):
            t = pl.intersect(C, rays_w)
            ok = (t > 0.5) & (t < 8.0) & np.isfinite(t)
            better = ok & (t < best_t)
            best_t = np.where(better, t, best_t)
            best_k = np.where(better, k, best_k)
        # rays_cam has z=1, so t is the depth along the optical axis
        depth[r0 - 0:r1 - 0] = np.where(np.isfinite(best_t), best_t, 0).astype(np.float32)
        # rays that hit nothing (t = inf) are never read back: give them t = 0
        X = C[None, None, :] + np.where(np.isfinite(best_t), best_t, 0.0)[..., None] * rays_w
        shade = np.zeros(u.shape)
        for k, pl in enumerate(all_planes):
            m = best_k == k
            if not np.any(m):
                continue
            pu, pv, nrm = pl.surface(X[m])
            shade[m] = value_noise(pu, pv, pl.tex_seed)
            normal[r0:r1][m] = nrm.astype(np.float32)
        img[r0:r1] = 20.0 + 215.0 * shade
    img += rng.normal(0.0, noise_sigma, img.shape)
    img8 = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    return img8, depth, normal


def pair_scores(cams, k):
    """Top-k neighbours by a baseline-angle score (peaks at 5 deg, like colmap2mvsnet)."""
    n = len(cams)
    pairs = []
    for i in range(n):
        sc = []
        for j in range(n):
            if j == i:
                continue
            a = cams[i].center
            b = cams[j].center
            ang = math.degrees(math.acos(np.clip((a @ b) / (np.linalg.norm(a) * np.linalg.norm(b)), -1, 1)))
            s = math.exp(-((ang - 5.0) ** 2) / (2 * (1.0 if ang <= 5.0 else 10.0) ** 2))
            sc.append((j, s))
        sc.sort(key=lambda x: -x[1])
        pairs.append(sc[:k])
    return pairs


def make_plane_scene(seed=1, n_planes=8, n_cams=10, width=640, height=480, f=525.0, cx=None, cy=None,
                     arc_deg=30.0, radius=3.0, offset_lo=2.0, offset_hi=4.0, max_tilt_deg=50.0,
                     noise_sigma=2.0, n_src=9, render_views=None) -> Scene:
    cx = width / 2 if cx is None else cx
    cy = height / 2 if cy is None else cy
    rng = np.random.default_rng(seed)
    cams = arc_cameras(n_cams, radius, arc_deg, f, cx, cy, width, height)
    mid = cams[len(cams) // 2].center
    planes = random_planes(rng, n_planes, (0.0, 0.0, mid[2]), offset_lo, offset_hi, max_tilt_deg)
    # fronto-parallel background 5 m in front of the middle camera
    planes.append(Plane(normal=np.array([0.0, 0.0, -1.0]), point=np.array([0.0, 0.0, mid[2] + 5.0]),
                        tex_seed=int(rng.integers(1 << 30))))
    images, depths, normals = [], [], []
    noise_rng = np.random.default_rng(seed + 2)
    for idx, c in enumerate(cams):
        if render_views is not None and idx not in render_views:
            images.append(None); depths.append(None); normals.append(None)
            continue
        im, d, nm = render(c, planes, noise_sigma, noise_rng)
        images.append(im); depths.append(d); normals.append(nm)
    return Scene(planes=planes, cameras=cams, images=images, depths=depths, normals=normals,
                 pairs=pair_scores(cams, n_src))


def _render_job(args):
    cam, planes, noise_sigma, seed = args
    return render(cam, planes, noise_sigma, np.random.default_rng(seed))


def render_all(cams, planes, noise_sigma, seed, workers=None, views=None):
    """Renders every camera, or those in `views` (the others give (None, None,
    None)).  Camera k's pixel noise is seeded by (seed, k), so the result does
    not depend on the number of worker processes or on which views are
    rendered.  The views are rendered on threads (numpy releases the GIL in
    the ray casting), not in forked processes: a caller may already have a
    HIP runtime initialised, which a forked child must not inherit."""
    keep = [k for k in range(len(cams)) if views is None or k in views]
    jobs = [(cams[k], planes, noise_sigma, [seed, k]) for k in keep]
    workers = workers if workers is not None else min(16, os.cpu_count() or 1, max(1, len(jobs)))
    if workers <= 1:
        done = [_render_job(j) for j in jobs]
    else:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(workers) as pool:
            done = list(pool.map(_render_job, jobs))
    out = [(None, None, None)] * len(cams)
    for k, r in zip(keep, done):
        out[k] = r
    return out


def random_patches(rng, n, lo=(-1.6, -1.2, -1.4), hi=(1.6, 1.2, 1.4), ext=(0.4, 1.2), max_tilt_deg=50.0):
    """Square textured plane patches: centres uniform in a box, half sides
    U[ext], normals within max_tilt of -z (facing the cameras)."""
    out = []
    cos_max = math.cos(math.radians(max_tilt_deg))
    for _ in range(n):
        cz = rng.uniform(cos_max, 1.0)
        phi = rng.uniform(0, 2 * math.pi)
        sz = math.sqrt(max(0.0, 1 - cz * cz))
        nrm = np.array([sz * math.cos(phi), sz * math.sin(phi), -cz])
        out.append(Plane(normal=nrm, point=rng.uniform(lo, hi), tex_seed=int(rng.integers(1 << 30)),
                         half_extent=float(rng.uniform(*ext))))
    return out


def random_spheres(rng, n, lo=(-1.0, -0.8, -1.6), hi=(1.0, 0.8, -0.4), r_lo=0.2, r_hi=0.6):
    return [Sphere(center=rng.uniform(lo, hi), radius=float(rng.uniform(r_lo, r_hi)), tex_seed=int(rng.integers(1 << 30)))
            for _ in range(n)]


def make_orbit_scene(seed=3, n_views=30, width=3072, height=2048, n_planes=24, n_spheres=6, radius=4.0,
                     elevations_deg=(15.0, 35.0), span_deg=60.0, n_src=10, noise_sigma=2.0, workers=None,
                     render_refs=None) -> Scene:
    """Full-resolution scene: n_planes textured square plane patches (half
    sides U[0.4, 1.2] m, normals within 50 deg of the scene axis) plus
    n_spheres textured spheres (radius U[0.2, 0.6] m) in front of a background
    plane, seen by n_views
    cameras on two rings (radius 4 m, elevations 15 and 35 deg) over a 60 deg
    arc, focal length 525 px per 640 px of width, each with its n_src
    top-scored sources.  Rays that leave the scene within 8 m (a few per cent
    near the top corners of the high ring) hit nothing: depth 0 there.
    render_refs: render only these reference views and their sources (the
    other views' images, depths and normals are None)."""
    rng = np.random.default_rng(seed)
    f = 525.0 * width / 640.0
    cams = ring_cameras(n_views, radius, elevations_deg, f, width / 2, height / 2, width, height, span_deg=span_deg)
    surfaces = random_patches(rng, n_planes)
    surfaces += random_spheres(rng, n_spheres)
    surfaces.append(Plane(normal=np.array([0.0, 0.0, -1.0]), point=np.array([0.0, 0.0, 2.5]),
                          tex_seed=int(rng.integers(1 << 30))))
    for c in cams:
        c.depth_min, c.depth_max = 1.0, 7.5
    pairs = pair_scores(cams, n_src)
    views = None
    if render_refs is not None:
        views = set(render_refs)
        for r in render_refs:
            views.update(j for j, _ in pairs[r])
    out = render_all(cams, surfaces, noise_sigma, seed + 2, workers, views)
    return Scene(planes=surfaces, cameras=cams, images=[o[0] for o in out], depths=[o[1] for o in out],
                 normals=[o[2] for o in out], pairs=pairs)


def scaled_camera(cam: Camera, width: int, height: int) -> Camera:
    """The camera of an image resized to width x height: K scaled per axis by
    the size ratio, as RescaleImageAndCamera (HPM.cpp:186-207) and
    InuputInitialization (HPM.cpp:668-674) scale it."""
    sx, sy = width / cam.width, height / cam.height
    K = cam.K.copy()
    K[0, 0] *= sx
    K[0, 2] *= sx
    K[1, 1] *= sy
    K[1, 2] *= sy
    return Camera(K=K, R=cam.R, t=cam.t, width=width, height=height, depth_min=cam.depth_min, depth_max=cam.depth_max)


def ground_truth_depth(scene: Scene, cam: Camera):
    """Ray-cast depth of the scene's surfaces for any camera (no texture)."""
    return render(cam, scene.planes, 0.0, np.random.default_rng(0))[1]


def make_batch_scene(seed=1, n_views=64, sizes=((640, 480), (1024, 768), (1600, 1200), (2048, 1536)),
                     n_planes=40, src_range=(6, 10), radius=3.0, elevations_deg=(10.0, 25.0), span_deg=60.0,
                     noise_sigma=2.0, workers=None) -> Scene:
    """Reference views of uneven sizes over one scene: n_views cameras on two
    rings aimed at the origin, each with a size drawn uniformly from `sizes`
    (focal length 525 px per 640 px of width, principal point at the centre)
    and a source count drawn uniformly from src_range (the top-scored
    neighbours).  Scene.pairs[i] holds exactly camera i's sources."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(sizes), n_views)
    n_src = rng.integers(src_range[0], src_range[1] + 1, n_views)
    base = ring_cameras(n_views, radius, elevations_deg, 1.0, 0.0, 0.0, 1, 1, span_deg=span_deg)
    cams = []
    for k, c in enumerate(base):
        w, h = sizes[pick[k]]
        f = 525.0 * w / 640.0
        cams.append(Camera(K=np.array([[f, 0.0, w / 2], [0.0, f, h / 2], [0.0, 0.0, 1.0]]), R=c.R, t=c.t,
                           width=int(w), height=int(h)))
    planes = random_planes(rng, n_planes, (0.0, 0.0, -radius), 2.0, 4.0, 50.0)
    planes.append(Plane(normal=np.array([0.0, 0.0, -1.0]), point=np.array([0.0, 0.0, 2.0]),
                        tex_seed=int(rng.integers(1 << 30))))
    out = render_all(cams, planes, noise_sigma, seed + 2, workers)
    pairs = [p[:int(n)] for p, n in zip(pair_scores(cams, int(src_range[1])), n_src)]
    return Scene(planes=planes, cameras=cams, images=[o[0] for o in out], depths=[o[1] for o in out],
                 normals=[o[2] for o in out], pairs=pairs)


def reference_view_problem(scene: Scene, ref: int, n_src: int):
    """(ref camera, source cameras, images as float32) for one RunPatchMatch."""
    ids = [ref] + [j for j, s in scene.pairs[ref][:n_src] if s > 0]
    cams = [scene.cameras[i] for i in ids]
    imgs = [scene.images[i].astype(np.float32) for i in ids]
    return ids, cams, imgs


def random_planes_per_pixel(cam: Camera, seed, depth_lo=1.0, depth_hi=5.0, max_angle_deg=60.0):
    """Per-pixel random planes in the reference-camera frame (normal within
    max_angle of the reversed view ray), stored as (nx, ny, nz, d) with
    n . X + d = 0 (HPM.cu:135-140)."""
    rng = np.random.default_rng(seed)
    H, W = cam.height, cam.width
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    K = cam.K
    ray = np.stack([(u - K[0, 2]) / K[0, 0], (v - K[1, 2]) / K[1, 1], np.ones_like(u)], -1)
    ray /= np.linalg.norm(ray, axis=-1, keepdims=True)
    depth = rng.uniform(depth_lo, depth_hi, (H, W))
    cos_max = math.cos(math.radians(max_angle_deg))
    cz = rng.uniform(cos_max, 1.0, (H, W))
    phi = rng.uniform(0, 2 * math.pi, (H, W))
    sz = np.sqrt(1 - cz * cz)
    # build a basis around -ray
    a = -ray
    tmp = np.where(np.abs(a[..., :1]) < 0.9, np.array([1.0, 0, 0]), np.array([0, 1.0, 0]))
    e1 = tmp - a * np.sum(tmp * a, -1, keepdims=True)
    e1 /= np.linalg.norm(e1, axis=-1, keepdims=True)
    e2 = np.cross(a, e1)
    n = cz[..., None] * a + sz[..., None] * (np.cos(phi)[..., None] * e1 + np.sin(phi)[..., None] * e2)
    X = np.stack([depth * (u - K[0, 2]) / K[0, 0], depth * (v - K[1, 2]) / K[1, 1], depth], -1)
    d = -np.sum(n * X, -1)
    out = np.concatenate([n, d[..., None]], -1).astype(np.float32)
    return out


# ----------------------------------------------------------------- file formats
def write_dense_folder(scene: Scene, folder: str, depth_lo=1.0, depth_hi=5.0, n_depth=192):
    """images/%08d.pgm, cams/%08d_cam.txt and pair.txt in the layout the
    reference reads (ReadCamera HPM.cpp:159-184, GenerateSampleList
    main.cpp:171-200).  Images are binary PGM; the pipeline also reads
    PPM and JPEG (images/%08d.ppm, images/%08d.jpg)."""
    os.makedirs(os.path.join(folder, "images"), exist_ok=True)
    os.makedirs(os.path.join(folder, "cams"), exist_ok=True)
    for i, (cam, im) in enumerate(zip(scene.cameras, scene.images)):
        with open(os.path.join(folder, "images", f"{i:08d}.pgm"), "wb") as fh:
            fh.write(f"P5\n{cam.width} {cam.height}\n255\n".encode())
            fh.write(np.ascontiguousarray(im).tobytes())
        with open(os.path.join(folder, "cams", f"{i:08d}_cam.txt"), "w") as fh:
            fh.write("extrinsic\n")
            for r in range(3):
                fh.write(" ".join(f"{x:.9g}" for x in cam.R[r]) + f" {cam.t[r]:.9g}\n")
            fh.write("0 0 0 1\n\nintrinsic\n")
            for r in range(3):
                fh.write(" ".join(f"{x:.9g}" for x in cam.K[r]) + "\n")
            fh.write(f"\n{depth_lo} {(depth_hi - depth_lo) / n_depth} {n_depth} {depth_hi}\n")
    with open(os.path.join(folder, "pair.txt"), "w") as fh:
        fh.write(f"{len(scene.cameras)}\n")
        for i, pr in enumerate(scene.pairs):
            fh.write(f"{i}\n{len(pr)} " + " ".join(f"{j} {s:.6f}" for j, s in pr) + "\n")
