"""ctypes binding of oracle/lib/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; the product never does.  See oracle_pm.cpp for what each entry
point restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")


class orc_camera(ctypes.Structure):
    _fields_ = [("K", ctypes.c_float * 9), ("R", ctypes.c_float * 9), ("t", ctypes.c_float * 3),
                ("height", ctypes.c_int32), ("width", ctypes.c_int32),
                ("depth_min", ctypes.c_float), ("depth_max", ctypes.c_float)]


class orc_params(ctypes.Structure):
    _fields_ = [("patch_size", ctypes.c_int), ("radius_increment", ctypes.c_int),
                ("sigma_spatial", ctypes.c_float), ("sigma_color", ctypes.c_float), ("top_k", ctypes.c_int),
                ("num_images", ctypes.c_int), ("depth_min", ctypes.c_float), ("depth_max", ctypes.c_float),
                ("scaled_cols", ctypes.c_float), ("scaled_rows", ctypes.c_float),
                ("geom_consistency", ctypes.c_int), ("prior_consistency", ctypes.c_int),
                ("hierarchy", ctypes.c_int), ("upsample", ctypes.c_int), ("mand_consistency", ctypes.c_int),
                ("num_iterations", ctypes.c_int)]


class orc_probe_out(ctypes.Structure):
    _fields_ = [("draws_used", ctypes.c_int), ("positions", ctypes.c_int * 8), ("flags", ctypes.c_int * 8),
                ("rows", (ctypes.c_float * 32) * 8), ("n_costvec", ctypes.c_int), ("costvec_pos", ctypes.c_int * 16),
                ("pdf", ctypes.c_float * 32), ("weights", ctypes.c_float * 32), ("sel", ctypes.c_uint32),
                ("weight_norm", ctypes.c_float), ("fc", ctypes.c_float * 8), ("rfc", ctypes.c_float * 8),
                ("n_cand", ctypes.c_int), ("cand", (ctypes.c_float * 4) * 5), ("cand_depth", ctypes.c_float * 5),
                ("cand_cost", ctypes.c_float * 5)]


COST_FN = ctypes.CFUNCTYPE(ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.POINTER(ctypes.c_float))


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        f32 = ctypes.c_float
        L.orc_expf.argtypes = [f32]; L.orc_expf.restype = f32
        L.orc_set_reverse_sweep.argtypes = [ctypes.c_int]; L.orc_set_reverse_sweep.restype = None
        L.orc_set_threads.argtypes = [ctypes.c_int]; L.orc_set_threads.restype = None
        L.orc_acosf.argtypes = [f32]; L.orc_acosf.restype = f32
        L.orc_reproj_error.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_void_p]
        L.orc_reproj_error.restype = None
        L.orc_support_points.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_prior_level.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p]
        L.orc_prior_plane.argtypes = [ctypes.c_void_p] * 4
        L.orc_prior_plane.restype = None
        L.orc_support_points_scan.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_prior_coarse.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_prior_reload.argtypes = [ctypes.c_void_p] * 5
        L.orc_cost_grad_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_cost_params_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_jbu_depth.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_void_p]
        L.orc_jbu_prior.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_consistency.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.orc_exp.argtypes = [ctypes.c_double]; L.orc_exp.restype = ctypes.c_double
        L.orc_sinf.argtypes = [f32]; L.orc_sinf.restype = f32
        L.orc_cosf.argtypes = [f32]; L.orc_cosf.restype = f32
        L.orc_bilinear.restype = f32
        L.orc_bilinear.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, f32, f32]
        L.orc_bilinear_f64.restype = ctypes.c_double
        L.orc_bilinear_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_uniforms.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p]
        L.orc_philox4x32_10.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_homography_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_depth_from_plane_f64.restype = f32
        L.orc_depth_from_plane_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_distance_to_origin_f64.restype = ctypes.c_double
        L.orc_distance_to_origin_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        L.orc_random_plane_f32.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
                                           f32, f32, ctypes.c_void_p]
        L.orc_probe_pixel.argtypes = ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
                                      + [ctypes.c_void_p] * 10 + [ctypes.c_int] * 3 + [ctypes.c_void_p, ctypes.c_int]
                                      + [COST_FN, COST_FN, ctypes.c_void_p, ctypes.POINTER(orc_probe_out)])
        for n in ("orc_run_patchmatch_f32", "orc_run_patchmatch_f64"):
            getattr(L, n).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_uint64] + [ctypes.c_void_p] * 10 + [ctypes.c_int] * 4
        for n in ("orc_cost_fixed_f32", "orc_cost_fixed_f64"):
            getattr(L, n).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_init_cost_f32.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_geom_cost_f32.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def camera(K, R, t, width, height, depth_min=1.0, depth_max=5.0) -> orc_camera:
    c = orc_camera()
    for i, v in enumerate(np.asarray(K, np.float32).reshape(9)):
        c.K[i] = float(v)
    for i, v in enumerate(np.asarray(R, np.float32).reshape(9)):
        c.R[i] = float(v)
    for i, v in enumerate(np.asarray(t, np.float32).reshape(3)):
        c.t[i] = float(v)
    c.width, c.height, c.depth_min, c.depth_max = int(width), int(height), float(depth_min), float(depth_max)
    return c


def camera_from_scene(cam) -> orc_camera:
    return camera(cam.K, cam.R, cam.t, cam.width, cam.height, cam.depth_min, cam.depth_max)


def params(num_images, cam0: orc_camera, **kw) -> orc_params:
    p = orc_params()
    p.patch_size, p.radius_increment, p.sigma_spatial, p.sigma_color, p.top_k = 11, 2, 5.0, 3.0, 4
    p.num_images = num_images
    # HPM.cpp:685-686, float products
    p.depth_min = float(np.float32(cam0.depth_min) * np.float32(0.6))
    p.depth_max = float(np.float32(cam0.depth_max) * np.float32(1.2))
    p.num_iterations = 3
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _cams_array(cams):
    return (orc_camera * len(cams))(*cams)


def _img_ptrs(images):
    imgs = [np.ascontiguousarray(im, np.float32) for im in images]
    arr = (ctypes.c_void_p * len(imgs))(*[a.ctypes.data for a in imgs])
    return imgs, arr


def _p(a):
    return None if a is None else a.ctypes.data


def run_patchmatch(cams, images, prm: orc_params, seed, dtype=np.float32, planes=None, costs=None, pre_costs=None,
                   selected=None, prior=None, masks=None, conf=None, canny=None, texture=None, scaled=None,
                   depth_images=None, n_halfpasses=-1, do_output=True, rows=(-1, -1)):
    """Runs the oracle's RunPatchMatch.  Buffers default to zeros; returns a dict."""
    W, H = cams[0].width, cams[0].height
    z = lambda shape, dt=dtype: np.zeros(shape, dt)
    planes = z((H, W, 4)) if planes is None else np.ascontiguousarray(planes, dtype)
    costs = z((H, W)) if costs is None else np.ascontiguousarray(costs, dtype)
    pre_costs = z((H, W)) if pre_costs is None else np.ascontiguousarray(pre_costs, dtype)
    selected = z((H, W), np.uint32) if selected is None else np.ascontiguousarray(selected, np.uint32)
    conf = z((H, W)) if conf is None else np.ascontiguousarray(conf, dtype)
    canny = z((H, W), np.uint32) if canny is None else np.ascontiguousarray(canny, np.uint32)
    texture = z((H, W)) if texture is None else np.ascontiguousarray(texture, dtype)
    prior = None if prior is None else np.ascontiguousarray(prior, dtype)
    masks = None if masks is None else np.ascontiguousarray(masks, np.uint32)
    scaled = None if scaled is None else np.ascontiguousarray(scaled, dtype)
    imgs, img_arr = _img_ptrs(images)
    darr = None
    if depth_images is not None:
        dimgs = [np.zeros((1, 1), np.float32) if d is None else np.ascontiguousarray(d, np.float32) for d in depth_images]
        darr = (ctypes.c_void_p * len(dimgs))(*[a.ctypes.data for a in dimgs])
    fn = lib().orc_run_patchmatch_f32 if dtype == np.float32 else lib().orc_run_patchmatch_f64
    rc = fn(_cams_array(cams), len(cams), img_arr, darr, ctypes.byref(prm), ctypes.c_uint64(seed), _p(planes),
            _p(costs), _p(pre_costs), _p(selected), _p(prior), _p(masks), _p(conf), _p(canny), _p(texture),
            _p(scaled), n_halfpasses, int(do_output), rows[0], rows[1])
    if rc != 0:
        raise RuntimeError(f"oracle run failed ({rc})")
    return dict(planes=planes, costs=costs, pre_costs=pre_costs, selected=selected, conf=conf, texture=texture)


def probe_pixel(op, cams, prm, maps, px, py, iter=0, draws=None, ncc=None, geom=None, ref_image=None):
    """One pixel's RandomInitialization (op 0) or red/black update (op 1) in
    the FP32 oracle, on the maps in `maps` (planes (H, W, 4), costs, pre_costs,
    selected, conf, texture, canny, and optionally prior, masks, scaled), which
    are updated in place.  draws: the uniforms the update consumes, in order;
    ncc / geom: python callables (view, px, py, plane[4]) -> cost replacing the
    bilateral NCC / the geometric cost.  Returns the probe record as a dict."""
    W, H = cams[0].width, cams[0].height
    for k, dt, shape in (("planes", np.float32, (H, W, 4)), ("costs", np.float32, (H, W)),
                         ("pre_costs", np.float32, (H, W)), ("selected", np.uint32, (H, W)),
                         ("conf", np.float32, (H, W)), ("texture", np.float32, (H, W)),
                         ("canny", np.uint32, (H, W))):
        if k not in maps:
            maps[k] = np.zeros(shape, dt)
        assert maps[k].dtype == dt and maps[k].shape == shape and maps[k].flags.c_contiguous, k
    for k, dt in (("prior", np.float32), ("masks", np.uint32), ("scaled", np.float32)):
        if k in maps:
            assert maps[k].dtype == dt and maps[k].flags.c_contiguous, k

    def wrap(fn):
        if fn is None:
            return COST_FN()
        return COST_FN(lambda user, v, x, y, pl: float(fn(v, x, y, np.array(pl[:4], np.float32))))
    cn, cg = wrap(ncc), wrap(geom)
    d = None if draws is None else np.ascontiguousarray(draws, np.float32)
    ref = None if ref_image is None else np.ascontiguousarray(ref_image, np.float32)
    out = orc_probe_out()
    rc = lib().orc_probe_pixel(op, _cams_array(cams), len(cams), _p(ref), ctypes.byref(prm), _p(maps["planes"]),
                               _p(maps["costs"]), _p(maps["pre_costs"]), _p(maps["selected"]), _p(maps.get("prior")),
                               _p(maps.get("masks")), _p(maps["conf"]), _p(maps["canny"]), _p(maps["texture"]),
                               _p(maps.get("scaled")), px, py, iter, _p(d), 0 if d is None else d.size, cn, cg, None,
                               ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"orc_probe_pixel failed ({rc})")
    r = {f: getattr(out, f) for f, _ in orc_probe_out._fields_}
    for f in ("positions", "flags", "costvec_pos", "pdf", "weights", "fc", "rfc", "cand_depth", "cand_cost"):
        r[f] = np.array(r[f][:])
    r["rows"] = np.array([row[:] for row in out.rows], np.float32)
    r["cand"] = np.array([c[:] for c in out.cand], np.float32)
    return r


def cost_fixed(cams, images, prm, planes, dtype=np.float32, rows=(-1, -1)):
    W, H = cams[0].width, cams[0].height
    V = len(cams) - 1
    pl = np.ascontiguousarray(planes, dtype)
    out = np.zeros((H, W, V), dtype)
    imgs, arr = _img_ptrs(images)
    fn = lib().orc_cost_fixed_f32 if dtype == np.float32 else lib().orc_cost_fixed_f64
    fn(_cams_array(cams), len(cams), arr, ctypes.byref(prm), _p(pl), _p(out), rows[0], rows[1])
    return out


def init_cost(cams, images, prm, planes):
    W, H = cams[0].width, cams[0].height
    pl = np.ascontiguousarray(planes, np.float32)
    c = np.zeros((H, W), np.float32)
    m = np.zeros((H, W), np.uint32)
    imgs, arr = _img_ptrs(images)
    lib().orc_init_cost_f32(_cams_array(cams), len(cams), arr, ctypes.byref(prm), _p(pl), _p(c), _p(m))
    return c, m


def geom_cost(cams, depth_images, planes):
    W, H = cams[0].width, cams[0].height
    V = len(cams) - 1
    pl = np.ascontiguousarray(planes, np.float32)
    out = np.zeros((H, W, V), np.float32)
    dimgs, arr = _img_ptrs([np.zeros((1, 1)) if d is None else d for d in depth_images])
    lib().orc_geom_cost_f32(_cams_array(cams), len(cams), arr, _p(pl), _p(out))
    return out


def expf(x):
    return lib().orc_expf(float(x))


def uniforms(seed, pixel, stage, n):
    out = np.zeros(n, np.float32)
    lib().orc_uniforms(ctypes.c_uint64(seed), pixel, stage, n, out.ctypes.data)
    return out


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def acosf(x):
    return lib().orc_acosf(float(x))


def reproj_error(c, r, u, v):
    """The fusion / confidence reprojection error (main.cpp:922,1064) per element."""
    c = np.ascontiguousarray(c, np.int32); r = np.ascontiguousarray(r, np.int32)
    u = np.ascontiguousarray(u, np.float32); v = np.ascontiguousarray(v, np.float32)
    out = np.empty(c.shape, np.float32)
    lib().orc_reproj_error(_p(c), _p(r), _p(u), _p(v), c.size, _p(out))
    return out


def jbu_depth(fine, coarse):
    f = np.ascontiguousarray(fine, np.float32)
    c = np.ascontiguousarray(coarse, np.float32)
    out = np.empty_like(f)
    lib().orc_jbu_depth(_p(f), f.shape[1], f.shape[0], _p(c), c.shape[1], c.shape[0], _p(out))
    return out


def jbu_prior(fine, coarse, coarse_normals):
    f = np.ascontiguousarray(fine, np.float32)
    c = np.ascontiguousarray(coarse, np.float32)
    cn = np.zeros(c.shape + (4,), np.float32)
    cn[..., :3] = np.asarray(coarse_normals)[..., :3]
    od = np.empty_like(f)
    on = np.empty(f.shape + (4,), np.float32)
    lib().orc_jbu_prior(_p(f), f.shape[1], f.shape[0], _p(c), _p(cn), c.shape[1], c.shape[0], _p(od), _p(on))
    return od, on


def consistency(mode, cams, depths, normals, sources):
    """RunFusion (mode 0) and RunFusion_Sky_Strict's pass (mode 2) ->
    (pairs[k, 2], points[k, 6]); ConfidenceEvaluation (mode 1) -> list of
    confidence maps."""
    n = len(cams)
    ca = _cams_array(cams)
    d = [np.ascontiguousarray(x, np.float32) for x in depths]
    nm = [np.ascontiguousarray(np.asarray(x, np.float32)[..., :3]) for x in normals]
    dptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in d])
    nptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in nm])
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum([len(s) for s in sources])
    ids = np.array([int(v) for s in sources for v in s] or [0], np.int32)
    count = ctypes.c_int64(0)
    if mode == 1:
        conf = [np.zeros(x.shape, np.float32) for x in d]
        cptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in conf])
        lib().orc_consistency(1, n, ca, dptr, nptr, _p(off), _p(ids), cptr, None, None, 0, ctypes.byref(count))
        return conf
    cap = sum(x.size for x in d)
    pairs = np.zeros((cap, 2), np.int32)
    pts = np.zeros((cap, 6), np.float32)
    lib().orc_consistency(mode, n, ca, dptr, nptr, _p(off), _p(ids), None, _p(pairs), _p(pts), cap,
                          ctypes.byref(count))
    return pairs[:count.value], pts[:count.value]


def cost_grad(cams, images, prm, params, weights):
    """FP64 aggregated cost and its gradient w.r.t. (depth, theta, phi).
    params: (h, w, 3 or 4) float32; weights: (h, w, V) float32."""
    h, w = params.shape[:2]
    pp = np.zeros((h, w, 4), np.float32)
    pp[..., :3] = params[..., :3]
    wt = np.ascontiguousarray(weights, np.float32)
    imgs, arr = _img_ptrs(images)
    cost = np.zeros((h, w), np.float64)
    grad = np.zeros((h, w, 3), np.float64)
    lib().orc_cost_grad_f64(_cams_array(cams), len(cams), arr, ctypes.byref(prm), _p(pp), _p(wt), _p(cost), _p(grad))
    return cost, grad


def cost_params(cams, images, prm, params, weights):
    """FP64 aggregated cost at (depth, theta, phi) given in double."""
    h, w = params.shape[:2]
    pp = np.zeros((h, w, 4), np.float64)
    pp[..., :3] = params[..., :3]
    wt = np.ascontiguousarray(weights, np.float32)
    imgs, arr = _img_ptrs(images)
    cost = np.zeros((h, w), np.float64)
    lib().orc_cost_params_f64(_cams_array(cams), len(cams), arr, ctypes.byref(prm), _p(pp), _p(wt), _p(cost))
    return cost

