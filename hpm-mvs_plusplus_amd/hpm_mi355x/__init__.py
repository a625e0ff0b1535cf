"""Python binding of libhpm_mi355x.so (include/hpm_mi355x.h).

`HPM` mirrors the reference's host class (HPM.h:57-140): the same method names
(InuputInitialization, CudaSpaceInitialization, RunPatchMatch,
GetPlaneHypothesis, GetCost, GetTexture, Set*Params, the prior/JBU entry
points) over the C ABI, with bulk NumPy getters added.  The HIP library is the
only compute path: if it is missing, importing works but every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# HPM_MI355X_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("HPM_MI355X_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libhpm_mi355x.so")

HPM_OK = 0
PREC_FP32_EXACT = 0
PREC_FP16 = 1
PREC_BF16 = 2
PREC_FP32_FAST = 3
PRECISIONS = {"fp32": PREC_FP32_EXACT, "fp16": PREC_FP16, "bf16": PREC_BF16, "fp32fast": PREC_FP32_FAST}


class HpmError(RuntimeError):
    pass


class hpm_camera(ctypes.Structure):
    _fields_ = [("K", ctypes.c_float * 9), ("R", ctypes.c_float * 9), ("t", ctypes.c_float * 3),
                ("height", ctypes.c_int32), ("width", ctypes.c_int32),
                ("depth_min", ctypes.c_float), ("depth_max", ctypes.c_float)]


class hpm_params(ctypes.Structure):
    _fields_ = [("patch_size", ctypes.c_int32), ("radius_increment", ctypes.c_int32),
                ("sigma_spatial", ctypes.c_float), ("sigma_color", ctypes.c_float),
                ("top_k", ctypes.c_int32), ("geom_consistency", ctypes.c_int32),
                ("prior_consistency", ctypes.c_int32), ("hierarchy", ctypes.c_int32),
                ("mand_consistency", ctypes.c_int32), ("num_iterations", ctypes.c_int32),
                ("precision", ctypes.c_int32)]


class hpm_stats(ctypes.Structure):
    _fields_ = [("ncc_init", ctypes.c_uint64), ("ncc_propagate", ctypes.c_uint64),
                ("ncc_reference", ctypes.c_uint64), ("ms_init", ctypes.c_double),
                ("ms_propagate", ctypes.c_double), ("ms_output", ctypes.c_double),
                ("propagate_launches", ctypes.c_int32)]


EXPORTS = [
    "hpm_default_params", "hpm_create", "hpm_destroy", "hpm_last_error", "hpm_set_params", "hpm_set_views",
    "hpm_set_src_depths", "hpm_set_canny", "hpm_set_confidence", "hpm_reload_hypotheses", "hpm_set_prior",
    "hpm_set_coarse", "hpm_run_patchmatch", "hpm_run_patchmatch_partial", "hpm_synchronize", "hpm_get_planes",
    "hpm_get_costs", "hpm_get_texture", "hpm_get_selected_views", "hpm_get_confidence", "hpm_get_size",
    "hpm_get_stats", "hpm_set_timing", "hpm_cost_fixed", "hpm_cost_fixed_device", "hpm_init_cost",
    "hpm_geom_cost", "hpm_jbu_depth", "hpm_jbu_prior", "hpm_consistency_check", "hpm_set_graph",
    "hpm_run_patchmatch_batch", "hpm_pool_create", "hpm_pool_destroy", "hpm_set_views_pooled",
    "hpm_cost_grad", "hpm_cost_grad_device", "hpm_refine_gn", "hpm_refine_random", "hpm_refine_hybrid",
    "hpm_refine_gn_smoothed", "hpm_prior_support_points", "hpm_prior_build", "hpm_get_prior", "hpm_canny", "hpm_get_depth_normals",
    "hpm_reload_hypotheses_maps", "hpm_set_canny_u8", "hpm_consistency_check_cb",
    "hpm_batch_create", "hpm_batch_destroy", "hpm_batch_last_error", "hpm_batch_run", "hpm_batch_set_graph",
    "hpm_batch_set_shape", "hpm_batch_get_shape", "hpm_batch_last_ms", "hpm_batch_autotune",
    "hpm_batch_set_workgroup", "hpm_batch_get_workgroup",
]

_lib = None


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HpmError(f"{LIB_PATH} is missing: run `python __graft_entry__.py build` (or build.py)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    up = ctypes.POINTER(ctypes.c_uint32)
    L.hpm_default_params.argtypes = [ctypes.POINTER(hpm_params)]
    L.hpm_default_params.restype = None
    L.hpm_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
    L.hpm_destroy.argtypes = [P]
    L.hpm_destroy.restype = None
    L.hpm_last_error.argtypes = [P]
    L.hpm_last_error.restype = ctypes.c_char_p
    L.hpm_set_params.argtypes = [P, ctypes.POINTER(hpm_params)]
    L.hpm_set_views.argtypes = [P, ctypes.c_int, ctypes.POINTER(hpm_camera), ctypes.POINTER(fp)]
    L.hpm_set_src_depths.argtypes = [P, ctypes.POINTER(fp)]
    L.hpm_set_canny.argtypes = [P, up]
    L.hpm_set_confidence.argtypes = [P, fp]
    L.hpm_reload_hypotheses.argtypes = [P, fp, fp]
    L.hpm_set_prior.argtypes = [P, fp, up]
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.hpm_prior_support_points.argtypes = [P, ctypes.c_int, i32p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.hpm_prior_build.argtypes = [P, i32p, ctypes.c_int]
    L.hpm_get_prior.argtypes = [P, fp, up]
    L.hpm_get_depth_normals.argtypes = [P, fp, fp]
    L.hpm_reload_hypotheses_maps.argtypes = [P, fp, fp, fp]
    L.hpm_set_canny_u8.argtypes = [P, ctypes.POINTER(ctypes.c_uint8)]
    u8p = ctypes.POINTER(ctypes.c_uint8)
    L.hpm_canny.argtypes = [ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                            ctypes.c_float, u8p]
    L.hpm_set_coarse.argtypes = [P, fp, ctypes.c_int, ctypes.c_int, fp]
    L.hpm_run_patchmatch.argtypes = [P, ctypes.c_uint64]
    L.hpm_run_patchmatch_partial.argtypes = [P, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    L.hpm_synchronize.argtypes = [P]
    L.hpm_set_graph.argtypes = [P, ctypes.c_int]
    L.hpm_cost_grad.argtypes = [P, fp, fp, fp, fp, fp]
    L.hpm_cost_grad_device.argtypes = [P, P, P, P, P, P]
    L.hpm_refine_gn.argtypes = [P, fp, fp, ctypes.c_int, fp]
    L.hpm_refine_random.argtypes = [P, fp, fp, ctypes.c_int, ctypes.c_uint64, fp]
    L.hpm_refine_hybrid.argtypes = [P, fp, fp, ctypes.c_int, ctypes.c_uint64, fp]
    L.hpm_refine_gn_smoothed.argtypes = [P, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_float, fp]
    L.hpm_pool_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_int, ctypes.POINTER(hpm_camera),
                                  ctypes.POINTER(fp)]
    L.hpm_pool_destroy.argtypes = [P]
    L.hpm_pool_destroy.restype = None
    L.hpm_set_views_pooled.argtypes = [P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
    L.hpm_run_patchmatch_batch.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    L.hpm_batch_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(P), ctypes.c_int]
    L.hpm_batch_destroy.argtypes = [P]
    L.hpm_batch_destroy.restype = None
    L.hpm_batch_last_error.argtypes = [P]
    L.hpm_batch_last_error.restype = ctypes.c_char_p
    L.hpm_batch_run.argtypes = [P, ctypes.POINTER(ctypes.c_uint64)]
    L.hpm_batch_set_graph.argtypes = [P, ctypes.c_int]
    L.hpm_batch_set_shape.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.hpm_batch_get_shape.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.hpm_batch_set_workgroup.argtypes = [P, ctypes.c_int]
    L.hpm_batch_get_workgroup.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    L.hpm_batch_last_ms.argtypes = [P]
    L.hpm_batch_last_ms.restype = ctypes.c_float
    L.hpm_batch_autotune.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), fp]
    for n in ("hpm_get_planes", "hpm_get_costs", "hpm_get_texture", "hpm_get_confidence"):
        getattr(L, n).argtypes = [P, fp]
    L.hpm_get_selected_views.argtypes = [P, up]
    L.hpm_get_size.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.hpm_get_stats.argtypes = [P, ctypes.POINTER(hpm_stats)]
    L.hpm_set_timing.argtypes = [P, ctypes.c_int]
    L.hpm_cost_fixed.argtypes = [P, fp, fp, ctypes.c_int]
    L.hpm_cost_fixed_device.argtypes = [P, P, P, ctypes.c_int]
    L.hpm_init_cost.argtypes = [P, fp, fp, up]
    L.hpm_geom_cost.argtypes = [P, fp, fp]
    L.hpm_jbu_depth.argtypes = [fp, ctypes.c_int, ctypes.c_int, fp, ctypes.c_int, ctypes.c_int, fp]
    L.hpm_jbu_prior.argtypes = [fp, ctypes.c_int, ctypes.c_int, fp, fp, ctypes.c_int, ctypes.c_int, fp, fp]
    ip = ctypes.POINTER(ctypes.c_int32)
    L.hpm_consistency_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(hpm_camera), P, P, ip, ip, P, ip,
                                        fp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    _lib = L
    return L


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def make_camera(K, R, t, width, height, depth_min=1.0, depth_max=5.0) -> hpm_camera:
    c = hpm_camera()
    K = np.asarray(K, np.float32).reshape(9)
    R = np.asarray(R, np.float32).reshape(9)
    t = np.asarray(t, np.float32).reshape(3)
    for i in range(9):
        c.K[i] = float(K[i])
        c.R[i] = float(R[i])
    for i in range(3):
        c.t[i] = float(t[i])
    c.width, c.height = int(width), int(height)
    c.depth_min, c.depth_max = float(depth_min), float(depth_max)
    return c


def camera_from_scene(cam) -> hpm_camera:
    return make_camera(cam.K, cam.R, cam.t, cam.width, cam.height, cam.depth_min, cam.depth_max)


def default_params(**kw) -> hpm_params:
    p = hpm_params()
    lib().hpm_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        setattr(p, k, v)
    return p


class Context:
    """Thin RAII wrapper of one hpm_ctx (one device, one stream)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.hpm_create(ctypes.byref(h), device)
        if rc != HPM_OK:
            raise HpmError(f"hpm_create failed ({rc}): no usable HIP device {device}")
        self._h = h
        self._L = L
        self.width = self.height = self.num_images = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.hpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != HPM_OK:
            msg = self._L.hpm_last_error(self._h)
            raise HpmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    @property
    def handle(self):
        return self._h

    def set_params(self, params: hpm_params):
        self._check(self._L.hpm_set_params(self._h, ctypes.byref(params)), "hpm_set_params")

    def set_views(self, cameras, images):
        n = len(cameras)
        cams = (hpm_camera * n)(*cameras)
        imgs = [np.ascontiguousarray(im, dtype=np.float32) for im in images]
        self._keep = imgs
        arr = (ctypes.POINTER(ctypes.c_float) * n)(*[_fptr(a) for a in imgs])
        self._check(self._L.hpm_set_views(self._h, n, cams, arr), "hpm_set_views")
        self.width, self.height, self.num_images = cameras[0].width, cameras[0].height, n

    def set_src_depths(self, depths):
        n = self.num_images
        ds = [None if d is None else np.ascontiguousarray(d, dtype=np.float32) for d in depths]
        self._keep_d = ds
        arr = (ctypes.POINTER(ctypes.c_float) * n)(*[(_fptr(a) if a is not None else ctypes.POINTER(ctypes.c_float)()) for a in ds])
        self._check(self._L.hpm_set_src_depths(self._h, arr), "hpm_set_src_depths")

    def set_canny(self, canny):
        a = np.ascontiguousarray(canny, dtype=np.uint32)
        self._check(self._L.hpm_set_canny(self._h, _uptr(a)), "hpm_set_canny")

    def set_canny_u8(self, canny):
        a = np.ascontiguousarray(canny, dtype=np.uint8)
        self._check(self._L.hpm_set_canny_u8(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
                    "hpm_set_canny_u8")

    def set_confidence(self, conf):
        a = np.ascontiguousarray(conf, dtype=np.float32)
        self._check(self._L.hpm_set_confidence(self._h, _fptr(a)), "hpm_set_confidence")

    def reload_hypotheses(self, planes, costs):
        p = np.ascontiguousarray(planes, dtype=np.float32)
        c = np.ascontiguousarray(costs, dtype=np.float32)
        self._check(self._L.hpm_reload_hypotheses(self._h, _fptr(p), _fptr(c)), "hpm_reload_hypotheses")

    def reload_maps(self, depth, normals, costs):
        """hpm_reload_hypotheses from a depth map, a (h, w, 3) normal map and costs."""
        d = np.ascontiguousarray(depth, dtype=np.float32)
        n = np.ascontiguousarray(normals, dtype=np.float32)
        c = np.ascontiguousarray(costs, dtype=np.float32)
        self._check(self._L.hpm_reload_hypotheses_maps(self._h, _fptr(d), _fptr(n), _fptr(c)),
                    "hpm_reload_hypotheses_maps")

    def set_prior(self, prior_planes, masks):
        p = np.ascontiguousarray(prior_planes, dtype=np.float32)
        m = np.ascontiguousarray(masks, dtype=np.uint32)
        self._check(self._L.hpm_set_prior(self._h, _fptr(p), _uptr(m)), "hpm_set_prior")

    def prior_support_points(self, distance=0):
        """Support points of the planar prior at `distance` 0, 1 or 2, built
        on the device from the reloaded hypotheses, costs and confidences:
        (n, 2) int32 (x, y) in the reference's order."""
        cap = 2 * ((self.width + 4) // 5) * ((self.height + 4) // 5)
        out = np.empty((cap, 2), np.int32)
        n = ctypes.c_int()
        self._check(self._L.hpm_prior_support_points(self._h, int(distance),
                                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap,
                                                     ctypes.byref(n)), "hpm_prior_support_points")
        return out[:n.value].copy()

    def prior_build(self, triangles):
        """The rest of the planar prior from the Delaunay triangles of the
        last support points ((n, 6) int32 corners)."""
        t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 6)
        self._check(self._L.hpm_prior_build(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(t)),
                    "hpm_prior_build")

    def prior(self):
        """The context's prior planes (h, w, 4) and triangle masks (h, w)."""
        p = np.empty((self.height, self.width, 4), np.float32)
        m = np.empty((self.height, self.width), np.uint32)
        self._check(self._L.hpm_get_prior(self._h, _fptr(p), _uptr(m)), "hpm_get_prior")
        return p, m

    def set_coarse(self, coarse_planes, fine_depth):
        cp = np.ascontiguousarray(coarse_planes, dtype=np.float32)
        fd = np.ascontiguousarray(fine_depth, dtype=np.float32)
        ch, cw = cp.shape[0], cp.shape[1]
        self._check(self._L.hpm_set_coarse(self._h, _fptr(cp), cw, ch, _fptr(fd)), "hpm_set_coarse")

    def run(self, seed: int, n_halfpasses: int = -1, do_output: bool = True, sync: bool = True):
        self._check(self._L.hpm_run_patchmatch_partial(self._h, ctypes.c_uint64(seed), n_halfpasses, int(do_output)),
                    "hpm_run_patchmatch")
        if sync:
            self.synchronize()

    def synchronize(self):
        self._check(self._L.hpm_synchronize(self._h), "hpm_synchronize")

    def set_views_pooled(self, pool, ids):
        """hpm_set_views on images of an ImagePool (ids[0] = reference)."""
        arr = (ctypes.c_int32 * len(ids))(*[int(i) for i in ids])
        self._check(self._L.hpm_set_views_pooled(self._h, pool.handle, len(ids), arr), "hpm_set_views_pooled")
        self._pool = pool  # keeps the pool alive while this context uses it
        self.width, self.height = pool.sizes[ids[0]]
        self.num_images = len(ids)

    def cost_grad(self, params, weights, gn=False):
        """hpm_cost_grad: (cost (h, w), grad (h, w, 3)[, gn (h, w, 6)])."""
        h, w = self.height, self.width
        pp = np.zeros((h, w, 4), np.float32)
        pp[..., :3] = np.asarray(params, np.float32)[..., :3]
        wt = np.ascontiguousarray(weights, np.float32)
        cost = np.empty((h, w), np.float32)
        grad = np.empty((h, w, 3), np.float32)
        g6 = np.empty((h, w, 6), np.float32) if gn else None
        self._check(self._L.hpm_cost_grad(self._h, _fptr(pp), _fptr(wt), _fptr(cost), _fptr(grad),
                                          _fptr(g6) if gn else None), "hpm_cost_grad")
        return (cost, grad, g6) if gn else (cost, grad)

    def cost_grad_device(self, params_ptr, weights_ptr, cost_ptr, grad_ptr, gn_ptr=None):
        """hpm_cost_grad_device on raw device pointers (ints); asynchronous."""
        self._check(self._L.hpm_cost_grad_device(self._h, params_ptr, weights_ptr, cost_ptr, grad_ptr, gn_ptr),
                    "hpm_cost_grad_device")

    def refine(self, params, weights, iterations, method="gn", seed=1, smoothed_iterations=None, sigma=2.0):
        """hpm_refine_gn / _random / _hybrid / _gn_smoothed: returns (params (h, w, 4), cost (h, w))."""
        h, w = self.height, self.width
        pp = np.zeros((h, w, 4), np.float32)
        pp[..., :3] = np.asarray(params, np.float32)[..., :3]
        wt = np.ascontiguousarray(weights, np.float32)
        cost = np.empty((h, w), np.float32)
        if method == "gn":
            rc = self._L.hpm_refine_gn(self._h, _fptr(pp), _fptr(wt), iterations, _fptr(cost))
        elif method == "random":
            rc = self._L.hpm_refine_random(self._h, _fptr(pp), _fptr(wt), iterations, ctypes.c_uint64(seed),
                                           _fptr(cost))
        elif method == "hybrid":
            rc = self._L.hpm_refine_hybrid(self._h, _fptr(pp), _fptr(wt), iterations, ctypes.c_uint64(seed),
                                           _fptr(cost))
        elif method == "gn_smoothed":  # the first half of the steps (by default) on the smoothed views
            k = (iterations + 1) // 2 if smoothed_iterations is None else smoothed_iterations
            rc = self._L.hpm_refine_gn_smoothed(self._h, _fptr(pp), _fptr(wt), iterations, k, sigma, _fptr(cost))
        else:
            raise ValueError(f"unknown refinement method {method!r}")
        self._check(rc, "hpm_refine_" + method)
        return pp, cost

    def set_graph(self, on=True):
        self._check(self._L.hpm_set_graph(self._h, int(on)), "hpm_set_graph")

    def set_timing(self, on=True):
        self._check(self._L.hpm_set_timing(self._h, int(on)), "hpm_set_timing")

    def stats(self) -> hpm_stats:
        s = hpm_stats()
        self._check(self._L.hpm_get_stats(self._h, ctypes.byref(s)), "hpm_get_stats")
        return s

    def planes(self):
        out = np.empty((self.height, self.width, 4), np.float32)
        self._check(self._L.hpm_get_planes(self._h, _fptr(out)), "hpm_get_planes")
        return out

    def costs(self):
        out = np.empty((self.height, self.width), np.float32)
        self._check(self._L.hpm_get_costs(self._h, _fptr(out)), "hpm_get_costs")
        return out

    def depth_normals(self):
        """The result as a depth map (h, w) and a world normal map (h, w, 3)."""
        d = np.empty((self.height, self.width), np.float32)
        n = np.empty((self.height, self.width, 3), np.float32)
        self._check(self._L.hpm_get_depth_normals(self._h, _fptr(d), _fptr(n)), "hpm_get_depth_normals")
        return d, n

    def texture(self):
        out = np.empty((self.height, self.width), np.float32)
        self._check(self._L.hpm_get_texture(self._h, _fptr(out)), "hpm_get_texture")
        return out

    def confidence(self):
        out = np.empty((self.height, self.width), np.float32)
        self._check(self._L.hpm_get_confidence(self._h, _fptr(out)), "hpm_get_confidence")
        return out

    def selected_views(self):
        out = np.empty((self.height, self.width), np.uint32)
        self._check(self._L.hpm_get_selected_views(self._h, _uptr(out)), "hpm_get_selected_views")
        return out

    def set_u8_pairs_min_bytes(self, min_bytes: int):
        """FP16 storage on 8-bit column pairs from this FP16 pair footprint on
        (-1 never, 0 always when the images are 8-bit); results are
        bit-identical either way."""
        self._check(self._L.hpm_set_u8_pairs_min_bytes(self._h, ctypes.c_longlong(int(min_bytes))),
                    "hpm_set_u8_pairs_min_bytes")

    def cost_fixed(self, planes, precision=PREC_FP32_EXACT):
        p = np.ascontiguousarray(planes, dtype=np.float32)
        out = np.empty((self.height, self.width, self.num_images - 1), np.float32)
        self._check(self._L.hpm_cost_fixed(self._h, _fptr(p), _fptr(out), int(precision)), "hpm_cost_fixed")
        return out

    def cost_fixed_device(self, planes_ptr: int, out_ptr: int, precision=PREC_FP32_EXACT):
        self._check(self._L.hpm_cost_fixed_device(self._h, ctypes.c_void_p(planes_ptr), ctypes.c_void_p(out_ptr),
                                                  int(precision)), "hpm_cost_fixed_device")

    def init_cost(self, planes):
        p = np.ascontiguousarray(planes, dtype=np.float32)
        c = np.empty((self.height, self.width), np.float32)
        m = np.empty((self.height, self.width), np.uint32)
        self._check(self._L.hpm_init_cost(self._h, _fptr(p), _fptr(c), _uptr(m)), "hpm_init_cost")
        return c, m

    def geom_cost(self, planes):
        p = np.ascontiguousarray(planes, dtype=np.float32)
        out = np.empty((self.height, self.width, self.num_images - 1), np.float32)
        self._check(self._L.hpm_geom_cost(self._h, _fptr(p), _fptr(out)), "hpm_geom_cost")
        return out


def jbu_depth(fine_gray, coarse_depth):
    f = np.ascontiguousarray(fine_gray, np.float32)
    c = np.ascontiguousarray(coarse_depth, np.float32)
    out = np.empty_like(f)
    rc = lib().hpm_jbu_depth(_fptr(f), f.shape[1], f.shape[0], _fptr(c), c.shape[1], c.shape[0], _fptr(out))
    if rc != HPM_OK:
        raise HpmError(f"hpm_jbu_depth failed ({rc})")
    return out


def jbu_prior(fine_gray, coarse_depth, coarse_normals):
    f = np.ascontiguousarray(fine_gray, np.float32)
    c = np.ascontiguousarray(coarse_depth, np.float32)
    cn = np.zeros(c.shape + (4,), np.float32)
    cn[..., :3] = coarse_normals[..., :3]
    od = np.empty_like(f)
    on = np.empty(f.shape + (4,), np.float32)
    rc = lib().hpm_jbu_prior(_fptr(f), f.shape[1], f.shape[0], _fptr(c), _fptr(cn), c.shape[1], c.shape[0],
                             _fptr(od), _fptr(on))
    if rc != HPM_OK:
        raise HpmError(f"hpm_jbu_prior failed ({rc})")
    return od, on


CHECK_FUSION = 0
CHECK_CONFIDENCE = 1
CHECK_FUSION_MASK = 2


def canny(image, low=50.0, high=150.0, device=0):
    """Edge map (255 = edge) of an (h, w) grey or (h, w, 3) R G B uint8 image
    computed on the device (hpm_canny)."""
    g = np.ascontiguousarray(image, np.uint8)
    channels = 1 if g.ndim == 2 else g.shape[2]
    out = np.zeros(g.shape[:2], np.uint8)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    rc = lib().hpm_canny(device, g.ctypes.data_as(u8p), g.shape[1], g.shape[0], channels, low, high,
                         out.ctypes.data_as(u8p))
    if rc != HPM_OK:
        raise HpmError(f"hpm_canny failed ({rc})")
    return out


def _pack_views(cameras, depths, normals, sources):
    n = len(cameras)
    cams = (hpm_camera * n)(*cameras)
    d = [np.ascontiguousarray(x, np.float32) for x in depths]
    nm = [np.ascontiguousarray(np.asarray(x, np.float32)[..., :3]) for x in normals]
    dptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in d])
    nptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in nm])
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum([len(s) for s in sources])
    ids = np.array([int(v) for s in sources for v in s] or [0], np.int32)
    return cams, d, nm, dptr, nptr, off, ids


def consistency_check(mode, cameras, depths, normals, sources, capacity=None):
    """RunFusion (mode CHECK_FUSION) / ConfidenceEvaluation (CHECK_CONFIDENCE).
    Returns (pairs[k, 2], points[k, 6]) for fusion, a list of confidence maps
    for confidence evaluation."""
    cams, d, nm, dptr, nptr, off, ids = _pack_views(cameras, depths, normals, sources)
    n = len(cameras)
    count = ctypes.c_int64(0)
    L = lib()
    if mode == CHECK_CONFIDENCE:
        conf = [np.zeros(x.shape, np.float32) for x in d]
        cptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in conf])
        rc = L.hpm_consistency_check(mode, n, cams, dptr, nptr, _iptr(off), _iptr(ids), cptr, None, None, 0,
                                     ctypes.byref(count))
        if rc != HPM_OK:
            raise HpmError(f"hpm_consistency_check failed ({rc})")
        return conf
    cap = int(capacity if capacity is not None else sum(x.size for x in d))
    pairs = np.zeros((max(cap, 1), 2), np.int32)
    pts = np.zeros((max(cap, 1), 6), np.float32)
    rc = L.hpm_consistency_check(mode, n, cams, dptr, nptr, _iptr(off), _iptr(ids), None, _iptr(pairs), _fptr(pts),
                                 cap, ctypes.byref(count))
    if rc != HPM_OK:
        raise HpmError(f"hpm_consistency_check failed ({rc})")
    k = min(count.value, cap)
    return pairs[:k], pts[:k]


def run_batch(contexts, seeds):
    """hpm_run_patchmatch_batch: RunPatchMatch on several contexts at once."""
    n = len(contexts)
    hs = (ctypes.c_void_p * n)(*[c.handle for c in contexts])
    sd = (ctypes.c_uint64 * n)(*[int(s) for s in seeds])
    rc = lib().hpm_run_patchmatch_batch(hs, n, sd)
    if rc != HPM_OK:
        msgs = [lib().hpm_last_error(c.handle) for c in contexts]
        raise HpmError(f"hpm_run_patchmatch_batch failed ({rc}): " + "; ".join(m.decode() for m in msgs if m))


class Batch:
    """hpm_batch: the RunPatchMatch schedules of several contexts fused into
    one launch per stage (views of any size; one precision)."""

    def __init__(self, contexts):
        L = lib()
        n = len(contexts)
        self._ctxs = list(contexts)  # the batch holds their handles
        hs = (ctypes.c_void_p * n)(*[c.handle for c in contexts])
        h = ctypes.c_void_p()
        rc = L.hpm_batch_create(ctypes.byref(h), hs, n)
        if rc != HPM_OK:
            raise HpmError(f"hpm_batch_create failed ({rc})")
        self._h = h

    def _check(self, rc, what):
        if rc != HPM_OK:
            raise HpmError(f"{what} failed ({rc}): {lib().hpm_batch_last_error(self._h).decode()}")

    def _seeds(self, seeds):
        if len(seeds) != len(self._ctxs):
            raise ValueError("one seed per context")
        return (ctypes.c_uint64 * len(seeds))(*[int(s) for s in seeds])

    def run(self, seeds):
        self._check(lib().hpm_batch_run(self._h, self._seeds(seeds)), "hpm_batch_run")

    def set_graph(self, on=True):
        self._check(lib().hpm_batch_set_graph(self._h, int(on)), "hpm_batch_set_graph")

    def set_shape(self, tile_w=32, strip=2):
        self._check(lib().hpm_batch_set_shape(self._h, int(tile_w), int(strip)), "hpm_batch_set_shape")

    def shape(self):
        w, s = ctypes.c_int(), ctypes.c_int()
        self._check(lib().hpm_batch_get_shape(self._h, ctypes.byref(w), ctypes.byref(s)), "hpm_batch_get_shape")
        return w.value, s.value

    def set_workgroup(self, threads=256):
        """Pixels (threads) per colour-pass workgroup: 256 or 128."""
        self._check(lib().hpm_batch_set_workgroup(self._h, int(threads)), "hpm_batch_set_workgroup")

    def workgroup(self):
        t = ctypes.c_int()
        self._check(lib().hpm_batch_get_workgroup(self._h, ctypes.byref(t)), "hpm_batch_get_workgroup")
        return t.value

    def last_ms(self):
        return float(lib().hpm_batch_last_ms(self._h))

    def autotune(self, seeds):
        """Times every launch shape, keeps the fastest; returns (tile_w, strip, ms)
        (the chosen pixels per workgroup: workgroup())."""
        ms = ctypes.c_float()
        self._check(lib().hpm_batch_autotune(self._h, self._seeds(seeds), ctypes.byref(ms)), "hpm_batch_autotune")
        return self.shape() + (ms.value,)

    def close(self):
        if getattr(self, "_h", None):
            lib().hpm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ImagePool:
    """hpm_pool: a scene's images on one device, shared by several contexts."""

    def __init__(self, cameras, images, device=0):
        L = lib()
        n = len(cameras)
        cams = (hpm_camera * n)(*cameras)
        self._imgs = [np.ascontiguousarray(i, np.float32) for i in images]
        ptrs = (ctypes.POINTER(ctypes.c_float) * n)(*[_fptr(i) for i in self._imgs])
        h = ctypes.c_void_p()
        rc = L.hpm_pool_create(ctypes.byref(h), device, n, cams, ptrs)
        if rc != HPM_OK:
            raise HpmError(f"hpm_pool_create failed ({rc})")
        self._h = h
        self.sizes = [(c.width, c.height) for c in cameras]
        self._imgs = None

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().hpm_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

