#!/usr/bin/env python3
"""PatchMatch throughput on one MI355X: views/s of RunPatchMatch at 1600x1200
with 10 source views (photometric pass: init + 3 red/black iterations + the
plane -> depth/normal output), inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|fp16|bf16]

One "step" is one RunPatchMatch of one reference view.  Prints one JSON line.
The roofline object prices the dominant kernel (the red/black propagation
kernel) by its algorithmic bytes: 181 texels per NCC evaluation (37 reference
texels + 36 bilinear samples x 4 source texels) x bytes per stored texel, plus
16 B plane in and 4 B per view cost out per hypothesis, divided by its average
launch time measured with HIP events on the context stream.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hpm-mvs_plusplus_amd"))

HBM_PEAK_GBS = 8000.0
W_FULL, H_FULL, NSRC = 1600, 1200, 10
SEED = 5


def make_workload(width=W_FULL, height=H_FULL, nsrc=NSRC):
    from hpm_mi355x import scenes
    f = 525.0 * width / 640.0
    sc = scenes.make_plane_scene(seed=1, n_cams=nsrc + 1, width=width, height=height, f=f, n_src=nsrc)
    ref = nsrc // 2
    ids, cams, imgs = scenes.reference_view_problem(sc, ref, nsrc)
    return sc, ids, cams, imgs


def canny_like(img):
    g = np.zeros_like(img, dtype=np.float32)
    g[:, 1:-1] = np.abs(img[:, 2:] - img[:, :-2])
    g[1:-1, :] += np.abs(img[2:, :] - img[:-2, :])
    return (g > 60).astype(np.uint32)


def cpu_baseline(cams, imgs, canny, rows=24):
    """The FP32 oracle on a band of `rows` rows of the same view (all host
    threads up to 16), scaled to views/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    oc = [O.camera_from_scene(c) for c in cams]
    prm = O.params(len(oc), oc[0])
    H = cams[0].height
    y0 = H // 2 - rows // 2
    t0 = time.perf_counter()
    O.run_patchmatch(oc, imgs, prm, SEED, canny=canny, rows=(y0, y0 + rows))
    dt = time.perf_counter() - t0
    return {"value": (rows / H) / dt, "unit": "views/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "port",
            "sample": f"FP32 oracle RunPatchMatch on rows {y0}-{y0 + rows} of the 1600x1200 view "
                      f"(+ init halo), {dt:.2f} s, scaled by {H}/{rows}"}


def calibration_copy(nbytes):
    """One plain device-to-device copy of `nbytes` (the runtime's copy kernel),
    so that profiled HBM counters can be scaled by a transfer of known size."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16", "fp32fast"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=400)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (views are independent; "
              "the multi-GPU design shards reference views, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import hpm_mi355x as Hm
    t_setup = time.perf_counter()
    sc, ids, cams, imgs = make_workload()
    canny = canny_like(imgs[0])
    ctx = Hm.Context(0)
    ctx.set_params(Hm.default_params(precision=args.precision))
    ctx.set_views([Hm.camera_from_scene(c) for c in cams], imgs)
    ctx.set_canny(canny)
    print(f"[bench] workload ready in {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        ctx.run(SEED + i)
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    ctx.set_timing(True)
    prop_ms, launches, ncc_prop, ncc_ref = 0.0, 0, 0, 0
    ctx.synchronize()
    t0 = time.perf_counter()
    per_step = []
    for k in range(args.steps):
        ts = time.perf_counter()
        ctx.run(SEED + 100 + k, sync=False)
        st = ctx.stats()  # synchronises the stream
        per_step.append(time.perf_counter() - ts)
        prop_ms += st.ms_propagate
        launches += st.propagate_launches
        ncc_prop += st.ncc_propagate
        ncc_ref += st.ncc_reference
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    views_s = args.steps / elapsed

    # accuracy of the last run against ground truth
    planes = ctx.planes()
    gt = sc.depths[ids[0]]
    rel = np.abs(planes[..., 3] - gt) / gt
    within1 = float(np.mean(rel <= 0.01))

    bytes_per_texel = {"fp32": 4, "fp32fast": 4, "fp16": 2, "bf16": 2}[args.precision]
    V = NSRC
    per_launch_evals = ncc_prop / max(launches, 1)
    alg_bytes = per_launch_evals * (181 * bytes_per_texel) + (per_launch_evals / V) * (16 + 4 * V)
    avg_launch_s = (prop_ms / max(launches, 1)) / 1e3
    achieved = alg_bytes / avg_launch_s / 1e9

    out = {
        "metric": "views/s @1600x1200, 10 src: plain FP32 vs FP16+batched, at <=1% depth err",
        "value": round(views_s, 4),
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp32fast": "f32"}.get(args.precision, args.precision),
        "data": "synthetic (seeded ray-cast textured planes, 11 cameras on a 30 deg arc)",
        "config": {"workload": "RunPatchMatch photometric, 1600x1200 reference + 10 source views, 3 iterations",
                   "precision": args.precision, "width": W_FULL, "height": H_FULL, "sources": NSRC},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_propagate", "avg_launch_ms": round(avg_launch_s * 1e3, 3),
                     "alg_bytes_per_launch": int(alg_bytes)},
        "mpix_hyp_view_per_s": round(ncc_prop / (prop_ms / 1e3) / 1e6, 1) if prop_ms > 0 else None,
        "mpix_hyp_view_per_s_reference_count": round(ncc_ref / elapsed / 1e6, 1),
        "depth_within_1pct_of_gt": round(within1, 4),
    }
    ctx.close()
    if args.calibrate_copy_mb:
        out["calibration_copy_bytes"] = calibration_copy(args.calibrate_copy_mb << 20)
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle band) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cams, imgs, canny, rows=args.cpu_rows)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
