"""Batched RunPatchMatch and hipGraph replay (needs an MI355X).

* hipGraph replay: a context with graphs enabled gives, run after run, the
  same bits as a context launching the kernels directly -- with other work on
  the GPU between replays, after a mode/precision change (recapture), and on
  both the exact and the mixed-precision engine.
* hpm_run_patchmatch_batch on contexts of different sizes and view counts
  gives the same bits as running each context alone.
* The fused batch (hpm_batch: one launch per stage over all views) gives the
  same bits, counters and texture maps as single runs -- for every launch
  shape, replayed from its hipGraph, in geometric-consistency mode, with a
  view beyond the LDS row limit (25 sources), and after autotuning.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]


def _ctx(cams, imgs, graph=False, **params):
    import hpm_mi355x as H
    ctx = H.Context(0)
    ctx.set_params(H.default_params(**params))
    ctx.set_views([H.camera_from_scene(c) for c in cams], imgs)
    ctx.set_graph(graph)
    return ctx


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_graph_replay_matches_direct_launches(small_scene, prec):
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    sc, ids, cams, imgs = small_scene
    print(f"case: graph replay {prec} 160x120 V=4", flush=True)
    direct = _ctx(cams, imgs, graph=False, precision=prec)
    graphed = _ctx(cams, imgs, graph=True, precision=prec)
    other = scenes.random_planes_per_pixel(cams[0], seed=9)
    for k, seed in enumerate([3, 4, 5, 6]):
        direct.run(seed)
        graphed.run(seed)
        np.testing.assert_array_equal(graphed.planes(), direct.planes())
        np.testing.assert_array_equal(graphed.costs(), direct.costs())
        # unrelated GPU work between replays
        graphed.cost_fixed(other, precision=H.PRECISIONS[prec])
        if k == 1:  # partial schedule and back: recapture twice
            direct.run(seed, n_halfpasses=3)
            graphed.run(seed, n_halfpasses=3)
            np.testing.assert_array_equal(graphed.planes(), direct.planes())


def test_graph_recaptures_on_mode_change(small_scene):
    import hpm_mi355x as H
    sc, ids, cams, imgs = small_scene
    direct = _ctx(cams, imgs, graph=False)
    graphed = _ctx(cams, imgs, graph=True)
    direct.run(1)
    graphed.run(1)
    base_p, base_c = direct.planes(), direct.costs()
    for prec, geom in [("fp16", 0), ("fp32", 1), ("fp32", 0)]:
        for c in (direct, graphed):
            c.set_params(H.default_params(precision=prec, geom_consistency=geom))
            if geom:
                c.set_src_depths([None] + [sc.depths[i] for i in ids[1:]])
                c.reload_hypotheses(base_p, base_c)
            c.run(7)
        np.testing.assert_array_equal(graphed.planes(), direct.planes())


def test_batch_matches_single_runs():
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    print("case: batch of 3 contexts (160x120 V=4, 128x96 V=6, 96x72 V=3)", flush=True)
    probs = []
    for (w, h, f, n, ref, seed) in [(160, 120, 131.25, 5, 2, 1), (128, 96, 105.0, 7, 3, 3), (96, 72, 78.75, 4, 1, 2)]:
        sc = scenes.make_plane_scene(seed=seed, n_cams=n, width=w, height=h, f=f, n_src=n - 1)
        probs.append(scenes.reference_view_problem(sc, ref, n - 1))
    batch = [_ctx(c, i, precision="fp16") for (_, c, i) in probs]
    single = [_ctx(c, i, precision="fp16") for (_, c, i) in probs]
    seeds = [11, 12, 13]
    H.run_batch(batch, seeds)
    for s, c, sd in zip(single, batch, seeds):
        s.run(sd)
        np.testing.assert_array_equal(c.planes(), s.planes())
        np.testing.assert_array_equal(c.costs(), s.costs())
    for c in batch:
        c.set_graph(True)
    H.run_batch(batch, seeds)
    H.run_batch(batch, seeds)
    for s, c in zip(single, batch):
        np.testing.assert_array_equal(c.planes(), s.planes())


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_pooled_views_match_copied_views(prec):
    """Contexts on a shared image pool give the bits of contexts holding their
    own copies, alone and batched."""
    import hpm_mi355x as H
    from hpm_mi355x import scenes
    print(f"case: image pool {prec} 128x96, 7 cameras, 3 reference views", flush=True)
    sc = scenes.make_plane_scene(seed=3, n_cams=7, width=128, height=96, f=105.0, n_src=4)
    pool = H.ImagePool([H.camera_from_scene(c) for c in sc.cameras], [im.astype(np.float32) for im in sc.images])
    pooled, copied, seeds = [], [], []
    for ref in (2, 3, 4):
        ids, cams, imgs = scenes.reference_view_problem(sc, ref, 4)
        c = H.Context(0)
        c.set_params(H.default_params(precision=prec))
        c.set_views_pooled(pool, ids)
        pooled.append(c)
        copied.append(_ctx(cams, imgs, precision=prec))
        seeds.append(40 + ref)
    H.run_batch(pooled, seeds)
    for c, d, sd in zip(pooled, copied, seeds):
        d.run(sd)
        np.testing.assert_array_equal(c.planes(), d.planes())
        np.testing.assert_array_equal(c.costs(), d.costs())
    pooled[1].run(seeds[1])  # alone, after the batch
    np.testing.assert_array_equal(pooled[1].planes(), copied[1].planes())


def _uneven_problems(n_views=4, extra=()):
    from hpm_mi355x import scenes
    specs = [(160, 120, 131.25, 5, 2, 1), (128, 96, 105.0, 7, 3, 3), (96, 72, 78.75, 4, 1, 2),
             (200, 150, 164.0, 6, 2, 5)][:n_views] + list(extra)
    out = []
    for (w, h, f, n, ref, seed) in specs:
        sc = scenes.make_plane_scene(seed=seed, n_cams=n, width=w, height=h, f=f, n_src=n - 1)
        out.append((sc,) + tuple(scenes.reference_view_problem(sc, ref, n - 1)))
    return out


def _same(a, b):
    np.testing.assert_array_equal(a.planes(), b.planes())
    np.testing.assert_array_equal(a.costs(), b.costs())
    np.testing.assert_array_equal(a.texture(), b.texture())
    sa, sb = a.stats(), b.stats()
    assert (sa.ncc_init, sa.ncc_propagate, sa.ncc_reference) == (sb.ncc_init, sb.ncc_propagate, sb.ncc_reference)


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
def test_fused_batch_matches_single_runs(prec):
    import hpm_mi355x as H
    print(f"case: fused batch {prec}, 4 views 96x72..200x150, V 3..6", flush=True)
    probs = _uneven_problems()
    fused = [_ctx(c, i, precision=prec) for (_, _, c, i) in probs]
    single = [_ctx(c, i, precision=prec) for (_, _, c, i) in probs]
    seeds = [21, 22, 23, 24]
    for s, sd in zip(single, seeds):
        s.run(sd)
    b = H.Batch(fused)
    # (tile width, strip, threads per workgroup)
    shapes = [(32, 2, 256), (32, 1, 256), (32, 4, 256)]
    if prec != "fp32":
        shapes += [(16, 1, 256), (64, 2, 256), (16, 4, 256), (16, 2, 128), (32, 1, 128), (64, 4, 128)]
    for graph in (False, True):
        b.set_graph(graph)
        for (w, st, nt) in shapes:
            b.set_shape(w, st)
            b.set_workgroup(nt)
            for _ in range(2 if graph else 1):  # capture, then replay
                b.run(seeds)
                for f, s in zip(fused, single):
                    _same(f, s)
    assert b.last_ms() > 0
    if prec == "fp32":
        with pytest.raises(H.HpmError):
            b.set_shape(16, 2)
            b.run(seeds)
        b.set_shape(32, 2)
        with pytest.raises(H.HpmError):
            b.set_workgroup(128)
            b.run(seeds)
    with pytest.raises(H.HpmError):
        b.set_workgroup(64)


def test_fused_batch_geometric_mode():
    """A geometric-consistency pass (source depth maps, reloaded hypotheses)
    through the fused batch equals the single-context pass."""
    import hpm_mi355x as H
    print("case: fused batch geom fp32 + fp16, 3 views", flush=True)
    probs = _uneven_problems(3)
    for prec in ("fp32", "fp16"):
        ctxs = {k: [_ctx(c, i, precision=prec) for (_, _, c, i) in probs] for k in ("fused", "single")}
        base = []
        for (sc, ids, c, i), s in zip(probs, ctxs["single"]):
            s.run(5)
            base.append((s.planes(), s.costs()))
        for k in ("fused", "single"):
            for (sc, ids, c, i), ctx, (p, cost) in zip(probs, ctxs[k], base):
                ctx.set_params(H.default_params(precision=prec, geom_consistency=1))
                ctx.set_src_depths([None] + [sc.depths[j] for j in ids[1:]])
                ctx.reload_hypotheses(p, cost)
        for s, sd in zip(ctxs["single"], (7, 8, 9)):
            s.run(sd)
        b = H.Batch(ctxs["fused"])
        b.set_graph(True)
        b.run([7, 8, 9])
        for f, s in zip(ctxs["fused"], ctxs["single"]):
            np.testing.assert_array_equal(f.planes(), s.planes())
            np.testing.assert_array_equal(f.costs(), s.costs())
        with pytest.raises(H.HpmError):  # reruns would not start from the same state
            b.autotune([7, 8, 9])


def test_fused_batch_private_rows_and_autotune():
    """A view with 25 sources (beyond the LDS rows) in the batch switches all
    views to private rows; results still equal single runs, and autotuning
    keeps a valid shape whose result equals them too."""
    import hpm_mi355x as H
    print("case: fused batch fp16 with a 25-source view + autotune", flush=True)
    probs = _uneven_problems(2, extra=[(64, 48, 52.5, 26, 12, 7)])
    fused = [_ctx(c, i, precision="fp16") for (_, _, c, i) in probs]
    single = [_ctx(c, i, precision="fp16") for (_, _, c, i) in probs]
    seeds = [31, 32, 33]
    for s, sd in zip(single, seeds):
        s.run(sd)
    b = H.Batch(fused)
    b.run(seeds)
    for f, s in zip(fused, single):
        _same(f, s)
    w, st, ms = b.autotune(seeds)
    assert w in (16, 32, 64) and st in (1, 2, 4) and b.workgroup() in (128, 256) and ms > 0
    for f, s in zip(fused, single):
        _same(f, s)
    # mixed precisions are refused
    fused[0].set_params(H.default_params(precision="fp32"))
    with pytest.raises(H.HpmError):
        b.run(seeds)

