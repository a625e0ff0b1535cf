"""Invariants of the PatchMatch schedule, on the oracle (CPU).

Red/black order independence: every pixel of one colour reads only pixels of
the other colour (HPM.cu:1367-1607, 1690 and the extended left-up reads of
1063-1126), so the order in which a colour's pixels are visited cannot change
any result.  The oracle visits them with OpenMP over rows; here the whole
run is repeated with every colour swept bottom-up and right to left, and
every output must be bit-identical, in the photometric, geometric and
mandatory-consistency passes.  The GPU engines visit the pixels of a colour
in yet other orders (tiles, workgroups, XCDs) and equal the oracle bit for
bit (test_gpu_parity.py).

View-selection masks: the propagation draws 15 samples of source views
(HPM.cu:1730-1740), so a pixel's mask has at most min(V, 15) bits, V the
number of sources; checked with V = 16, where 15 binds."""
import numpy as np
import pytest

import oracle_py as O

SEED = 11


def _scene(n_src, w=64, h=48, seed=2):
    from hpm_mi355x import scenes
    sc = scenes.make_plane_scene(seed=seed, n_cams=n_src + 1, width=w, height=h, f=0.82 * w, n_src=n_src)
    ids, cams, imgs = scenes.reference_view_problem(sc, n_src // 2, n_src)
    return sc, ids, [O.camera_from_scene(c) for c in cams], imgs


def _run(cams, imgs, reverse, **kw):
    O.lib().orc_set_reverse_sweep(int(reverse))
    try:
        extra = kw.pop("extra", {})
        return O.run_patchmatch(cams, imgs, O.params(len(cams), cams[0], **kw), SEED, **extra)
    finally:
        O.lib().orc_set_reverse_sweep(0)


def _same(a, b):
    for k in ("planes", "costs", "selected", "texture", "pre_costs"):
        assert np.array_equal(np.asarray(a[k]).view(np.uint8), np.asarray(b[k]).view(np.uint8)), k


def test_reversed_sweep_photometric():
    _, _, cams, imgs = _scene(4)
    fwd, rev = _run(cams, imgs, False), _run(cams, imgs, True)
    _same(fwd, rev)
    # the sweep did something: hypotheses moved away from the random start
    init = O.run_patchmatch(cams, imgs, O.params(len(cams), cams[0]), SEED, n_halfpasses=0)
    assert not np.array_equal(init["planes"], fwd["planes"])


def test_reversed_sweep_geometric_and_mandatory():
    sc, ids, cams, imgs = _scene(4)
    base = _run(cams, imgs, False)
    depths = [None] + [sc.depths[i] * np.float32(1.004) for i in ids[1:]]
    rng = np.random.default_rng(1)
    conf = rng.random(imgs[0].shape).astype(np.float32)
    for kw in (dict(geom_consistency=1), dict(geom_consistency=1, mand_consistency=1)):
        runs = []
        for reverse in (False, True):  # the pass updates its inputs in place: fresh copies for each run
            extra = dict(planes=base["planes"].copy(), costs=base["costs"].copy(), depth_images=depths,
                         conf=conf.copy())
            runs.append(_run(cams, imgs, reverse, extra=extra, **kw))
        _same(*runs)
        assert np.array_equal(runs[0]["conf"].view(np.uint8), runs[1]["conf"].view(np.uint8))
        assert not np.array_equal(runs[0]["planes"], base["planes"])
        if kw.get("mand_consistency"):  # the adoption branch ran and passed confidences on
            assert not np.array_equal(runs[0]["conf"], conf)


@pytest.mark.parametrize("n_src", [4, 16])
def test_selected_views_popcount(n_src):
    _, _, cams, imgs = _scene(n_src, w=48, h=32, seed=4)
    r = _run(cams, imgs, False)
    sel = r["selected"].astype(np.uint32)
    bits = np.unpackbits(sel.view(np.uint8)).reshape(sel.shape + (32,)).sum(-1)
    assert bits.max() <= min(n_src, 15)
    assert (sel >> np.uint32(n_src)).max() == 0  # no bit beyond the last source
    if n_src == 16:
        assert bits.max() > 4  # more than top_k: the propagation's samples, not the initial top-k
