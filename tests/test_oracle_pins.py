"""Pins the oracle's propagation family on the CPU, one pixel at a time.

orc_probe_pixel runs a single pixel's RandomInitialization or red/black
update of the FP32 oracle with the pixel's uniform draws given by the test
and the bilateral NCC (and geometric cost) of each (view, plane) given by a
python callback, and records what the update computed: the neighbour
positions and cost rows, the view-selection probabilities and weights, the
direction costs, the refinement candidates.  Each test states what the
reference prescribes for that step -- read from HPM.cu, written here as
explicit candidate tables, closed forms or special cases, not as a copy of
the oracle's loops -- and checks the oracle against it:

* neighbour search (HPM.cu:1385-1608): the four axial directions scan 5
  pixels at steps of 2 from 5 px, the four diagonals a 7-step zig-zag from
  (-5,-6), (6,-5), (5,6), (-6,5); a direction exists when its first pixel is
  in the image and the scan stops at the first pixel outside it; the first
  strict minimum of `costs` wins, except right_far, which keeps the first
  strict maximum (Q4, HPM.cu:1464).  A missing direction keeps the cost
  row of `float cost_array[8][32] = {2.0f}`: row 0 = (2, 0, ...), other
  rows 0 (Q3, HPM.cu:1378);
* JudgeExtend (HPM.cu:1322-1345) and the left-up extension (HPM.cu:1063-
  1126, 1609-1680): a direction extends while it has no cost below
  0.8 exp(-iter^2 (3 - round) / 90) or more than 2 above 1.2; round k
  rescans from (-5 - 8k, -6 - 8k) and keeps the old pixel when the scan's
  minimum is above its cost.  Only left-up ever extends (Q2);
* joint view selection (HPM.cu:1686-1751): prior 0.9 / 0.1 per axial
  neighbour's mask bit, mean exp(-c^2/0.18) over good directions, the
  exp(-tau^2/0.32) fallback, zero with 3 bad directions, 15 draws over the
  CDF with the FLT_EPSILON shift;
* refinement (HPM.cu:765-841): the five (depth, normal) candidates from
  the draw stream, the +-2 % depth perturbation, the 0.02 pi normal
  perturbation, sequential adoption of cheaper in-range candidates;
* geometric cost (HPM.cu:539-564) on a rectified pair, texture factor
  (HPM.cu:349-373), random normals (HPM.cu:159-185), the initialisation
  modes (HPM.cu:567-763), and the mandatory-consistency adoption with its
  confidence encoding and x0.5 hand-over (HPM.cu:2322-2790).
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_py as O

F32_EPS = float(np.finfo(np.float32).eps)
f32 = np.float32


def pinhole(W, H, f=100.0, t=(0.0, 0.0, 0.0), dmin=1.0, dmax=5.0):
    K = np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]])
    return O.camera(K, np.eye(3), t, W, H, dmin, dmax)


def depth_of(cam, plane, px, py):
    """ComputeDepthfromPlaneHypothesis (HPM.cu:154-157) in the FP64 oracle."""
    pl = np.ascontiguousarray(plane, np.float64)
    return float(O.lib().orc_depth_from_plane_f64(ctypes.byref(cam), pl.ctypes.data, px, py))


def cams_for(W, H, V):
    return [pinhole(W, H)] + [pinhole(W, H, t=(-0.1 * (i + 1), 0.0, 0.0)) for i in range(V)]


class PlaneTable:
    """Fronto-parallel planes, one per pixel, each with its own depth, and a
    table of per-view costs per pixel: the NCC callback looks the plane up."""

    def __init__(self, W, H, V, rows=None, default=1.9):
        self.W, self.H, self.V, self.default = W, H, V, default
        self.planes = np.zeros((H, W, 4), np.float32)
        self.planes[..., 2] = -1.0
        # depth = w for a fronto-parallel plane (HPM.cu:154-157): distinct per pixel
        self.planes[..., 3] = (2.0 + 1e-4 * np.arange(W * H)).reshape(H, W)
        self.key = {self.planes.reshape(-1, 4)[q].tobytes(): q for q in range(W * H)}
        self.rows = np.full((W * H, V), 1.0, np.float32) if rows is None else rows
        self.calls = []

    def ncc(self, v, x, y, pl):
        q = self.key.get(np.asarray(pl, np.float32).tobytes())
        self.calls.append((v, q))
        return float(self.rows[q, v - 1]) if q is not None else self.default


# ------------------------------------------------------------------ neighbour search

# The pixels each direction scans, relative to the pixel, in scan order
# (HPM.cu:1385-1608).  Index = direction: 0 left_up, 1 up_far, 2 right_up,
# 3 down_far, 4 right_down, 5 left_far, 6 left_down, 7 right_far.
SCAN = {
    0: [(-5, -6), (-7, -6), (-7, -8), (-9, -8), (-9, -10), (-11, -10), (-11, -12), (-13, -12)],
    1: [(0, -5), (0, -7), (0, -9), (0, -11), (0, -13)],
    2: [(6, -5), (6, -7), (8, -7), (8, -9), (10, -9), (10, -11), (12, -11), (12, -13)],
    3: [(0, 5), (0, 7), (0, 9), (0, 11), (0, 13)],
    4: [(5, 6), (7, 6), (7, 8), (9, 8), (9, 10), (11, 10), (11, 12), (13, 12)],
    5: [(-5, 0), (-7, 0), (-9, 0), (-11, 0), (-13, 0)],
    6: [(-6, 5), (-6, 7), (-8, 7), (-8, 9), (-10, 9), (-10, 11), (-12, 11), (-12, 13)],
    7: [(5, 0), (7, 0), (9, 0), (11, 0), (13, 0)],
}


def expected_search(costs, W, H, px, py):
    """(valid, pixel) per direction by the rule above."""
    out = []
    for d in range(8):
        pts = []
        for dx, dy in SCAN[d]:
            x, y = px + dx, py + dy
            if not (0 <= x < W and 0 <= y < H):
                break
            pts.append(y * W + x)
        if not pts:
            out.append((False, None))
            continue
        best = pts[0]
        for q in pts[1:]:
            if (costs.flat[q] > costs.flat[best]) if d == 7 else (costs.flat[q] < costs.flat[best]):
                best = q
        out.append((True, best))
    return out


def _maps(tab, costs):
    return dict(planes=tab.planes.copy(), costs=np.ascontiguousarray(costs, np.float32))


def test_neighbour_search_matches_the_scan_tables():
    W, H, V = 44, 40, 3
    rng = np.random.default_rng(3)
    rows = rng.uniform(0.2, 1.0, (W * H, V)).astype(np.float32)
    rows[:, 0] = 0.5  # a good cost in every row: left-up never extends
    tab = PlaneTable(W, H, V, rows)
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0])
    pixels = [(x, y) for x in (0, 3, 5, 6, 7, 12, 20, 30, 37, 38, 39, 43) for y in (0, 4, 5, 6, 11, 20, 33, 34, 35, 39)]
    for trial, (px, py) in enumerate(pixels):
        # trial 0 mod 3: distinct costs; 1: three levels (ties); 2: all equal
        if trial % 3 == 0:
            costs = rng.permutation(W * H).astype(np.float32).reshape(H, W) / (W * H)
        elif trial % 3 == 1:
            costs = rng.integers(0, 3, (H, W)).astype(np.float32) * 0.25
        else:
            costs = np.full((H, W), 0.5, np.float32)
        r = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, 0, draws=np.full(40, 0.5), ncc=tab.ncc)
        for d, (valid, q) in enumerate(expected_search(costs, W, H, px, py)):
            assert bool(r["flags"][d]) == valid, (px, py, d)
            if valid:
                assert r["positions"][d] == q, (px, py, d, r["positions"][d], q)
                np.testing.assert_array_equal(r["rows"][d][:V], rows[q])
            else:  # Q3: `float cost_array[8][32] = {2.0f}`
                expect = np.zeros(32, np.float32)
                if d == 0:
                    expect[0] = 2.0
                np.testing.assert_array_equal(r["rows"][d], expect)


def test_right_far_keeps_the_maximum_and_up_far_the_minimum():
    W, H, V = 40, 40, 2
    tab = PlaneTable(W, H, V, np.full((W * H, V), 0.5, np.float32))
    cams = cams_for(W, H, V)
    px, py = 20, 20
    costs = np.full((H, W), 1.0, np.float32)
    vals = [0.5, 0.9, 0.3, 0.7, 0.6]
    for i, v in enumerate(vals):
        costs[py, px + 5 + 2 * i] = v  # right_far scan
        costs[py - 5 - 2 * i, px] = v  # up_far scan
    r = O.probe_pixel(1, cams, O.params(len(cams), cams[0]), _maps(tab, costs), px, py, 0,
                      draws=np.full(40, 0.5), ncc=tab.ncc)
    assert r["positions"][7] == py * W + px + 7  # the maximum, 0.9
    assert r["positions"][1] == (py - 9) * W + px  # the minimum, 0.3


# ------------------------------------------------------------------ extension

def judge_threshold(it, rnd):
    # good_threshold = 0.8 * exp(-iter * iter * iter_tmp / 90.0), iter_tmp = 3 - round
    return f32(0.8 * math.exp(-it * it * (3 - rnd) / 90.0))


def expected_extension(costs, rows, W, H, px, py, it, pos0):
    """Left-up position and row after JudgeExtend / ExtendedLeftUpPropagation,
    and the number of cost-vector evaluations the rounds made."""
    pos, row, evals = pos0, rows[pos0], 0
    for rnd in range(3):
        good = int(np.sum(row < judge_threshold(it, rnd)))
        bad = int(np.sum(row > f32(1.2)))
        if good >= 1 and bad <= 2:
            break
        sx, sy = px - 5 - 8 * rnd, py - 6 - 8 * rnd
        if not (px > 4 + 8 * rnd and py > 5 + 8 * rnd):
            continue  # keeps the position and the row
        pts = [sy * W + sx]
        x, y = sx, sy
        for i in range(7):
            if i % 2 == 0:
                x -= 2
            else:
                y -= 2
            if not (0 <= x < W and 0 <= y < H):
                break
            pts.append(y * W + x)
        best = pts[0]
        for q in pts[1:]:
            if costs.flat[q] < costs.flat[best]:
                best = q
        if costs.flat[best] > costs.flat[pos]:
            continue
        pos, row, evals = best, rows[best], evals + 1
    return pos, row, evals


def test_extension_rounds_and_judge_thresholds():
    W, H, V = 64, 64, 4
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0])
    px, py = 40, 44
    rng = np.random.default_rng(5)
    for it in range(3):
        for case in range(60):
            rows = rng.uniform(0.5, 1.0, (W * H, V)).astype(np.float32)
            rows[:, 0] = 0.5
            costs = rng.uniform(0, 1, (H, W)).astype(np.float32)
            # the left-up region: rows that sit on the three rounds' thresholds or are bad
            lu = [(x, y) for x in range(px - 30, px - 4) for y in range(py - 30, py - 5)]
            for (x, y) in lu:
                q = y * W + x
                kind = rng.integers(0, 4)
                if kind == 0:
                    rows[q] = [1.5, 1.5, 1.5, 1.0]  # 3 bad: extends
                elif kind == 1:
                    rows[q] = [0.95, 0.9, 1.0, 1.0]  # no good cost: extends
                else:  # one cost just below / above a threshold of this iteration
                    rnd = rng.integers(0, 3)
                    t = float(judge_threshold(it, rnd))
                    rows[q] = [t * (1 - 2e-4) if kind == 2 else t * (1 + 2e-4), 1.0, 1.0, 1.0]
            tab = PlaneTable(W, H, V, rows)
            r = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, it, draws=np.full(40, 0.5), ncc=tab.ncc)
            base = expected_search(costs, W, H, px, py)
            pos, row, evals = expected_extension(costs, rows, W, H, px, py, it, base[0][1])
            assert r["positions"][0] == pos, (it, case)
            np.testing.assert_array_equal(r["rows"][0][:V], row)
            assert r["n_costvec"] == 8 + evals, (it, case)
            for d in range(1, 8):  # only left-up extends (Q2)
                assert r["positions"][d] == base[d][1]


def test_judge_threshold_boundary_per_iteration():
    """A left-up row whose one good cost sits just below the threshold of
    iteration `it` stops the extension; just above, round 0 runs (its rescan
    re-evaluates the same pixel: one more cost vector)."""
    W, H, V = 40, 40, 4
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0])
    px, py = 20, 20
    costs = np.full((H, W), 1.0, np.float32)
    for it, thr in ((0, 0.8), (1, 0.8 * math.exp(-3 / 90.0)), (2, 0.8 * math.exp(-12 / 90.0))):
        for scale, extends in ((1 - 1e-4, False), (1 + 1e-4, True)):
            rows = np.full((W * H, V), 1.0, np.float32)
            rows[:, 0] = 0.5
            q = (py - 6) * W + (px - 5)  # left_up's first pixel wins on equal costs
            rows[q] = [thr * scale, 1.0, 1.0, 1.0]
            tab = PlaneTable(W, H, V, rows)
            r = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, it, draws=np.full(40, 0.5), ncc=tab.ncc)
            assert (r["n_costvec"] > 8) == extends, (it, scale)
            if not extends:
                assert r["positions"][0] == q
            pos, _, evals = expected_extension(costs, rows, W, H, px, py, it, q)
            assert r["positions"][0] == pos and r["n_costvec"] == 8 + evals
    # three bad costs extend although one is good; two do not
    for nbad, extends in ((2, False), (3, True)):
        rows = np.full((W * H, V), 0.5, np.float32)
        q = (py - 6) * W + (px - 5)
        rows[q] = [0.1] + [1.3] * nbad + [1.0] * (3 - nbad)
        tab = PlaneTable(W, H, V, rows)
        r = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, 0, draws=np.full(40, 0.5), ncc=tab.ncc)
        assert (r["n_costvec"] > 8) == extends, nbad


# ------------------------------------------------------------------ view selection

def _vs_setup(W=40, H=40, V=4, px=20, py=20):
    """Every direction picks its first scanned pixel: `costs` is 0 there
    (2 for right_far, which keeps the maximum) and 1 elsewhere."""
    costs = np.full((H, W), 1.0, np.float32)
    first = {}
    for d in range(8):
        dx, dy = SCAN[d][0]
        first[d] = (py + dy) * W + (px + dx)
        costs.flat[first[d]] = 2.0 if d == 7 else 0.0
    return W, H, V, px, py, costs, first


def vs_threshold(it):
    # float cost_threshold = 0.8 * expf((iter) * (iter) / (-90.0f)) (HPM.cu:1706)
    return float(f32(0.8 * float(f32(math.exp(it * it / -90.0)))))


def expected_pdf(rows8, flags, nbr_masks, it, V):
    """Sampling probabilities times priors (HPM.cu:1686-1728), float64."""
    prior = np.zeros(V)
    for i, d in enumerate((1, 3, 5, 7)):
        if flags[d]:
            for j in range(V):
                prior[j] += 0.9 if (nbr_masks[i] >> j) & 1 else 0.1
    tau = vs_threshold(it)
    p = np.zeros(V)
    for j in range(V):
        c = rows8[:, j].astype(np.float64)
        good = c < tau
        nbad = int(np.sum(c > 1.2))
        if good.sum() > 2 and nbad < 3:
            p[j] = np.mean(np.exp(-c[good] ** 2 / 0.18))
        elif nbad < 3:
            p[j] = math.exp(-tau * tau / 0.32)
        p[j] *= prior[j]
    return p, prior


def expected_weights(pdf32, draws):
    """Draw counts from the float32 CDF (TransformPDFToCDF, HPM.cu:98-112) and
    15 draws shifted by FLT_EPSILON (HPM.cu:1731-1740)."""
    s = f32(0)
    for v in pdf32:
        s = f32(s + v)
    inv = f32(f32(1) / s)
    cdf, cum = [], f32(0)
    for v in pdf32:
        cum = f32(cum + f32(v * inv))
        cdf.append(cum)
    w = np.zeros(len(pdf32))
    for r in draws:
        rr = f32(f32(r) - f32(F32_EPS))
        for j, c in enumerate(cdf):
            if c > rr:
                w[j] += 1
                break
    return w, cdf


def test_view_selection_probabilities_and_draws():
    W, H, V, px, py, costs, first = _vs_setup()
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0])
    rows = np.full((W * H, V), 1.0, np.float32)
    rng = np.random.default_rng(9)
    for it in range(3):
        # view 0: every direction at 0 -> mean exp(0) = 1
        # view 1: every direction 1.0 (neither good nor bad) -> the fallback exp(-tau^2 / 0.32)
        # view 2: three directions above 1.2 -> 0
        # view 3: random costs
        for d in range(8):
            q = first[d]
            rows[q] = [0.0, 1.0, 1.3 if d < 3 else 0.2, rng.uniform(0, 1.5)]
        tab = PlaneTable(W, H, V, rows)
        maps = _maps(tab, costs)
        nbr = [0b0001, 0b0011, 0b1001, 0b0000]  # up, down, left, right axial neighbours' masks
        for (dx, dy), m in zip(((0, -5), (0, 5), (-5, 0), (5, 0)), nbr):
            maps.setdefault("selected", np.zeros((H, W), np.uint32))[py + dy, px + dx] = m
        draws = rng.uniform(0, 1, 40).astype(np.float32)
        r = O.probe_pixel(1, cams, prm, maps, px, py, it, draws=draws, ncc=tab.ncc)
        rows8 = np.array([rows[first[d]] for d in range(8)])
        p, prior = expected_pdf(rows8, [1] * 8, nbr, it, V)
        tau = vs_threshold(it)
        assert p[1] == pytest.approx(math.exp(-tau * tau / 0.32) * prior[1], rel=1e-9)
        assert prior[1] == pytest.approx(0.1 + 0.9 + 0.1 + 0.1)  # bit 1 set in the down neighbour's mask only
        assert p[2] == 0.0 and p[0] == pytest.approx(prior[0], rel=1e-12)
        np.testing.assert_allclose(r["pdf"][:V], p, rtol=2e-6, atol=1e-12)
        w, _ = expected_weights(r["pdf"][:V].astype(np.float32), draws[:15])
        np.testing.assert_array_equal(r["weights"][:V], w)
        assert r["sel"] == sum(1 << j for j in range(V) if w[j] > 0)
        assert r["weight_norm"] == 15.0


def test_view_selection_border_zero_rows_count_as_good():
    """At the image border the missing directions' rows are 0 (Q3) and count
    as good costs: a view with costs of 1.0 everywhere else gets exp(0) = 1
    from them instead of the fallback."""
    W, H, V = 40, 40, 2
    px, py = 2, 2  # up_far, left_far, left_up, right_up, left_down missing
    cams = cams_for(W, H, V)
    rows = np.full((W * H, V), 1.0, np.float32)
    tab = PlaneTable(W, H, V, rows)
    costs = np.full((H, W), 1.0, np.float32)
    r = O.probe_pixel(1, cams, O.params(len(cams), cams[0]), _maps(tab, costs), px, py, 0,
                      draws=np.full(40, 0.5), ncc=tab.ncc)
    assert list(r["flags"]) == [0, 0, 0, 1, 1, 0, 0, 1]
    # view 0: row 0 of a missing left_up is 2 (bad), rows 1, 2, 5, 6 are 0 (good):
    # 4 good, 1 bad -> mean of exp(0) = 1; view 1: 5 zeros -> 1.  prior = 2 axial
    # neighbours (down, right) x 0.1
    np.testing.assert_allclose(r["pdf"][:2], [0.2, 0.2], rtol=1e-6)


def test_one_good_view_takes_every_draw_and_the_epsilon_shift():
    W, H, V, px, py, costs, first = _vs_setup()
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0])
    rows = np.full((W * H, V), 1.0, np.float32)
    for d in range(8):
        rows[first[d]] = [1.5, 0.1, 1.5, 1.5]  # views 0, 2, 3 have 8 bad directions
    tab = PlaneTable(W, H, V, rows)
    r = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, 0, draws=np.linspace(0.01, 1.0, 40), ncc=tab.ncc)
    np.testing.assert_array_equal(r["weights"][:V], [0, 15, 0, 0])
    assert r["sel"] == 0b10
    # two views with probabilities p0, p1: a draw r with F0 < r < F0 + FLT_EPSILON
    # still selects view 0, because the reference compares F > r - FLT_EPSILON
    for d in range(8):
        rows[first[d]] = [0.0, 1.0, 1.5, 1.5]
    tab = PlaneTable(W, H, V, rows)
    r0 = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, 0, draws=np.full(40, 0.5), ncc=tab.ncc)
    _, cdf = expected_weights(r0["pdf"][:V].astype(np.float32), [])
    f0 = cdf[0]
    inside = np.nextafter(f0, f32(2))  # F0 < r < F0 + eps
    assert f32(inside - f32(F32_EPS)) < f0 < inside
    draws = np.concatenate([np.full(15, inside, np.float32), np.full(25, 0.5, np.float32)])
    r = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, 0, draws=draws, ncc=tab.ncc)
    np.testing.assert_array_equal(r["weights"][:V], [15, 0, 0, 0])
    above = f32(f0 + f32(2 * F32_EPS))
    draws[:15] = above
    r = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, 0, draws=draws, ncc=tab.ncc)
    np.testing.assert_array_equal(r["weights"][:V], [0, 15, 0, 0])


# ------------------------------------------------------------------ refinement

def view_dir(cam, px, py, depth):
    K = np.array(cam.K).reshape(3, 3)
    X = depth * np.array([(px - K[0, 2]) / K[0, 0], (py - K[1, 2]) / K[1, 1], 1.0])
    return X / np.linalg.norm(X), X


def rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def rot_y(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def rot_z(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def random_normal(cam, px, py, depth, u1, u2):
    """Marsaglia (1972) from one accepted pair, turned to face the camera (HPM.cu:159-185)."""
    q1, q2 = 2 * u1 - 1, 2 * u2 - 1
    s = q1 * q1 + q2 * q2
    assert s < 1
    n = np.array([2 * q1 * math.sqrt(1 - s), 2 * q2 * math.sqrt(1 - s), 1 - 2 * s])
    vd, _ = view_dir(cam, px, py, depth)
    if n @ vd > 0:
        n = -n
    return n / np.linalg.norm(n)


def perturbed_normal(cam, px, py, n, u, pert):
    """Euler rotation Rz(a3) Ry(a2) Rx(a1), a_k = (u_k - 1/2) pert, kept only if
    it still faces the camera (HPM.cu:187-222)."""
    a = [(uk - 0.5) * pert for uk in u]
    m = rot_z(a[2]) @ rot_y(a[1]) @ rot_x(a[0]) @ n
    vd, _ = view_dir(cam, px, py, 1.0)
    if m @ vd >= 0:
        m = n
    return m / np.linalg.norm(m)


def _refine_case(u_pert_depth, cand_costs, cur_cost=0.6):
    W, H, V = 40, 40, 1
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0])
    px, py = 20, 20
    costs = np.full((H, W), 1.0, np.float32)
    tab = PlaneTable(W, H, V, np.full((W * H, V), 0.7, np.float32))  # neighbours: 0.7, good, not adopted
    maps = _maps(tab, costs)
    n0 = np.array([0.1, -0.05, -1.0]); n0 /= np.linalg.norm(n0)
    d0 = 3.0
    _, X = view_dir(cams[0], px, py, d0)
    cur = np.array([n0[0], n0[1], n0[2], -(n0 @ X)], np.float32)
    maps["planes"][py, px] = cur
    cur_key = cur.tobytes()
    seen = []

    def ncc(v, x, y, pl):
        k = np.asarray(pl, np.float32).tobytes()
        if k == cur_key:
            return cur_cost
        if k in tab.key:
            return 0.7
        seen.append(np.asarray(pl, np.float32).copy())
        return cand_costs[len(seen) - 1]
    # 15 view-selection draws (one view), then depth_rand, the normal pair,
    # depth_perturbed, three angles
    ref = [0.8, 0.6, 0.7, u_pert_depth, 0.9, 0.2, 0.65]
    draws = np.concatenate([np.full(15, 0.5), ref]).astype(np.float32)
    r = O.probe_pixel(1, cams, prm, maps, px, py, 0, draws=draws, ncc=ncc)
    return r, maps, cams, prm, (px, py), n0, ref, seen


@pytest.mark.parametrize("u3", [0.0, 1.0, 0.25])
def test_refinement_candidates(u3):
    r, maps, cams, prm, (px, py), n0, u, seen = _refine_case(u3, [0.9] * 5)
    assert r["draws_used"] == 22 and r["n_cand"] == 5 and len(seen) == 5
    dmin, dmax = prm.depth_min, prm.depth_max
    depth = depth_of(cams[0], maps["planes"][py, px], px, py)
    depth_rand = u[0] * (dmax - dmin) + dmin
    n_rand = random_normal(cams[0], px, py, depth, u[1], u[2])
    depth_pert = u[3] * (1.02 * depth - 0.98 * depth) + 0.98 * depth  # +-2 %, HPM.cu:789-794
    n_pert = perturbed_normal(cams[0], px, py, n0, u[4:7], 0.02 * math.pi)  # HPM.cu:796
    expect = [(depth_rand, n0), (depth, n_rand), (depth_rand, n_rand), (depth, n_pert), (depth_pert, n0)]
    for i, (d, n) in enumerate(expect):
        assert r["cand_depth"][i] == pytest.approx(d, rel=2e-6), i
        np.testing.assert_allclose(r["cand"][i][:3], n, atol=2e-6, err_msg=str(i))
        _, X = view_dir(cams[0], px, py, d)
        assert r["cand"][i][3] == pytest.approx(-(n @ X), rel=1e-5, abs=1e-6), i
    # no candidate is cheaper than the current 0.6: the plane stays
    assert maps["costs"][py, px] == pytest.approx(0.6, rel=1e-6)


def test_refinement_adopts_each_cheaper_candidate_in_turn():
    r, maps, cams, prm, (px, py), *_ = _refine_case(0.5, [0.55, 0.5, 0.58, 0.4, 0.45])
    # 0.55 < 0.6, 0.5 < 0.55, 0.58 no, 0.4 < 0.5, 0.45 no: candidate 3 wins
    np.testing.assert_array_equal(maps["planes"][py, px], r["cand"][3])
    assert maps["costs"][py, px] == pytest.approx(0.4, rel=1e-6)
    # a cheaper candidate out of the depth range is not adopted: depth_rand = dmax + 20 %
    r, maps, cams, prm, (px, py), *_ = _refine_case(0.5, [0.1, 0.9, 0.9, 0.9, 0.9])
    np.testing.assert_array_equal(maps["planes"][py, px], r["cand"][0])  # u0 = 0.8: in range, adopted


# ------------------------------------------------------------------ geometric cost

def test_geometric_cost_on_a_rectified_pair():
    """Reference and source differ by a baseline b along x, looking down +z at
    a fronto-parallel plane at depth Z: the plane's source position is
    u_s = x - 8.25 (disparity f b / Z).  The source depth D is read at the
    truncated column (int)u_s (clamped to the image), but the back-projection
    uses u_s itself (HPM.cu:549-556): it lands at u_s + f b / D, so the cost is
    min(3, 8.25 |1 - Z / D|) -- 0 for a consistent map -- and 3 where D = 0."""
    W, H, f, Z = 40, 16, 100.0, 4.0
    b = 8.25 * Z / f
    cams = [pinhole(W, H, f), pinhole(W, H, f, t=(-b, 0.0, 0.0))]
    planes = np.zeros((H, W, 4), np.float32)
    planes[..., 2] = -1.0
    planes[..., 3] = Z
    zero = np.zeros((H, W), np.float32)
    g = O.geom_cost(cams, [zero, np.full((H, W), Z, np.float32)], planes)[..., 0]
    assert np.abs(g).max() < 2e-5
    k = np.arange(W) % 5
    D = np.broadcast_to(Z * (1 + 0.02 * k), (H, W)).astype(np.float32)
    g = O.geom_cost(cams, [zero, np.ascontiguousarray(D)], planes)[..., 0]
    x = np.arange(W)
    col = np.where(x - 8.25 > -1, np.trunc(x - 8.25), 0).astype(int)  # (int) toward zero, clamped at 0
    expect = np.minimum(3.0, 8.25 * np.abs(1 - Z / D[0, col]))
    np.testing.assert_allclose(g, np.broadcast_to(expect, (H, W)), atol=5e-5)
    rounded = np.minimum(3.0, 8.25 * np.abs(1 - Z / D[0, np.clip(np.rint(x - 8.25), 0, W - 1).astype(int)]))
    assert np.mean(np.abs(rounded - expect) > 1e-3) > 0.5  # truncation and rounding differ here
    assert np.all(O.geom_cost(cams, [zero, zero], planes)[..., 0] == 3.0)
    assert np.all(O.geom_cost(cams, [zero, np.full((H, W), 2 * Z, np.float32)], planes)[..., 0] == 3.0)


def test_geometric_term_of_the_direction_costs():
    """In the geometric pass a direction's cost per selected view is
    NCC + 0.2 geom; a missing direction adds 0.1 * 3 instead (HPM.cu:1753-1771)."""
    W, H, V = 40, 40, 2
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0], geom_consistency=1)
    px, py = 2, 20  # left_up, left_far, left_down missing
    rows = np.full((W * H, V), 0.5, np.float32)
    tab = PlaneTable(W, H, V, rows)
    costs = np.full((H, W), 1.0, np.float32)
    r = O.probe_pixel(1, cams, prm, _maps(tab, costs), px, py, 0, draws=np.full(40, 0.3), ncc=tab.ncc,
                      geom=lambda v, x, y, pl: 1.5 if v == 1 else 0.25)
    w = r["weights"][:V]
    assert w.sum() == 15
    for d in range(8):
        row = r["rows"][d][:V].astype(np.float64)
        if r["flags"][d]:
            expect = np.sum(w * (row + 0.2 * np.array([1.5, 0.25]))) / 15
        else:
            expect = np.sum(w * (row + 0.3)) / 15
        assert r["fc"][d] == pytest.approx(expect, rel=1e-6), d


# ------------------------------------------------------------------ texture, normals, init

def test_texture_factor():
    """t = (e + 5e-5) / (e + 1e-4), e the share of edge pixels among the 6 x 6
    taps inside the image (HPM.cu:349-373): no edge gives 0.5."""
    W, H, V = 24, 20, 1
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0])
    for (px, py, edges, n) in ((10, 10, [], 36), (10, 10, [(-5, -5), (5, 5), (1, -3)], 36),
                               (0, 0, [(1, 1), (5, 3)], 9), (23, 0, [(-1, 1)], 9), (10, 10, "all", 36)):
        canny = np.zeros((H, W), np.uint32)
        if edges == "all":
            canny[:] = 255
            k = n
        else:
            for dx, dy in edges:
                canny[py + dy, px + dx] = 255
            k = len(edges)
            if px + 2 < W:
                canny[py, px + 2] = 255  # an even offset is never a tap
        maps = dict(canny=canny)
        O.probe_pixel(0, cams, prm, maps, px, py, draws=np.full(10, 0.5), ncc=lambda *a: 1.0)
        e = k / n
        assert maps["texture"][py, px] == pytest.approx((e + 5e-5) / (e + 1e-4), rel=1e-6), (px, py, edges)
    maps = dict()
    O.probe_pixel(0, cams, prm, maps, 10, 10, draws=np.full(10, 0.5), ncc=lambda *a: 1.0)
    assert maps["texture"][10, 10] == 0.5


def test_random_normals_face_the_camera_uniformly():
    """Random planes (HPM.cu:159-230): unit normals facing the camera, their
    component along the viewing ray uniform on [0, 1] (a uniform direction on
    the facing hemisphere), the azimuth uniform, depth uniform in the range."""
    W, H = 64, 48
    cam = pinhole(W, H, 60.0)
    out = np.zeros(4, np.float32)
    t, az, dep = [], [], []
    rng = np.random.default_rng(2)
    for k in range(6000):
        px, py = int(rng.integers(0, W)), int(rng.integers(0, H))
        O.lib().orc_random_plane_f32(ctypes.byref(cam), px, py, 17, k, 1.0, 5.0, out.ctypes.data)
        n = out[:3].astype(np.float64)
        assert abs(np.linalg.norm(n) - 1) < 2e-6
        vd, _ = view_dir(cam, px, py, 1.0)
        c = -(n @ vd)
        assert c >= -1e-6
        t.append(c)
        e1 = np.cross(vd, [0, 1, 0]); e1 /= np.linalg.norm(e1)
        e2 = np.cross(vd, e1)
        az.append(math.atan2(n @ e2, n @ e1))
        dep.append(depth_of(cam, out, px, py))
    from scipy import stats
    assert stats.kstest(t, "uniform", args=(0, 1)).statistic < 0.025
    assert stats.kstest(az, "uniform", args=(-math.pi, 2 * math.pi)).statistic < 0.025
    assert stats.kstest(dep, "uniform", args=(1.0, 4.0)).statistic < 0.025


def test_random_initialisation_draws_and_top_k_cost():
    """Photometric initialisation (HPM.cu:585-588, 450-488): depth = u0 (max -
    min) + min, the normal from (u1, u2); the cost is the mean of the
    min(#valid, 4) smallest view costs and the mask holds every view at or
    below the k-th smallest (ties included); no valid view gives 2 and 0."""
    W, H = 30, 20
    px, py = 12, 7
    for vec, cost, mask in (([0.3, 2.0, 0.1, 0.3, 0.5], (0.1 + 0.3 + 0.3 + 0.5) / 4, 0b11101),
                            ([0.3] * 5, 0.3, 0b11111),
                            ([2.0, 2.0, 0.4, 2.0], 0.4, 0b0100),
                            ([2.0, 2.0], 2.0, 0)):
        V = len(vec)
        cams = cams_for(W, H, V)
        prm = O.params(len(cams), cams[0])
        maps = dict()
        u = [0.25, 0.6, 0.7]
        O.probe_pixel(0, cams, prm, maps, px, py, draws=u, ncc=lambda v, x, y, pl: vec[v - 1])
        assert maps["costs"][py, px] == pytest.approx(cost, rel=1e-6)
        assert maps["selected"][py, px] == mask
        pl = maps["planes"][py, px].astype(np.float64)
        d = depth_of(cams[0], pl, px, py)
        assert d == pytest.approx(u[0] * (prm.depth_max - prm.depth_min) + prm.depth_min, rel=1e-5)
        np.testing.assert_allclose(pl[:3], random_normal(cams[0], px, py, d, u[1], u[2]), atol=2e-6)


def test_prior_and_mandatory_initialisation_special_cases():
    W, H, V = 30, 20, 2
    px, py = 12, 7
    cams = cams_for(W, H, V)
    n = np.array([0.2, 0.1, -1.0]); n /= np.linalg.norm(n)
    _, X = view_dir(cams[0], px, py, 2.5)

    def fresh(mask, conf, cost):
        m = dict(planes=np.zeros((H, W, 4), np.float32), masks=np.zeros((H, W), np.uint32),
                 conf=np.zeros((H, W), np.float32), costs=np.zeros((H, W), np.float32),
                 prior=np.zeros((H, W, 4), np.float32))
        m["planes"][py, px] = [n[0], n[1], n[2], 2.5]  # the reloaded depth sits in w
        m["prior"][py, px] = [n[0], n[1], n[2], 2.5]
        m["masks"][py, px], m["conf"][py, px], m["costs"][py, px] = mask, conf, cost
        return m
    prm = O.params(len(cams), cams[0], prior_consistency=1)
    # confident pixel with a prior: kept, cost 0.1, w = distance to the origin (HPM.cu:591-597)
    m = fresh(3, 0.5, 0.7)
    O.probe_pixel(0, cams, prm, m, px, py, draws=np.full(8, 0.5), ncc=lambda *a: 0.9)
    assert m["costs"][py, px] == pytest.approx(0.1)
    np.testing.assert_allclose(m["planes"][py, px], [n[0], n[1], n[2], -(n @ X)], rtol=1e-6)
    # prior, not confident, cost above 0.1: depth within +-6 % of the prior, normal perturbed by
    # at most 3 * 0.02 pi / 2 per Euler angle; u = 1/2 leaves the normal; w holds the depth (HPM.cu:609)
    for u0, dexp in ((0.0, 0.94 * 2.5), (1.0, 1.06 * 2.5), (0.5, 2.5)):
        m = fresh(3, 0.2, 0.7)
        O.probe_pixel(0, cams, prm, m, px, py, draws=[u0, 0.5, 0.5, 0.5], ncc=lambda *a: 0.9)
        assert m["planes"][py, px][3] == pytest.approx(dexp, rel=1e-6)
        np.testing.assert_allclose(m["planes"][py, px][:3], n, atol=1e-6)
        assert m["costs"][py, px] == pytest.approx(0.9)
    m = fresh(3, 0.2, 0.7)
    O.probe_pixel(0, cams, prm, m, px, py, draws=[0.5, 1.0, 0.5, 0.5], ncc=lambda *a: 0.9)
    a = 3 * 0.02 * math.pi / 2  # (1 - 1/2) * pert about x
    np.testing.assert_allclose(m["planes"][py, px][:3], perturbed_normal(cams[0], px, py, n, [1.0, 0.5, 0.5],
                                                                        3 * 0.02 * math.pi), atol=2e-6)
    # a rotation by a about x keeps n_x and turns the rest: n . n' = n_x^2 + (1 - n_x^2) cos a
    assert m["planes"][py, px][:3] @ n == pytest.approx(n[0] ** 2 + (1 - n[0] ** 2) * math.cos(a), abs=2e-6)
    assert m["planes"][py, px][0] == pytest.approx(n[0], abs=1e-6)
    # prior, cost already <= 0.1: plane from the reload, cost from the views
    m = fresh(3, 0.2, 0.05)
    O.probe_pixel(0, cams, prm, m, px, py, draws=np.full(8, 0.5), ncc=lambda *a: 0.9)
    assert m["costs"][py, px] == pytest.approx(0.9)
    # no prior: plane from the reload; confident -> cost 0.1, else the cost stays
    m = fresh(0, 0.5, 0.7)
    O.probe_pixel(0, cams, prm, m, px, py, draws=np.full(8, 0.5), ncc=lambda *a: 0.9)
    assert m["costs"][py, px] == pytest.approx(0.1)
    m = fresh(0, 0.2, 0.7)
    O.probe_pixel(0, cams, prm, m, px, py, draws=np.full(8, 0.5), ncc=lambda *a: 0.9)
    assert m["costs"][py, px] == pytest.approx(0.7)
    # mandatory pass: unconfident pixel on an edge-rich window (texture > 0.5) gets
    # confidence 0.1 (HPM.cu:717-722); without edges (texture 0.5) it keeps its own
    prm = O.params(len(cams), cams[0], geom_consistency=1, mand_consistency=1)
    for edges, conf, expect in ((True, 0.2, 0.1), (False, 0.2, 0.2), (True, 0.4, 0.4)):
        m = fresh(0, conf, 0.7)
        m["canny"] = np.zeros((H, W), np.uint32)
        if edges:
            m["canny"][py - 1, px - 1] = 255
        O.probe_pixel(0, cams, prm, m, px, py, draws=np.full(8, 0.5), ncc=lambda *a: 0.9,
                      geom=lambda *a: 0.0)
        assert m["conf"][py, px] == pytest.approx(expect), (edges, conf)


# ------------------------------------------------------------------ mandatory consistency

def test_mandatory_consistency_adoption_and_confidence_hand_over():
    """An unconfident pixel (confidence < 0.3) adopts the direction whose
    restricted cost is largest: (c_q > 0 ? max(3, m) : m) exp(-fc^2 / 0.18) with
    m = 0.5 + exp(-max(0, M - c_q)^2 / M^2), M = 0.5 x selected views; the
    adopted cost is fc x min(exp(-2 c_q), 0.35) and the pixel's confidence
    becomes half the source's (HPM.cu:2704-2773)."""
    W, H, V = 40, 40, 2
    cams = cams_for(W, H, V)
    prm = O.params(len(cams), cams[0], geom_consistency=1, mand_consistency=1)
    W_, H_, V_, px, py, costs, first = _vs_setup(W, H, V)
    rows = np.full((W * H, V), 0.4, np.float32)
    rows[first[3]] = [0.5, 0.6]
    tab = PlaneTable(W, H, V, rows)
    maps = _maps(tab, costs)
    conf = np.zeros((H, W), np.float32)  # the pixel itself: confidence 0
    conf.flat[first[3]] = 0.8  # down_far's pixel is confident
    maps["conf"] = conf
    cur = maps["planes"][py, px].copy()
    draws = np.concatenate([np.full(7, 0.2), np.full(8, 0.9), np.full(20, 0.5)]).astype(np.float32)
    r = O.probe_pixel(1, cams, prm, maps, px, py, 0, draws=draws, ncc=tab.ncc, geom=lambda *a: 0.5)
    w = r["weights"][:V].astype(np.float64)
    nsel = int(np.sum(w > 0))
    M = 0.5 * nsel
    fc_expect, rfc_expect = [], []
    for d in range(8):
        q = r["positions"][d]
        fc = np.sum(w * (rows[q] + 0.2 * 0.5)) / w.sum()
        cq = float(conf.flat[q])
        m = 0.5 + math.exp(-max(0.0, M - cq) ** 2 / (M * M))
        rfc_expect.append((max(3.0, m) if cq > 0 else m) * math.exp(-fc * fc / 0.18))
        fc_expect.append(fc * min(math.exp(-2 * cq), 0.35) if cq > 0 else fc)
    np.testing.assert_allclose(r["rfc"], rfc_expect, rtol=2e-6)
    np.testing.assert_allclose(r["fc"], fc_expect, rtol=2e-6)
    assert int(np.argmax(r["rfc"])) == 3
    np.testing.assert_array_equal(maps["planes"][py, px], tab.planes.reshape(-1, 4)[first[3]])
    assert maps["conf"][py, px] == f32(0.5 * 0.8)
    assert maps["costs"][py, px] == pytest.approx(fc_expect[3], rel=2e-6)
    assert maps["selected"][py, px] == r["sel"]
    # nothing beats the pixel's own restricted cost: it keeps its plane and confidence, and
    # its cost becomes cost_now x exp(0.2 + 2 c) (HPM.cu:2752-2757)
    conf.flat[first[3]] = 0.0
    conf[py, px] = 0.1
    maps = _maps(tab, costs)
    maps["conf"] = conf.copy()
    rows_cur = rows.copy()
    tab2 = PlaneTable(W, H, V, rows_cur)
    q0 = py * W + px
    rows_cur[q0] = [0.05, 0.05]
    maps["planes"] = tab2.planes.copy()
    r = O.probe_pixel(1, cams, prm, maps, px, py, 0, draws=draws, ncc=tab2.ncc, geom=lambda *a: 0.5)
    w = r["weights"][:V].astype(np.float64)
    cost_now = np.sum(w * (0.05 + 0.2 * 0.5)) / w.sum()
    np.testing.assert_array_equal(maps["planes"][py, px], tab2.planes[py, px])
    assert maps["conf"][py, px] == f32(0.1)
    assert maps["costs"][py, px] == pytest.approx(cost_now * math.exp(0.2 + 2 * 0.1), rel=2e-6)
    del cur


# ------------------------------------------------------------------ planar prior (main.cpp:324-441)

def test_prior_plane_through_three_points():
    """GetPriorPlaneParams (HPM.cpp:1134-1167): the plane n . X + w = 0 through
    three points, |n| = 1 and w >= 0 whatever their order."""
    rng = np.random.default_rng(8)
    for _ in range(200):
        n = rng.normal(size=3); n /= np.linalg.norm(n)
        w = rng.uniform(0.5, 5.0)
        # three points on the plane: a point on it plus two in-plane directions
        p0 = -w * n
        a = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0]); a /= np.linalg.norm(a)
        b = np.cross(n, a)
        pts = [p0 + rng.uniform(-1, 1) * a + rng.uniform(-1, 1) * b for _ in range(3)]
        for order in ((0, 1, 2), (2, 1, 0), (1, 2, 0)):
            X = [np.ascontiguousarray(pts[k], np.float32) for k in order]
            out = np.zeros(4, np.float32)
            O.lib().orc_prior_plane(X[0].ctypes.data, X[1].ctypes.data, X[2].ctypes.data, out.ctypes.data)
            assert out[3] >= 0
            np.testing.assert_allclose(out, np.r_[n, w], atol=2e-5 * max(1.0, w))


def test_support_points_by_blocks():
    """GetSupportPoints_Classify_Check (HPM.cpp:1026-1073), restated per block
    with numpy: over 5 x 5 blocks in column-major order, among pixels with
    cost < 2 the first (column by column) minimum of cost + 0.2 - conf and of
    cost - conf, each kept below 0.1, the first one first."""
    rng = np.random.default_rng(2)
    h, w = 23, 31
    costs = rng.uniform(-0.3, 2.5, (h, w)).astype(np.float32)
    costs[rng.random((h, w)) < 0.2] = 2.0
    conf = np.where(rng.random((h, w)) < 0.5, rng.uniform(0, 1, (h, w)), 0).astype(np.float32)
    costs[5:10, 10:15] = 0.05  # one block of ties: its first pixel wins
    conf[5:10, 10:15] = 0.0
    xy = np.zeros((2 * 7 * 5, 2), np.int32)
    cnt = ctypes.c_int(0)
    O.lib().orc_support_points(costs.ctypes.data, conf.ctypes.data, w, h, xy.ctypes.data, len(xy), ctypes.byref(cnt))
    expect = []
    for c0 in range(0, w, 5):
        for r0 in range(0, h, 5):
            cand = [(c, r) for c in range(c0, min(w, c0 + 5)) for r in range(r0, min(h, r0 + 5))
                    if costs[r, c] < 2.0]
            if not cand:
                continue
            shifted = [np.float32(np.float64(costs[r, c]) + 0.2 - np.float64(conf[r, c])) for c, r in cand]
            plain = [np.float32(costs[r, c] - conf[r, c]) for c, r in cand]
            k1, k2 = int(np.argmin(shifted)), int(np.argmin(plain))
            if shifted[k1] < 0.1:
                expect.append(cand[k1])
            if plain[k2] < 0.1:
                expect.append(cand[k2])
    assert [tuple(p) for p in xy[:cnt.value]] == expect
    assert (10, 5) in expect  # the tie block: its first pixel


def test_prior_level_raster_planes_and_depth_test():
    """main.cpp:385-431 on a fronto-parallel depth map at Z: every pixel the
    barycentric walk reaches gets the triangle's index and the plane
    (0, 0, -1, Z); a triangle with a corner outside the image is skipped and
    takes no index; the depth test against [0.6 dmin, 1.2 dmax] drops a
    triangle whose plane lies outside it."""
    W, H, Z = 40, 30, 3.0
    cam = pinhole(W, H, 50.0)
    depth = np.full((H, W), Z, np.float32)
    tris = np.array([[2, 2, 30, 4, 10, 25], [0, 0, 45, 0, 0, 10], [30, 4, 38, 28, 10, 25]], np.int32)
    prior = np.zeros((H, W, 4), np.float32)
    masks = np.zeros((H, W), np.uint32)
    O.lib().orc_prior_level(ctypes.byref(cam), depth.ctypes.data, tris.ctypes.data, len(tris), prior.ctypes.data,
                            masks.ctypes.data)
    assert set(np.unique(masks)) == {0, 1, 2}  # the second triangle leaves the image: index 2 is the third
    assert masks[10, 12] == 1 and masks[20, 33] == 2 and masks[28, 2] == 0
    for (y, x) in zip(*np.nonzero(masks)):
        np.testing.assert_allclose(prior[y, x], [0, 0, -1, Z], atol=1e-6)
    assert np.all(prior[masks == 0] == 0)
    # the corners of a triangle are rasterised (p = 1, q = 1 - p, 1 - p - q = 1)
    assert masks[2, 2] == 1 and masks[25, 10] >= 1
    far = pinhole(W, H, 50.0, dmin=6.0, dmax=9.0)  # range [3.6, 10.8]: Z = 3 fails the test
    O.lib().orc_prior_level(ctypes.byref(far), depth.ctypes.data, tris.ctypes.data, len(tris), prior.ctypes.data,
                            masks.ctypes.data)
    assert not masks.any() and not prior.any()
