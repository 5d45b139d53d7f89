"""GPU parity of the conjunction path (ClosestApproach, closest_approach.cu:47-210) against the
oracle: the reference's analytic encounters, the sampled screen bit-exact on the same stored
trajectory, the Hermite/secant refinement, and the candidate pair set of a forced-encounter
population propagated independently on the GPU and by the oracle.
"""
import numpy as np
import pytest

import opengpop_amd as ogp
from opengpop_amd import scenarios

pytestmark = pytest.mark.gpu


def _gpu_phases(case):
    with ogp.Context(force=ogp.FORCE_NONE) as c:
        ca = ogp.ClosestApproach(c, case["h"])
        ca.initialize(case["pos"], case["vel"], case["sets"])
        out = []
        for n in case["phases"]:
            ca.propagate(n)
            out.append(ca.find_closest())
        return out


def test_analytic_encounters(golden):
    # closest_approach_test.cu:60-230: straight-line encounters with no acceleration
    for case in golden["encounters"]["cases"]:
        res = _gpu_phases(case)
        tol = case["tol"]
        if "expect_phases" in case:
            for out, (dist, tim) in zip(res, case["expect_phases"]):
                ap = out.get(0, 1)
                assert abs(ap["distance"] - dist) <= tol, case["name"]
                if tim is not None:
                    assert abs(ap["time"] - tim) <= tol, case["name"]
            continue
        for a, b, dist, tim in case["expect"]:
            ap = res[-1].get(a, b)
            if tol == 0:
                assert ap["distance"] == dist, case["name"]
            else:
                assert abs(ap["distance"] - dist) <= tol, case["name"]
            if tim is not None:
                assert abs(ap["time"] - tim) <= tol, case["name"]
            # Get(j, i) swaps the positions (types.cuh:153-169)
            sw = res[-1].get(b, a)
            assert sw["distance"] == ap["distance"]
            assert np.array_equal(sw["pos_a"], ap["pos_b"])


def test_get_invalid_and_same_object():
    case = dict(h=0.5, phases=[10], pos=[[0, 0, 0], [10, 0, 0], [20, 0, 0]],
                vel=[[1, 0, 0], [-1, 0, 0], [0, 0, 0]], sets=[0, 1, 1])
    out = _gpu_phases(case)[-1]
    assert out.get(1, 1)["distance"] == 0.0
    assert out.get(1, 2)["distance"] == -1.0  # same set: not a pair
    assert len(out) == 2


def test_reinitialise_replaces_the_search():
    # a second initialise with more objects starts a new search in the same context
    small = _population(128, seed=1)
    big = _population(640, seed=2)
    with ogp.Context(degree=20) as c:
        ca = ogp.ClosestApproach(c, 60.0)
        ca.initialize(*small)
        ca.propagate(40)
        ca.find_closest()
        ca.initialize(*big)
        ca.propagate(40)
        again = ca.screen()
    with ogp.Context(degree=20) as c:
        ca = ogp.ClosestApproach(c, 60.0)
        ca.initialize(*big)
        ca.propagate(40)
        fresh = ca.screen()
    assert np.array_equal(again[0], fresh[0]) and np.array_equal(again[1], fresh[1])


def _population(n=384, seed=7):
    r, v = scenarios.leo(n, seed=seed)
    sets = np.repeat(np.arange(3, dtype=np.uint32), [n // 4, n // 4, n - n // 2])
    return r, v, sets


@pytest.fixture(scope="module")
def screened():
    r, v, sets = _population()
    with ogp.Context(degree=20) as c:
        ca = ogp.ClosestApproach(c, 60.0)
        ca.initialize(r, v, sets)
        ca.propagate(150)
        ca.propagate(90)  # second phase continues the same trajectory
        traj = ca.trajectory()
        mind, mink = ca.screen()
        res = ca.find_closest()
        thr = ca.find_closest(threshold=500e3)
        yield dict(r=r, v=v, sets=sets, traj=traj, mind=mind, mink=mink, res=res, thr=thr)


def test_screen_bit_exact_on_same_trajectory(screened, po):
    o = po.Oracle(degree=20)
    mind, mink = o.screen(screened["traj"], screened["sets"])
    assert screened["traj"].shape[0] == 240
    assert np.array_equal(mind, screened["mind"])
    assert np.array_equal(mink, screened["mink"])


def test_screen_ties_bit_exact(po):
    # No forces. Objects 0-9 (set 0) and 160-179 (set 1) are static, so their distances repeat
    # exactly at every sample (the first sample must win); 160-169 are copies of 0-9 (distance
    # 0). The others share one velocity and sit ~10^7 m apart, so their distances move only by
    # the rounding of the integrator's position updates, an ulp or so per step: the roots tie
    # or nearly tie with the running minimum at many samples. 101 samples end in a partial
    # 8-step stage.
    rng = np.random.default_rng(11)
    n = 200
    u = rng.normal(size=(n, 3))
    r = u / np.linalg.norm(u, axis=1)[:, None] * rng.uniform(6.5e6, 7.5e6, (n, 1))
    v = np.tile([0.0, 7.5e3, 1.0e3], (n, 1))
    static = np.r_[0:10, 160:180]
    v[static] = 0.0
    r[160:170] = r[0:10]
    sets = np.repeat(np.arange(2, dtype=np.uint32), [70, 130])
    with ogp.Context(force=ogp.FORCE_NONE) as c:
        ca = ogp.ClosestApproach(c, 60.0)
        ca.initialize(r, v, sets)
        ca.propagate(101)
        traj = ca.trajectory()
        mind, mink = ca.screen()
    ref_d, ref_k = po.Oracle(degree=2).screen(traj, sets)
    assert np.array_equal(mind, ref_d)
    assert np.array_equal(mink, ref_k)
    idx = lambda a, b: a * 130 + (b - 70)  # pair order: (a in set 0, b in set 1)
    ss = [idx(a, b) for a in range(10) for b in range(160, 180)]
    assert np.all(mink[ss] == 0)
    assert np.all(mind[[idx(a, 160 + a) for a in range(10)]] == 0.0)
    # the near-ties are really there for the moving pairs, and minima fall in later stages
    p = traj[:, :, :3]
    mov = [a for a in range(10, 70)], [b for b in range(70, 200) if not 160 <= b < 180]
    d = np.sqrt(((p[:, mov[0], None, :] - p[:, None, mov[1], :]) ** 2).sum(-1)).reshape(len(p), -1)
    near = (d <= d.min(0) * (1 + 1e-15)).sum(0)
    assert (near >= 2).mean() > 0.5
    mk = mink[[idx(a, b) for a in mov[0] for b in mov[1]]]
    assert np.any(mk >= 8) and np.any(mk >= 96)


def test_refine_matches_oracle_on_same_trajectory(screened, po):
    o = po.Oracle(degree=20)
    res, sets, traj = screened["res"], screened["sets"], screened["traj"]
    assert len(res) == len(screened["mind"])
    rng = np.random.default_rng(3)
    for q in rng.choice(len(res), 400, replace=False):
        a, b = int(res.a[q]), int(res.b[q])
        ok, pa, pb = o.partition_pair(sets, int(o.partition_idx(sets, a, b)))
        assert ok and (pa, pb) == (a, b)
        ref = o.refine(traj, 60.0, a, b, int(res.step[q]), float(screened["mind"][
            o.partition_idx(sets, a, b)]))
        assert abs(res.distance[q] - ref["distance"]) <= 1e-6
        assert abs(res.time[q] - ref["time"]) <= 1e-6
        if 0 < res.step[q] < traj.shape[0] - 1:
            assert np.abs(res.pos_a[q] - ref["pos_a"]).max() <= 1e-6


def test_threshold_selects_sampled_minimum_below(screened):
    thr, mind = screened["thr"], screened["mind"]
    assert len(thr) == int((mind < 500e3).sum())
    assert np.all(np.diff(thr.a.astype(np.int64) * 10**6 + thr.b) > 0)  # sorted, unique


def test_candidate_pairs_bit_exact_vs_oracle_end_to_end(po):
    # forced encounters; GPU and oracle each propagate, screen at 10 km and refine
    d = scenarios.config4(n_random=360, n_pairs=40, n_setA=80, days=0.25)
    with ogp.Context(degree=20) as c:
        ca = ogp.ClosestApproach(c, d["step"])
        ca.initialize(d["r"], d["v"], d["sets"])
        ca.propagate(d["steps"])
        gpu = ca.find_closest(threshold=10e3)
    o = po.Oracle(degree=20)
    _, _, traj = o.propagate(po.GRAVITY, d["r"], d["v"], d["step"], d["steps"], trajectory=True)
    mind, mink = o.screen(traj, d["sets"])
    # no pair sits within 1 m of the threshold, so the propagation difference (< 1 mm) cannot
    # move a pair across it
    assert np.all(np.abs(mind - 10e3) > 1.0)
    cand = np.nonzero(mind < 10e3)[0]
    pairs = [o.partition_pair(d["sets"], int(q))[1:] for q in cand]  # pair order = (a, b) order
    assert len(pairs) >= 15
    assert [(int(a), int(b)) for a, b in zip(gpu.a, gpu.b)] == pairs
    assert np.array_equal(gpu.step, mink[cand])
    for (a, b), q in zip(pairs, cand):
        ref = o.refine(traj, d["step"], a, b, int(mink[q]), float(mind[q]))
        g = gpu.get(a, b)
        assert abs(g["distance"] - ref["distance"]) <= 1.0
        assert abs(g["time"] - ref["time"]) <= 1e-3


@pytest.mark.parametrize("sizes", [(1, 1199), (3, 597, 600), (5, 40, 1155), (70, 1130)])
def test_screen_tile_shapes_bit_identical(sizes, po):
    # square (64 x 64) and narrow (4 x 256) tiles on the same trajectory: every pair's sampled
    # minimum and step equal between the shapes and to the oracle's per-pair screen, and the
    # threshold search returns the same records; a one-object set, small sets and three sets
    n = sum(sizes)
    r, v = scenarios.leo(n, seed=31)
    r[1:4] = r[0] + [[50.0, 0, 0], [0, 80.0, 0], [0, 0, 120.0]]  # near pairs across the sets
    sets = np.repeat(np.arange(len(sizes), dtype=np.uint32), sizes)
    out = {}
    with ogp.Context(degree=20) as c:
        ca = ogp.ClosestApproach(c, 60.0)
        ca.initialize(r, v, sets)
        ca.propagate(77)  # ends in a partial stage of 8 steps
        traj = ca.trajectory()
        for shape in (1, 2):
            c.set_screen_shape(shape)
            assert c.get_screen_shape(sizes[0]) == shape
            out[shape] = (ca.screen(), ca.find_closest(threshold=2e6))
    (d1, k1), f1 = out[1]
    (d2, k2), f2 = out[2]
    assert np.array_equal(d1, d2) and np.array_equal(k1, k2)
    assert len(f1) == len(f2) > 0
    for a in ("a", "b", "step", "distance", "time", "pos_a", "pos_b"):
        assert np.array_equal(getattr(f1, a), getattr(f2, a)), a
    ref_d, ref_k = po.Oracle(degree=2).screen(traj, sets)
    assert np.array_equal(d1, ref_d) and np.array_equal(k1, ref_k)


@pytest.mark.parametrize("sizes", [(1, 1199), (3, 597, 600), (70, 1130), (600, 600)])
@pytest.mark.parametrize("threshold", [20e3, 500e3])
def test_radial_filter_same_candidates(sizes, threshold, po):
    # threshold screens with the radial filter skip the tiles of radially separated objects:
    # the candidates, steps and distances equal the oracle's per-pair screen of the same
    # trajectory below the threshold, and every record equals the unfiltered search's
    print(f"[radial filter] sets {sizes} threshold {threshold}", flush=True)
    n = sum(sizes)
    r, v = scenarios.leo(n, seed=53)
    j = sizes[0]  # near neighbours of object 0 at the start of set 1
    r[j:j + 3] = r[0] + [[50.0, 0, 0], [0, 8e3, 0], [0, 0, 15e3]]
    v[j:j + 3] = v[0]
    sets = np.repeat(np.arange(len(sizes), dtype=np.uint32), sizes)
    out, tim = {}, {}
    with ogp.Context(degree=20) as c:
        ca = ogp.ClosestApproach(c, 60.0)
        ca.initialize(r, v, sets)
        ca.propagate(77)
        traj = ca.trajectory()
        for mode in (2, 0, 1):  # always, never, automatic
            c.set_radial_filter(mode)
            out[mode] = ca.find_closest(threshold=threshold)
            tim[mode] = ca.timing()
    ref_d, ref_k = po.Oracle(degree=2).screen(traj, sets)
    pa, pb = [], []
    for s in range(len(sizes)):
        for i in range(sum(sizes[:s]), sum(sizes[:s + 1])):
            for j in range(sum(sizes[:s + 1]), n):
                pa.append(i)
                pb.append(j)
    sel = ref_d < threshold
    f = out[2]
    assert len(f) == int(sel.sum()) > 0
    assert np.array_equal(f.a, np.asarray(pa)[sel]) and np.array_equal(f.b, np.asarray(pb)[sel])
    assert np.array_equal(f.step, ref_k[sel])
    for a in ("a", "b", "step", "distance", "time", "pos_a", "pos_b"):
        assert np.array_equal(getattr(f, a), getattr(out[0], a)), a
    assert tim[0]["screened_pair_steps"] == tim[0]["pair_steps"]
    # automatic: too few pairs per object here for the sort to pay, so every tile is screened
    assert tim[1]["screened_pair_steps"] == tim[1]["pair_steps"]
    for a in ("a", "b", "step", "distance", "time", "pos_a", "pos_b"):
        assert np.array_equal(getattr(out[1], a), getattr(out[0], a)), a
    if threshold == 20e3:  # some tiles lie radially apart (few where a 64-row tile spans a set)
        assert tim[2]["screened_pair_steps"] < tim[2]["pair_steps"]


def test_screen_tuner_records_choice():
    # a one-object set against 2,047: the tuner times both shapes once and keeps one; the
    # results do not depend on it
    r, v = scenarios.leo(2048, seed=5)
    sets = np.ones(2048, np.uint32)
    sets[0] = 0
    with ogp.Context(degree=20) as c:
        ca = ogp.ClosestApproach(c, 60.0)
        ca.initialize(r, v, sets)
        ca.propagate(40)
        tuned = ca.screen()
        assert c.get_screen_shape(1) in (1, 2)
        c.set_screen_shape(1)
        square = ca.screen()
    assert np.array_equal(tuned[0], square[0]) and np.array_equal(tuned[1], square[1])
