"""GPU parity of the force model, the frames and the integrator against the FP64 oracle.

Bars: gravity within 1e-13 relative of the oracle (and the reference's golden vectors at 1e-10);
time-node matrices within 1e-13; integrator known answers at the reference's tolerances;
propagation final state within 1 mm / 1e-6 m/s of the oracle.
"""
import numpy as np
import pytest

import opengpop_amd as ogp
from opengpop_amd import scenarios

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx70():
    with ogp.Context(degree=70) as c:
        yield c


@pytest.fixture(scope="module")
def ctx20():
    with ogp.Context(degree=20) as c:
        yield c


def _rand_points(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return d * rng.uniform(6.5e6, 4.3e7, (n, 1))


def test_gravity_golden_vectors(ctx70, golden):
    g = golden["gravity_70x70"]
    a = ctx70.gravity_itrs(np.array(g["positions"]).reshape(10, 3))
    assert np.abs(a - np.array(g["accelerations"]).reshape(10, 3)).max() < g["tolerance"]


def test_gravity_matches_oracle(ctx70, oracle70):
    r = _rand_points(4096, 1)
    # include the poles and the equator exactly
    r[:4] = [[0, 0, 7.0e6], [0, 0, -7.2e6], [7.1e6, 0, 0], [0, -6.9e6, 0]]
    a = ctx70.gravity_itrs(r)
    ref = oracle70.gravity_itrs(r)
    rel = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < 1e-13


def test_gravity_degree_truncation_matches_oracle(ctx20, po):
    o = po.Oracle(degree=20)
    r = _rand_points(512, 2)
    rel = np.linalg.norm(ctx20.gravity_itrs(r) - o.gravity_itrs(r), axis=1) / np.linalg.norm(
        o.gravity_itrs(r), axis=1)
    assert rel.max() < 1e-13


def test_split_precision_gravity_error_bound(oracle70):
    with ogp.Context(degree=70, precision="split") as c:
        r = _rand_points(4096, 3)
        a = c.gravity_itrs(r)
    ref = oracle70.gravity_itrs(r)
    # degree >= 3 terms in FP32 with FP64 column sums: the error stays at the 1e-10 m/s^2
    # level (about 1e-11 relative) even at 130 km altitude, where degree-70 terms are largest
    err = np.linalg.norm(a - ref, axis=1)
    assert np.all(np.isfinite(a))
    assert err.max() < 5e-10
    assert np.median(err) < 1e-12


def test_split_precision_position_error_two_days():
    """Split precision against the plain FP64 GPU path: 2 days of LEO at 60 s."""
    r0, v0 = scenarios.leo(256, seed=2027)
    with ogp.Context(degree=70) as c:
        r64, _ = c.propagate(r0, v0, 60.0, 2880)
    with ogp.Context(degree=70, precision="split") as c:
        r32, _ = c.propagate(r0, v0, 60.0, 2880)
    e = np.linalg.norm(r32 - r64, axis=1)
    assert e.max() < 0.25


def test_split_precision_seven_days_within_a_metre_of_oracle(oracle70):
    # the split-precision bar: 7 days of LEO at 60 s (10,080 steps), degree 70, final
    # positions within 1 m of the FP64 oracle
    r0, v0 = scenarios.leo(64, seed=2031)
    with ogp.Context(degree=70, precision="split") as c:
        r, _ = c.propagate(r0, v0, 60.0, 10080)
    ro, _ = oracle70.propagate(0, r0, v0, 60.0, 10080, threads=16)
    assert np.linalg.norm(r - ro, axis=1).max() < 1.0


def test_time_nodes_match_oracle(ctx70, oracle70):
    offs = np.concatenate([np.arange(0, 86400.0 * 3, 977.0), [1e-3, 59.999]])
    M, mjd = ctx70.time_nodes(60.0, 4, offs)
    Mo, mjdo = oracle70.time_nodes(scenarios.EPOCH, 0.0 - 4 * 60.0, offs)
    assert np.abs(mjd - mjdo).max() == 0.0
    assert np.abs(M - Mo).max() < 1e-13


def test_predictor_and_corrector_kat(ctx70, golden):
    k = golden["predictor_kat"]
    r, v, _ = ctx70.kat_predict(k["acc"], k["s"], k["S"], k["h"])
    assert np.allclose(r, k["r"], rtol=1e-7, atol=0)
    assert np.allclose(v, k["v"], rtol=2e-15, atol=0)
    k = golden["corrector_kat"]
    r, v = ctx70.kat_finalize_correction(k["r"], k["v"], k["acc"], k["s"], k["S"], k["h"])
    assert np.allclose(r, k["r_out"], rtol=1e-7, atol=0)
    assert np.allclose(v, k["v_out"], rtol=1e-7, atol=0)


def _exact7(t):
    return t ** 7 - t ** 3


def _exact8(t):
    return t ** 8 - t ** 7 + 2 * t ** 6 - 5 * t ** 5 - 10 * t ** 4 + t ** 3 - t ** 2


@pytest.mark.parametrize("force,exact,tol", [(ogp.FORCE_POLY7, _exact7, 1e-13),
                                             (ogp.FORCE_POLY8, _exact8, 2e-13)])
def test_gj8_exact_for_polynomials(force, exact, tol):
    with ogp.Context(force=force, n_corrections=1) as c:
        err = 0.0
        for n in range(1, 21):
            r, _ = c.propagate([0, 0, 0], [0, 0, 0], 0.1, n)
            err = max(err, np.linalg.norm(r[0] - exact(n * 0.1)))
    assert err < tol


@pytest.mark.parametrize("force,exact,tol", [(ogp.FORCE_POLY7, _exact7, 1e-10),
                                             (ogp.FORCE_POLY8, _exact8, 1e-13)])
def test_startup_exact_for_polynomials(force, exact, tol):
    with ogp.Context(force=force, n_corrections=5) as c:
        st = c.startup([[0, 0, 0]], [[0, 0, 0]], 1.0)
    e = np.array([exact(i) for i in range(-4, 5)])
    assert np.all(np.abs(st["pos"][0, :, 0] - e) <= np.maximum(tol, 2 * np.spacing(np.abs(e))))


def test_stationary_and_constant_speed():
    with ogp.Context(force=ogp.FORCE_NONE) as c:
        p = ogp.GpuPropagator(c, [[0, 0, 0], [1, 1, 1]], [[1, 1, 1], [1, 1, 1]], 1.0, 10)
        for n in range(1, 11):
            r, v, _ = p.step()
            assert np.allclose(r, np.array([[0, 0, 0], [1, 1, 1]]) + n, rtol=4 * 2.2e-16, atol=4e-16)
            assert np.allclose(v, 1.0, rtol=4 * 2.2e-16, atol=0)
        r, v = c.propagate([[0, 0, 0], [1, 1, 1]], [[0, 0, 0], [0, 0, 0]], 1.0, 10)
        assert np.array_equal(r, [[0, 0, 0], [1, 1, 1]])


def test_gj8_error_decreases_with_corrections():
    def damped(t):
        return np.exp(-t) * (np.sin(t) + np.cos(t))
    prev = np.inf
    for nc in range(1, 6):
        with ogp.Context(force=ogp.FORCE_DAMPED, n_corrections=nc) as c:
            p = ogp.GpuPropagator(c, [1, 1, 1], [0, 0, 0], 0.1, 40)
            err = 0.0
            for n in range(1, 41):
                r, _, _ = p.step()
                err = max(err, np.linalg.norm(r[0] - damped(n * 0.1)))
        assert err < prev
        prev = err


def test_startup_matches_oracle_with_gravity(ctx70, oracle70):
    d = scenarios.config0()
    st = ctx70.startup(d["r"][:8], d["v"][:8], 60.0)
    for i in range(8):
        o = oracle70.startup(0, 60.0, d["r"][i], d["v"][i])
        assert np.abs(st["pos"][i] - o["pos"]).max() < 1e-6
        assert np.abs(st["vel"][i] - o["vel"]).max() < 1e-9
        assert np.abs(st["acc"][i] - o["acc"]).max() < 1e-12


def test_config0_matches_oracle(ctx20, po):
    """64 LEO objects, degree 20, 60 s, 6 h: final state within 1 mm and 1e-6 m/s."""
    d = scenarios.config0()
    r, v = ctx20.propagate(d["r"], d["v"], d["step"], d["steps"])
    o = po.Oracle(degree=20)
    ro, vo = o.propagate(0, d["r"], d["v"], d["step"], d["steps"])
    assert np.abs(r - ro).max() < 1e-3
    assert np.abs(v - vo).max() < 1e-6


def test_24h_degree70_matches_oracle(ctx70, oracle70):
    r0, v0 = scenarios.leo(6, seed=11)
    r, v = ctx70.propagate(r0, v0, 60.0, 1440)
    ro, vo = oracle70.propagate(0, r0, v0, 60.0, 1440)
    assert np.abs(r - ro).max() < 1e-3
    assert np.abs(v - vo).max() < 1e-6


def test_mixed_regimes_degree70_match_oracle(ctx70, oracle70):
    # the headline workload's shape at its parity bar: 256 objects, 80% LEO at 30 s and 20%
    # MEO/GEO at 300 s, spans U(6, 24) h, degree 70; about 4 x 10^5 object-steps of the oracle
    # on 16 threads
    d = scenarios.mixed(256, seed=17)
    r, v = ctx70.propagate(d["r"], d["v"], d["step"], d["steps"])
    ro, vo = oracle70.propagate(0, d["r"], d["v"], d["step"], d["steps"], threads=16)
    assert np.abs(r - ro).max() < 1e-3
    assert np.abs(v - vo).max() < 1e-6


def test_mixed_steps_and_uneven_spans_match_oracle(ctx20, po):
    d = scenarios.mixed(40, seed=5)
    steps = np.minimum(d["steps"], 300)  # keep the oracle run short
    steps[:3] = [0, 2, 4]                # spans that end inside the startup
    r, v = ctx20.propagate(d["r"], d["v"], d["step"], steps)
    o = po.Oracle(degree=20)
    ro, vo = o.propagate(0, d["r"], d["v"], d["step"], steps)
    assert np.abs(r - ro).max() < 1e-3
    assert np.abs(v - vo).max() < 1e-6
    assert np.array_equal(r[0], d["r"][0])


def test_stepwise_equals_bulk(ctx20):
    d = scenarios.config0()
    p = ogp.GpuPropagator(ctx20, d["r"][:16], d["v"][:16], 60.0, 30)
    for _ in range(30):
        rs, vs, _ = p.step()
    rb, vb = ctx20.propagate(d["r"][:16], d["v"][:16], 60.0, 30)
    assert np.array_equal(rs, rb) and np.array_equal(vs, vb)


def test_identical_objects_bit_identical(ctx70):
    r0 = np.tile([-3884872.0, -5749111.0, 14723.0], (8, 1))
    v0 = np.tile([-795.304, 573.772, 7518.179], (8, 1))
    r, v = ctx70.propagate(r0, v0, 30.0, 10)
    assert all(np.array_equal(r[0], r[i]) and np.array_equal(v[0], v[i]) for i in range(8))


def test_million_objects_match_subsets():
    # 1,048,600 objects (16,385 groups of 64, the last one partial): state offsets, grids and
    # tuning buckets at this size; an object's result does not depend on the batch it runs in
    n = 1_048_600
    r0, v0 = scenarios.leo(n, seed=41)
    with ogp.Context(degree=20) as c:
        r, v = c.propagate(r0, v0, 60.0, 12)
        head = c.propagate(r0[:640], v0[:640], 60.0, 12)
        tail = c.propagate(r0[-600:], v0[-600:], 60.0, 12)
    assert np.isfinite(r).all() and np.isfinite(v).all()
    assert np.array_equal(r[:640], head[0]) and np.array_equal(v[:640], head[1])
    assert np.array_equal(r[-600:], tail[0]) and np.array_equal(v[-600:], tail[1])


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_launch_shape_does_not_change_results(precision):
    # 300 objects = four full 64-object groups and a partial one; 140 steps = three chunks
    # with chunk_steps 64, the last one partial
    d = scenarios.mixed(300, seed=11)
    steps = np.minimum(d["steps"], 140)
    with ogp.Context(degree=70, precision=precision) as c:
        out = []
        for w in (1, 2, 4, 8):
            c.set_tuning(w, 64)
            out.append(c.propagate(d["r"], d["v"], d["step"], steps))
            assert c.get_tuning(len(steps))[0] == w
        for r, v in out[1:]:
            assert np.array_equal(r, out[0][0]) and np.array_equal(v, out[0][1])


@pytest.mark.parametrize("degree", [2, 3, 4, 5, 7, 9, 12, 33, 47])
def test_every_degree_shapes_and_oracle(degree, po):
    # The 8 column parts depend on the degree: below degree 7 some parts are empty, at low
    # degrees a part can be the zonal column alone or a single column, and the tables are read
    # one column past the last. At each degree the acceleration matches the oracle and every
    # launch shape gives the same bits.
    o = po.Oracle(degree=degree)
    r = _rand_points(256, 40 + degree)
    r[:2] = [[0, 0, 7.0e6], [7.1e6, 0, 0]]
    d = scenarios.mixed(200, seed=degree)
    steps = (np.arange(200) * 13) % 90
    with ogp.Context(degree=degree) as c:
        ref = o.gravity_itrs(r)
        rel = np.linalg.norm(c.gravity_itrs(r) - ref, axis=1) / np.linalg.norm(ref, axis=1)
        assert rel.max() < 1e-13
        out = []
        for w in (1, 2, 4, 8):
            c.set_tuning(w, 64)
            out.append(c.propagate(d["r"], d["v"], d["step"], steps))
        for rr, vv in out[1:]:
            assert np.array_equal(rr, out[0][0]) and np.array_equal(vv, out[0][1])


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("n", [330, 300])
def test_ragged_spans_every_wave_count(precision, n):
    # Spans 0 .. 140 spread over every group, so objects end in the startup, at every point of
    # the three 64-step chunks, and next to objects that run on. With 2..8 waves per group only
    # the first wave integrates, and the others must follow it step for step to the group's
    # last step; with two groups in lockstep per workgroup (330 objects: 6 groups, 300: 5, the
    # last workgroup then holding one group past the objects) every wave follows the
    # workgroup's last step. The result equals the one-wave launch bit for bit, and a span of 0
    # leaves the state unchanged.
    d = scenarios.mixed(n, seed=17)
    steps = (np.arange(n) * 37) % 141
    with ogp.Context(degree=20, precision=precision) as c:
        out = []
        for w, g in ((1, 1), (2, 1), (4, 1), (8, 1), (4, 2), (8, 2)):
            c.set_tuning(w, 64)
            c.set_lockstep(g)
            out.append(c.propagate(d["r"], d["v"], d["step"], steps))
            assert c.get_tuning(n)[0] == w and c.get_lockstep(n) == g
        for r, v in out[1:]:
            assert np.array_equal(r, out[0][0]) and np.array_equal(v, out[0][1])
        zero = steps == 0
        assert zero.any()
        assert np.array_equal(out[0][0][zero], d["r"][zero])
        assert np.array_equal(out[0][1][zero], d["v"][zero])


def test_concurrent_slices_do_not_change_results():
    # 900 objects in two step-size classes (720 at 30 s, 180 at 300 s) with uneven spans, cut
    # into 1..7 slices per class on concurrent streams; also with the autotuner's trial chunks
    # (which join the slices) and with the stored trajectory of the conjunction search
    d = scenarios.mixed(900, seed=13)
    steps = np.minimum(d["steps"], 150)
    with ogp.Context(degree=20) as c:
        c.set_tuning(1, 16)
        c.set_slices(1)
        r1, v1 = c.propagate(d["r"], d["v"], d["step"], steps)
        for s, w in ((2, 2), (3, 4), (7, 8), (3, 0)):
            c.set_slices(s)
            c.set_tuning(w, 16)
            r, v = c.propagate(d["r"], d["v"], d["step"], steps)
            assert np.array_equal(r, r1) and np.array_equal(v, v1), (s, w)
        traj = []
        for s in (1, 3):
            c.set_slices(s)
            ca = ogp.ClosestApproach(c, 30.0)
            ca.initialize(d["r"][:700], d["v"][:700], np.arange(700) >= 350)
            ca.propagate(40)
            ca.propagate(30)
            traj.append(np.array(ca.trajectory()))
        assert traj[0].shape == traj[1].shape and np.array_equal(traj[0], traj[1])
        with pytest.raises(ogp.OgpError):
            c.set_slices(17)


def test_late_started_class_matches_separate_runs():
    # 1,600 LEO at 30 s for 300 steps, 400 MEO/GEO at 300 s for 72-288 steps (under a third of
    # the LEO object-steps, so that class starts late, during the LEO class's last chunks) and
    # 40 objects at 120 s with only startup steps (never started late): every class gives the
    # bits of its own propagation
    d = scenarios.mixed(2000, seed=17)
    r0 = np.concatenate([d["r"], d["r"][:40]])
    v0 = np.concatenate([d["v"], d["v"][:40]])
    step = np.concatenate([d["step"], np.full(40, 120.0)])
    steps = np.concatenate([np.minimum(d["steps"], 300), np.arange(40) % 5])
    with ogp.Context(degree=20) as c:
        r, v = c.propagate(r0, v0, step, steps)
        for h in (30.0, 300.0, 120.0):
            m = step == h
            rs, vs = c.propagate(r0[m], v0[m], step[m], steps[m])
            assert np.array_equal(r[m], rs) and np.array_equal(v[m], vs), h


def test_autotuner_keeps_results_and_records_choice():
    d = scenarios.config0()
    # 640 objects (10 groups): only the widest shape can give every SIMD a wave, so it is taken
    # without trials; 16,640 objects (260 groups): 4 and 8 waves are timed against each other;
    # 33,280 objects (520 groups): 2, 4 and 8 waves
    for reps, choices in ((10, (8,)), (260, (4, 8)), (520, (2, 4, 8))):
        r0, v0 = np.tile(d["r"], (reps, 1)), np.tile(d["v"], (reps, 1))
        with ogp.Context(degree=20) as c:
            c.set_tuning(1, 16)
            r1, v1 = c.propagate(r0, v0, 60.0, 120)
            c.set_tuning(0, 16)  # autotune: 115 GJ steps = 8 chunks of 16
            ra, va = c.propagate(r0, v0, 60.0, 120)
            assert np.array_equal(r1, ra) and np.array_equal(v1, va)
            w, chunk = c.get_tuning(len(r0))
            assert w in choices and chunk == 16
            rb, vb = c.propagate(r0, v0, 60.0, 120)  # tuned table reused
            assert np.array_equal(r1, rb) and np.array_equal(v1, vb)


def test_errors_are_reported_not_fatal(ctx20):
    with pytest.raises(ogp.OgpError):
        ctx20.propagate([[7e6, 0, 0]], [[0, 7.5e3, 0]], -1.0, 10)
    with pytest.raises(ogp.OgpError):
        ctx20.propagate([[np.nan, 0, 0]], [[0, 7.5e3, 0]], 60.0, 10)
    with pytest.raises(ogp.OgpError):
        ogp.Context(degree=71)
    with pytest.raises(ogp.OgpError):
        ogp.Context(data_dir="/nonexistent")
    # the context is still usable afterwards
    r, _ = ctx20.propagate([[7e6, 0, 0]], [[0, 7.5e3, 0]], 60.0, 5)
    assert np.all(np.isfinite(r))


def test_device_resident_equals_host_arrays(ctx20):
    # ogp_propagate_device (SoA device arrays, the headline's path) and ogp_propagate (AoS
    # host arrays) give the same bits for a mixed population with per-object steps and spans
    torch = pytest.importorskip("torch")
    d = scenarios.mixed(300, seed=23)
    steps = np.minimum(d["steps"], 150)
    r, v = ctx20.propagate(d["r"], d["v"], d["step"], steps)
    dev = torch.device("cuda")
    R0 = torch.tensor(d["r"].T.copy(), device=dev)
    V0 = torch.tensor(d["v"].T.copy(), device=dev)
    Ro, Vo = torch.empty_like(R0), torch.empty_like(V0)
    torch.cuda.synchronize()
    ctx20.propagate_device(R0, V0, len(steps), d["step"], steps, Ro, Vo)
    assert np.array_equal(Ro.cpu().numpy().T, r) and np.array_equal(Vo.cpu().numpy().T, v)
