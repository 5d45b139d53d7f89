"""GPU parity of the acceleration sources added to gravity (the IAccelerationSource extension
point, iacceleration_source.cuh:19-33, summed by force_model.cu:73-86) against the oracle:
the force-model boundary (ogp_accel) at propagation nodes, whole propagations with the Sun, the
Moon and drag, the gravity-only default left bit-identical, and the error paths.
"""
import numpy as np
import pytest

import opengpop_amd as ogp
from opengpop_amd import scenarios

pytestmark = pytest.mark.gpu

ALL = ogp.SOURCE_SUN | ogp.SOURCE_MOON | ogp.SOURCE_DRAG
TT_MINUS_UTC = 69.184 / 86400.0  # 2025: 37 leap seconds + 32.184 s


@pytest.fixture(scope="module")
def ctx20():
    with ogp.Context(degree=20) as c:
        yield c


def _population(seed=3, n=24):
    d = scenarios.mixed(n, seed=seed)
    return d["r"], d["v"], d["step"]


def test_accel_sources_match_oracle(ctx20, po):
    o = po.Oracle(degree=None)
    r, v, _ = _population()
    step, node = 60.0, 7
    _, mjd = ctx20.time_nodes(step, 4, np.array([node * step]))
    sun, moon = o.sun_moon(mjd[0] + TT_MINUS_UTC)
    tb = np.concatenate([sun, moon])
    try:
        grav = ctx20.accel(r, step, node).reshape(-1, 3)
        for mask, bc in ((ogp.SOURCE_SUN, 0.0), (ogp.SOURCE_MOON, 0.0), (ogp.SOURCE_DRAG, 0.02),
                         (ALL, 0.01)):
            ctx20.set_sources(mask, bc)
            total = ctx20.accel(r, step, node, v=v).reshape(-1, 3)
            want = grav + np.array([o.sources_accel(mask, bc, tb, r[i], v[i])
                                    for i in range(len(r))])
            # gravity is identical in both calls; the sources agree to the rounding of the sum
            assert np.all(np.abs(total - want) <= 4 * np.spacing(np.abs(grav)) + 1e-18), mask
            assert np.any(total != grav)
    finally:
        ctx20.set_sources(0, 0.0)


def test_propagation_with_sources_matches_oracle(ctx20, po):
    r0, v0, step = _population(seed=8, n=40)
    steps = np.where(step > 100.0, 72, 360)  # 6 h
    o = po.Oracle(degree=20)
    try:
        ctx20.set_sources(ALL, 0.015)
        o.set_sources(ALL, 0.015)
        r, v = ctx20.propagate(r0, v0, step, steps)
        ro, vo = o.propagate(po.GRAVITY, r0, v0, step, steps)
    finally:
        ctx20.set_sources(0, 0.0)
        o.set_sources(0, 0.0)
    assert np.abs(r - ro).max() < 1e-3
    assert np.abs(v - vo).max() < 1e-6
    rg, _ = ctx20.propagate(r0, v0, step, steps)
    shift = np.linalg.norm(r - rg, axis=1)
    assert shift.min() > 1e-3  # every object feels the Sun and the Moon


def test_gravity_only_default_is_unchanged(ctx20):
    r0, v0, step = _population(seed=9, n=32)
    r1, v1 = ctx20.propagate(r0, v0, step, 120)
    ctx20.set_sources(ALL, 0.02)
    ctx20.propagate(r0, v0, step, 120)
    ctx20.set_sources(0, 0.02)
    assert ctx20.get_sources() == (0, 0.02)
    r2, v2 = ctx20.propagate(r0, v0, step, 120)
    assert np.array_equal(r1, r2) and np.array_equal(v1, v2)
    ctx20.set_sources(0, 0.0)


def test_drag_lowers_the_orbit(ctx20):
    # circular 300 km orbit for one day: drag (B = 0.02 m^2/kg) removes energy
    rad = 6678137.0
    vc = np.sqrt(3.986004415e14 / rad)
    r0 = np.array([[rad, 0.0, 0.0]])
    v0 = np.array([[0.0, vc * np.cos(0.5), vc * np.sin(0.5)]])
    try:
        rg, vg = ctx20.propagate(r0, v0, 60.0, 1440)
        ctx20.set_sources(ogp.SOURCE_DRAG, 0.02)
        rd, vd = ctx20.propagate(r0, v0, 60.0, 1440)
    finally:
        ctx20.set_sources(0, 0.0)
    e = lambda r, v: 0.5 * (v ** 2).sum() - 3.986004415e14 / np.linalg.norm(r)
    assert e(rd[0], vd[0]) < e(rg[0], vg[0])
    assert np.linalg.norm(rd - rg) > 100.0


def test_per_object_drag_coefficients(ctx20):
    # a batch with per-object C_D A / m equals, object by object, the same population propagated
    # with each coefficient as the single value (an object's arithmetic does not depend on the
    # others, whatever the launch shape); also through ogp_accel
    r0, v0, step = _population(seed=12, n=48)
    bc = np.array([0.004, 0.02, 0.0])[np.arange(48) % 3]
    try:
        ctx20.set_sources(ogp.SOURCE_SUN | ogp.SOURCE_DRAG, 0.5)
        ctx20.set_drag_coefficients(bc)
        r, v = ctx20.propagate(r0, v0, step, 90)
        a = ctx20.accel(r0, 60.0, 6, v=v0).reshape(-1, 3)
        ctx20.set_drag_coefficients(None)
        for value in (0.004, 0.02, 0.0):
            ctx20.set_sources(ogp.SOURCE_SUN | ogp.SOURCE_DRAG, value)
            rs, vs = ctx20.propagate(r0, v0, step, 90)
            as_ = ctx20.accel(r0, 60.0, 6, v=v0).reshape(-1, 3)
            sel = bc == value
            assert np.array_equal(r[sel], rs[sel]) and np.array_equal(v[sel], vs[sel]), value
            assert np.array_equal(a[sel], as_[sel]), value
        ctx20.set_drag_coefficients(bc)
        with pytest.raises(ogp.OgpError):
            ctx20.propagate(r0[:10], v0[:10], step[:10], 20)  # another object count
        with pytest.raises(ogp.OgpError):
            ctx20.set_drag_coefficients(-bc)
    finally:
        ctx20.set_drag_coefficients(None)
        ctx20.set_sources(0, 0.0)


def test_split_precision_with_sources(po):
    r0, v0, step = _population(seed=10, n=16)
    out = {}
    for prec in ("fp64", "split"):
        with ogp.Context(degree=20, precision=prec) as c:
            c.set_sources(ALL, 0.01)
            out[prec] = c.propagate(r0, v0, step, 120)
    assert np.abs(out["split"][0] - out["fp64"][0]).max() < 1e-2


def test_source_errors(ctx20):
    with pytest.raises(ogp.OgpError):
        ctx20.set_sources(8, 0.0)  # unknown bit
    with pytest.raises(ogp.OgpError):
        ctx20.set_sources(ogp.SOURCE_DRAG, -1.0)
    r0, v0, _ = _population(seed=11, n=4)
    try:
        ctx20.set_sources(ogp.SOURCE_DRAG, 0.01)
        with pytest.raises(ogp.OgpError):
            ctx20.accel(r0, 60.0, 5)  # drag needs the velocities
        ctx20.set_sources(ogp.SOURCE_SUN, 0.0)
        with pytest.raises(ogp.OgpError):
            ctx20.stm(r0, v0, 60.0, 10)  # derivatives linearise gravity alone
    finally:
        ctx20.set_sources(0, 0.0)
    with ogp.Context(force=ogp.FORCE_NONE) as c:
        with pytest.raises(ogp.OgpError):
            c.set_sources(ogp.SOURCE_SUN, 0.0)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_conjunction_trajectory_is_the_propagation_with_sources(precision):
    # the conjunction search stores the states the propagation produces, sources included
    # (its own copy of the context's sources and drag coefficients): after one phase of 40
    # steps every stored row equals a propagation of that many steps, bit for bit
    r0, v0 = scenarios.leo(96, seed=51)
    sets = np.repeat(np.arange(2, dtype=np.uint32), [32, 64])
    with ogp.Context(degree=20, precision=precision) as c:
        c.set_sources(ogp.SOURCE_SUN | ogp.SOURCE_MOON | ogp.SOURCE_DRAG, drag_bc=0.02)
        ca = ogp.ClosestApproach(c, 60.0)
        ca.initialize(r0, v0, sets)
        ca.propagate(40)
        traj = ca.trajectory()
        ends = {k: c.propagate(r0, v0, 60.0, k) for k in (1, 4, 5, 23, 40)}
    assert traj.shape[:2] == (40, 96)
    for k, (r, v) in ends.items():
        assert np.array_equal(traj[k - 1, :, :3], r) and np.array_equal(traj[k - 1, :, 3:6], v), k


def test_stepwise_with_sources_equals_bulk():
    # the stepwise propagator (its own copy of the drag coefficients) gives, after every step,
    # the states of a bulk propagation of that many steps, with all sources and per-object drag
    r0, v0 = scenarios.leo(40, seed=52)
    bc = np.linspace(0.0, 0.03, 40)
    with ogp.Context(degree=20) as c:
        c.set_sources(ogp.SOURCE_SUN | ogp.SOURCE_MOON | ogp.SOURCE_DRAG, 0.01)
        c.set_drag_coefficients(bc)
        p = ogp.GpuPropagator(c, r0, v0, 60.0, 12)
        steps = [p.step()[:2] for _ in range(12)]
        for k in (1, 4, 5, 12):
            r, v = c.propagate(r0, v0, 60.0, k)
            assert np.array_equal(steps[k - 1][0], r) and np.array_equal(steps[k - 1][1], v), k


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_sources_launch_shape_does_not_change_results(precision):
    # With 2..8 waves per group only the group's first wave integrates and adds the sources;
    # the result must equal the one-wave launch bit for bit. 200 objects = three full groups
    # and a partial one; per-object drag; spans end inside the second and third 64-step chunks.
    d = scenarios.mixed(200, seed=41)
    steps = np.minimum(d["steps"], 150)
    steps[::7] = 70
    bc = np.linspace(0.0, 0.03, 200)
    with ogp.Context(degree=20, precision=precision) as c:
        c.set_sources(ALL, 0.01)
        c.set_drag_coefficients(bc)
        out = []
        for w in (1, 2, 4, 8):
            c.set_tuning(w, 64)
            out.append(c.propagate(d["r"], d["v"], d["step"], steps))
        for r, v in out[1:]:
            assert np.array_equal(r, out[0][0]) and np.array_equal(v, out[0][1])
