"""The GJ8 corrector's fixed-point exit (ogp_set_fixed_point_exit): a step of a 64-object group
ends its corrector evaluations once the state to evaluate next repeats the state just evaluated,
bit for bit, for every object that stores the step. The reference always runs n_corrections
evaluations (gpu_propagator.cu:102-117); every further evaluation would repeat the last one
exactly, so each result here must equal the all-evaluations result bit for bit, for every launch
shape, precision, source set and entry point that runs the GJ8 chunks, while the evaluation count
shows that evaluations were in fact skipped.
"""
import numpy as np
import pytest

import opengpop_amd as ogp
from opengpop_amd import scenarios

pytestmark = pytest.mark.gpu

SHAPES = ((1, 1), (2, 1), (4, 1), (8, 1), (4, 2), (8, 2))


def _run(c, d, steps, on):
    c.set_fixed_point_exit(on)
    r, v = c.propagate(d["r"], d["v"], d["step"], steps)
    return r, v, c.timing()


@pytest.mark.parametrize("precision,degree", [("fp64", 20), ("fp64", 70), ("split", 70)])
def test_fixed_point_exit_bit_identical_every_shape(precision, degree):
    # 330 mixed objects (LEO at 30 s, MEO and GEO at 300 s) with ragged spans up to 160 steps,
    # so groups hold objects that end inside chunks next to objects that run on
    print(f"[fixed-point] {precision} degree {degree}", flush=True)
    d = scenarios.mixed(330, seed=41)
    steps = 5 + (np.arange(330) * 53) % 156
    with ogp.Context(degree=degree, precision=precision) as c:
        c.set_tuning(1, 64)
        r_all, v_all, t_all = _run(c, d, steps, False)
        gj = t_all["gj_object_steps"]
        assert gj > 0 and t_all["evaluations"] == 7 * gj
        for w, g in SHAPES:
            c.set_tuning(w, 64)
            c.set_lockstep(g)
            for on in (True, False):
                r, v, t = _run(c, d, steps, on)
                assert np.array_equal(r, r_all) and np.array_equal(v, v_all), (w, g, on)
                if on and g == 1:
                    # typically 2-4 evaluations of 7 (lockstep groups always run 7)
                    assert t["evaluations"] < 0.7 * 7 * gj, (w, t["evaluations"] / gj)
                else:
                    assert t["evaluations"] == 7 * gj, (w, g, on)
        c.set_lockstep(0)


@pytest.mark.parametrize("drag", [True, False])
def test_fixed_point_exit_with_sources_bit_identical(drag):
    # Sun, Moon and (per-object) drag: each evaluation is corrected, and a step ends when the
    # corrected position repeats the evaluated one, and the velocity too when drag reads it
    print(f"[fixed-point] sources, drag {drag}", flush=True)
    d = scenarios.mixed(200, seed=43)
    steps = 5 + (np.arange(200) * 29) % 120
    bc = np.linspace(0.002, 0.03, 200)
    with ogp.Context(degree=20) as c:
        c.set_sources(ogp.SOURCE_SUN | ogp.SOURCE_MOON | (ogp.SOURCE_DRAG if drag else 0), 0.01)
        if drag:
            c.set_drag_coefficients(bc)
        for w in (1, 8):
            c.set_tuning(w, 32)
            r0, v0, t0 = _run(c, d, steps, False)
            r1, v1, t1 = _run(c, d, steps, True)
            assert np.array_equal(r0, r1) and np.array_equal(v0, v1), w
            assert t1["evaluations"] < t0["evaluations"] == 7 * t0["gj_object_steps"]


def test_fixed_point_exit_conjunction_and_stepwise():
    # the conjunction search's trajectory and candidates, and the stepwise propagator
    print("[fixed-point] conjunction and stepwise", flush=True)
    r, v = scenarios.leo(300, seed=45)
    r[150:] = r[:150] + 500.0  # close neighbours, so the candidate set is not empty
    v[150:] = v[:150]
    sets = np.r_[np.zeros(150, np.uint32), np.ones(150, np.uint32)]
    out = {}
    with ogp.Context(degree=20) as c:
        for on in (False, True):
            c.set_fixed_point_exit(on)
            ca = ogp.ClosestApproach(c, 60.0)
            ca.initialize(r, v, sets)
            ca.propagate(90)
            ca.propagate(47)
            res = ca.find_closest(5000.0)
            p = ogp.GpuPropagator(c, r[:100], v[:100], 60.0, 40)
            for _ in range(40):
                p.step()
            sr, sv = p.copy_state_to_host()
            p.close()
            out[on] = (ca.trajectory(), res.a, res.b, res.step, res.distance, res.time, sr, sv)
    assert len(out[True][1]) > 0
    for x, y in zip(out[False], out[True]):
        assert np.array_equal(x, y)


def test_fixed_point_exit_derivatives_bit_identical():
    # the derivative record pass stores the fixed point's Jacobian for the skipped iterates: the
    # adjoint Jacobian (whole-span record and checkpointed), the VJP and the STM are unchanged
    print("[fixed-point] derivatives", flush=True)
    r0, v0 = scenarios.leo(150, seed=47)
    steps = np.r_[np.full(75, 300), 5 + (np.arange(75) * 7) % 290]
    rng = np.random.default_rng(5)
    gr, gv = rng.normal(size=(150, 3)), rng.normal(size=(150, 3))
    out = {}
    with ogp.Context(degree=20) as c:
        for on in (False, True):
            c.set_fixed_point_exit(on)
            out[on] = (c.adjoint_jacobian(r0, v0, 60.0, steps),
                       c.adjoint_jacobian(r0, v0, 60.0, steps, checkpoint_steps=64),
                       c.adjoint_vjp(r0, v0, 60.0, steps, gr, gv),
                       c.stm(r0, v0, 60.0, steps))
    for x, y in zip(out[False], out[True]):
        for a, b in zip(x, y):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("force", ["gravity", "sun-moon", "sources"])
def test_fixed_point_exit_startup_bit_identical(force):
    # GpuStartup::Startup's work areas (nodes, s, S and the acceleration matrix): the startup
    # corrections stop once the corrected nodes repeat the evaluated ones, with the same results
    print(f"[fixed-point] startup {force}", flush=True)
    d = scenarios.mixed(150, seed=49)
    out = {}
    with ogp.Context(degree=20) as c:
        if force == "sources":
            c.set_sources(ogp.SOURCE_SUN | ogp.SOURCE_MOON | ogp.SOURCE_DRAG, 0.02)
        elif force == "sun-moon":
            c.set_sources(ogp.SOURCE_SUN | ogp.SOURCE_MOON, 0.0)
        for on in (False, True):
            c.set_fixed_point_exit(on)
            res = [c.startup(d["r"], d["v"], h) for h in (30.0, 300.0)]
            out[on] = [res[j][k] for j in range(2) for k in sorted(res[j])]
    for x, y in zip(out[False], out[True]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_register_budgets_bit_identical(precision):
    # launches of more than two 8-wave groups per CU take the kernel instance compiled for 6
    # waves per SIMD (fewer registers, a few spills), smaller launches the one for 4: 40,000
    # objects (625 groups) as one slice and as four slices of ~156 groups, and at W = 4
    print(f"[register budgets] {precision}", flush=True)
    d = scenarios.mixed(40_000, seed=57)
    steps = 5 + (np.arange(40_000) * 7) % 36
    out = []
    with ogp.Context(degree=20, precision=precision) as c:
        for waves, slices in ((8, 1), (8, 4), (4, 1)):
            c.set_tuning(waves, 64)
            c.set_slices(slices)
            out.append(c.propagate(d["r"], d["v"], d["step"], steps))
    for r, v in out[1:]:
        assert np.array_equal(r, out[0][0]) and np.array_equal(v, out[0][1])
