"""Pins the oracle's RK8 startup, startup corrector and GJ8 predictor/corrector.

Known answers of the reference's own tests (integrator_prediction_test.cu, integrator_test.cu,
startup_test.cu, rk8_test.cu), polynomial exactness, the Berry & Healy table identities, RK8
order conditions and an analytic Kepler orbit (degree <= 1 field).
"""
import numpy as np
import pytest
from fractions import Fraction


def exact7(t):
    return t ** 7 - t ** 3


def exact8(t):
    return t ** 8 - t ** 7 + 2 * t ** 6 - 5 * t ** 5 - 10 * t ** 4 + t ** 3 - t ** 2


def damped(t):
    return np.exp(-t) * (np.sin(t) + np.cos(t))


def test_predictor_kat(oracle, golden):
    k = golden["predictor_kat"]
    r, v, _ = oracle.gj_predict(k["acc"], k["s"], k["S"], k["h"])
    assert np.allclose(r, k["r"], rtol=1e-7, atol=0)  # EXPECT_FLOAT_EQ
    assert np.allclose(v, k["v"], rtol=4 * 2.2e-16 * 8, atol=0)  # EXPECT_DOUBLE_EQ-level


def test_corrector_kat(oracle, golden):
    k = golden["corrector_kat"]
    r, v = oracle.gj_finalize_correction(k["r"], k["v"], k["acc"], k["s"], k["S"], k["h"])
    assert np.allclose(r, k["r_out"], rtol=1e-7, atol=0)
    assert np.allclose(v, k["v_out"], rtol=1e-7, atol=0)


def test_startup_s0_kat(oracle, po, golden):
    k = golden["startup_s0_kat"]
    o = oracle.startup_piece(0, po.EXP, k["h"], k["ramp"], k["ramp"])
    assert np.allclose(o["s"][12:15], k["s0"], rtol=1e-15, atol=0)
    assert np.allclose(o["S"][12:15], k["S0"], rtol=1e-7, atol=0)


def test_startup_correct_kat(oracle, po, golden):
    k = golden["startup_correct_kat"]
    ramp = golden["startup_s0_kat"]["ramp"]
    o = oracle.startup_piece(1, po.EXP, k["h"], ramp, ramp)
    assert np.allclose(o["s"].reshape(9, 3), np.array(k["s"])[:, None], rtol=1e-7, atol=0)
    assert np.allclose(o["S"].reshape(9, 3), np.array(k["S"])[:, None], rtol=1e-7, atol=0)


def test_startup_acc_matrix_kat(oracle, po, golden):
    k = golden["startup_acc_matrix_kat"]
    o = oracle.startup_piece(2, po.DAMPED, k["h"], np.repeat(k["pos"], 3), np.repeat(k["vel"], 3))
    assert np.array_equal(o["acc"].reshape(9, 3), np.repeat(np.array(k["acc"], float)[:, None], 3, 1))


def test_rk8_stationary(oracle, po):
    pos, vel = oracle.rk8_nodes(po.NONE, 1.0, [1, 1, 1], [0, 0, 0])
    assert np.array_equal(pos, np.ones((9, 3)))
    assert np.array_equal(vel, np.zeros((9, 3)))


@pytest.mark.parametrize("kind,exact", [(3, exact7), (4, exact8)])
def test_rk8_exact_for_time_only_polynomials(oracle, kind, exact):
    pos, _ = oracle.rk8_nodes(kind, 1.0, [0, 0, 0], [0, 0, 0])
    e = np.array([exact(i) for i in range(-4, 5)])
    assert np.abs(pos[:, 0] - e).max() <= 4 * np.spacing(np.abs(e)).max()


def test_rk8_not_exact_for_poly9(oracle, po):
    def exact9(t):
        return t ** 9 + exact8(t)
    pos, _ = oracle.rk8_nodes(po.POLY9, 1.0, [0, 0, 0], [0, 0, 0])
    e = np.array([exact9(i) for i in range(-4, 5)])
    assert np.abs(pos[:, 0] - e).max() > 1e-6


@pytest.mark.parametrize("kind,exact", [(3, exact7), (4, exact8)])
def test_startup_exact_for_polynomials(oracle, kind, exact):
    st = oracle.startup(kind, 1.0, [0, 0, 0], [0, 0, 0], n_corr=5)
    e = np.array([exact(i) for i in range(-4, 5)])
    # reference tolerances 1e-10 / 1e-13 (startup_test.cu:111-139) on values up to 9e4;
    # one ulp of the node values is allowed on top (summation order is not the reference's)
    tol = 1e-10 if kind == 3 else 1e-13
    assert np.all(np.abs(st["pos"][:, 0] - e) <= np.maximum(tol, 2 * np.spacing(np.abs(e))))


def test_startup_stationary_is_exact(oracle, po):
    st = oracle.startup(po.NONE, 1.0, [0, 0, 0], [0, 0, 0], n_corr=5)
    assert np.array_equal(st["pos"], np.zeros((9, 3)))


def test_startup_converges_for_damped_oscillator(oracle, po):
    prev = np.inf
    for nc in range(1, 5):
        st = oracle.startup(po.DAMPED, 0.1, [1, 1, 1], [0, 0, 0], n_corr=nc)
        e = damped(np.arange(-4, 5) * 0.1)
        err = np.linalg.norm(st["pos"] - e[:, None])
        assert err < prev
        prev = err


def test_stationary_and_constant_speed(oracle, po):
    r, v = oracle.propagate(po.NONE, [[0, 0, 0], [1, 1, 1]], [[0, 0, 0], [0, 0, 0]], 1.0, 10)
    assert np.array_equal(r, [[0, 0, 0], [1, 1, 1]]) and np.array_equal(v, np.zeros((2, 3)))
    for n in range(1, 11):
        r, v = oracle.propagate(po.NONE, [[0, 0, 0], [1, 1, 1]], [[1, 1, 1], [1, 1, 1]], 1.0, n)
        assert np.allclose(r, np.array([[0, 0, 0], [1, 1, 1]]) + n, rtol=4 * 2.2e-16, atol=4e-16)
        assert np.allclose(v, 1.0, rtol=4 * 2.2e-16, atol=0)


@pytest.mark.parametrize("kind,exact,tol", [(3, exact7, 1e-13), (4, exact8, 2e-13)])
def test_gj8_exact_for_polynomials(oracle, kind, exact, tol):
    err = 0.0
    for n in range(1, 21):
        r, _ = oracle.propagate(kind, [0, 0, 0], [0, 0, 0], 0.1, n, n_corr=1)
        err = max(err, np.linalg.norm(r[0] - exact(n * 0.1)))
    assert err < tol


def test_gj8_error_decreases_with_corrections(oracle, po):
    prev = np.inf
    for nc in range(1, 6):
        err = 0.0
        for n in range(1, 41):
            r, _ = oracle.propagate(po.DAMPED, [1, 1, 1], [0, 0, 0], 0.1, n, n_corr=nc)
            err = max(err, np.linalg.norm(r[0] - damped(n * 0.1)))
        assert err < prev
        prev = err


def _gj_tables():
    """The ordinate tables as exact fractions, parsed from the oracle source."""
    import os, re
    src = open(os.path.join(os.path.dirname(__file__), "..", "oracle", "oracle.cpp")).read()
    out = {}
    for name in ("kGJ", "kSA"):
        body = src[src.index(f"const double {name}[10][9]"):]
        body = body[body.index("{") + 1: body.index("};")]
        nums = re.findall(r"(-?\d+\.0)\s*/\s*(\d+\.0)", body.replace("0.0,", "0.0 / 1.0,"))
        vals = [Fraction(int(float(a)), int(float(b))) for a, b in nums]
        assert len(vals) == 90, name
        out[name] = [vals[9 * i: 9 * i + 9] for i in range(10)]
    return out


def test_ordinate_table_identities():
    t = _gj_tables()
    # Gauss-Jackson ordinate rows sum to 1/12 (double integration of a constant, h^2/12
    # offset of the summed form); summed-Adams rows sum to 0 except the predictor row (1/2).
    for i in range(10):
        assert sum(t["kGJ"][i]) == Fraction(1, 12), i
        assert sum(t["kSA"][i]) == (Fraction(1, 2) if i == 9 else 0), i
    # mid-corrector rows are mirror images (time reversal symmetry)
    for i in range(9):
        assert t["kGJ"][i] == list(reversed(t["kGJ"][8 - i]))
        assert t["kSA"][i] == [-x for x in reversed(t["kSA"][8 - i])]


def test_kepler_orbit_with_point_mass_field(po):
    o = po.Oracle(degree=1)
    gm = o.gm
    r0 = np.array([7000e3, 0, 0])
    v0 = np.array([0, np.sqrt(gm / 7000e3) * 1.01, 1000.0])
    r, _ = o.propagate(po.GRAVITY, r0, v0, 60.0, 360)
    assert np.linalg.norm(r[0] - _kepler(gm, r0, v0, 21600.0)) < 1e-3


def _kepler(gm, r0, v0, t):
    rn = np.linalg.norm(r0)
    a = 1 / (2 / rn - (v0 @ v0) / gm)
    n = np.sqrt(gm / a ** 3)
    h = np.cross(r0, v0)
    e = np.linalg.norm(np.cross(v0, h) / gm - r0 / rn)
    E0 = np.arctan2((r0 @ v0) / (e * np.sqrt(gm * a)), (1 - rn / a) / e)
    M = E0 - e * np.sin(E0) + n * t
    E = M
    for _ in range(60):
        E = E - (E - e * np.sin(E) - M) / (1 - e * np.cos(E))
    f = 1 - a / rn * (1 - np.cos(E - E0))
    g = t - (E - E0 - np.sin(E - E0)) / n
    return f * r0 + g * v0


def test_identical_objects_give_identical_results(oracle70):
    # force_model_test.cu:38-88 / propagator_test.cu:38-83
    r0 = np.tile([-3884872.0, -5749111.0, 14723.0], (4, 1))
    v0 = np.tile([-795.304, 573.772, 7518.179], (4, 1))
    r, v = oracle70.propagate(0, r0, v0, 30.0, 10, epoch=(2015, 5, 4, 22, 22, 2.0))
    assert all(np.array_equal(r[0], r[i]) and np.array_equal(v[0], v[i]) for i in range(4))


def test_startup_fixed_point_absorbs_the_reference_quirks(po):
    # The startup corrector iterates to a fixed point that does not depend on the RK8 guess:
    # with the reference's 7 corrections, RK8 row 7 (-234 vs -231) and the mirrored backward
    # stage times leave the 24 h state bit-identical; with one correction they do not.
    from opengpop_amd import scenarios
    oracle20 = po.Oracle(degree=20)
    d = scenarios.config0()
    r, v = d["r"][:16], d["v"][:16]
    ref = oracle20.propagate(po.GRAVITY, r, v, 60.0, 1440, n_corr=7, quirks=3)
    for q in (0, 1, 2):
        alt = oracle20.propagate(po.GRAVITY, r, v, 60.0, 1440, n_corr=7, quirks=q)
        assert np.array_equal(ref[0], alt[0]) and np.array_equal(ref[1], alt[1])
    one = oracle20.propagate(po.GRAVITY, r, v, 60.0, 1440, n_corr=1, quirks=3)
    alt = oracle20.propagate(po.GRAVITY, r, v, 60.0, 1440, n_corr=1, quirks=2)
    assert np.abs(one[0] - alt[0]).max() > 1e-2
