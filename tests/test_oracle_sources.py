"""Pins the oracle's acceleration sources beyond gravity (the IAccelerationSource extension point,
iacceleration_source.cuh:19-33, summed by force_model.cu:73-86; the reference implements gravity
only).

* Sun and Moon positions (Montenbruck and Gill 2000, eq. 3.43-3.49) against 2025 almanac events:
  the March equinox and June solstice, perihelion and aphelion distances, a new and a full moon,
  and the total lunar eclipse of 2025-03-14. Each event fixes a longitude to the series'
  ~0.1 deg accuracy, so a dropped or mis-signed large term fails.
* Point-mass accelerations against closed forms (zero at the geocentre, exact on the body line,
  the tidal tensor for small r).
* The exponential atmosphere: exact at the band bases, continuous across them (each band's base
  density is the previous band extrapolated to it), decreasing with altitude.
* Drag: zero for air at rest in the rotating frame, opposite to the relative velocity with
  magnitude 1/2 B rho v_r^2; over a propagated orbit the energy decreases.
"""
import datetime

import numpy as np
import pytest

EPS0 = np.radians(23.43929111)
TT_UTC_2025 = 69.184  # s: 37 leap seconds + 32.184


def mjd_tt(y, mo, d, h, mi):
    t = datetime.datetime(y, mo, d, h, mi) - datetime.datetime(1858, 11, 17)
    return t.total_seconds() / 86400.0 + TT_UTC_2025 / 86400.0


def ecliptic(p):
    """J2000 equatorial -> (longitude, latitude) of the J2000 ecliptic, degrees."""
    x, y, z = p
    ye = np.cos(EPS0) * y + np.sin(EPS0) * z
    ze = -np.sin(EPS0) * y + np.cos(EPS0) * z
    return np.degrees(np.arctan2(ye, x)) % 360.0, np.degrees(np.arcsin(ze / np.linalg.norm(p)))


def dlon(a, b):
    return (a - b + 180.0) % 360.0 - 180.0


# precession of the equinox from J2000 to 2025.2-2025.5: 50.29"/yr x 25.3 yr = 0.353 deg
PREC = 0.353


@pytest.mark.parametrize("when, lon", [((2025, 3, 20, 9, 1), 0.0), ((2025, 6, 21, 2, 42), 90.0)])
def test_sun_equinox_and_solstice(oracle, when, lon):
    sun, _ = oracle.sun_moon(mjd_tt(*when))
    lam, beta = ecliptic(sun)
    assert abs(dlon(lam, lon - PREC)) < 0.15
    assert abs(beta) < 1e-9


@pytest.mark.parametrize("when, dist_km", [((2025, 1, 4, 13, 28), 147_103_000.0),
                                           ((2025, 7, 3, 19, 55), 152_087_000.0)])
def test_sun_perihelion_aphelion(oracle, when, dist_km):
    sun, _ = oracle.sun_moon(mjd_tt(*when))
    assert abs(np.linalg.norm(sun) / 1e3 / dist_km - 1.0) < 5e-4


@pytest.mark.parametrize("when, elong, max_lat", [((2025, 1, 29, 12, 36), 0.0, 5.3),
                                                   ((2025, 2, 12, 13, 53), 180.0, 5.3),
                                                   ((2025, 3, 14, 6, 58), 180.0, 0.45)])
def test_moon_phases_and_eclipse(oracle, when, elong, max_lat):
    # new moon, full moon, greatest total lunar eclipse: the Moon's ecliptic longitude is the
    # Sun's (+180 deg); in the eclipse the Moon is also within the Earth's shadow (|beta| small)
    sun, moon = oracle.sun_moon(mjd_tt(*when))
    ls, _ = ecliptic(sun)
    lm, bm = ecliptic(moon)
    assert abs(dlon(lm, ls + elong)) < 0.3
    assert abs(bm) < max_lat
    assert 356_000e3 < np.linalg.norm(moon) < 407_000e3


def test_point_mass_closed_forms(oracle):
    s = np.array([1.2e11, -8.0e10, 3.0e10])  # "Sun" position
    tb = np.concatenate([s, [0.0, 0.0, 0.0]])
    v = np.zeros(3)
    assert np.array_equal(oracle.sources_accel(1, 0.0, tb, np.zeros(3), v), np.zeros(3))
    gm = 1.32712440018e20
    d = np.linalg.norm(s)
    u = s / d
    for x in (7.0e6, 4.2e7, -7.0e6):
        a = oracle.sources_accel(1, 0.0, tb, x * u, v)
        exact = gm * (1.0 / (d - x) ** 2 - 1.0 / d ** 2) * u
        assert np.allclose(a, exact, rtol=1e-6, atol=0)  # cancellation: |a| ~ 1e-7 |GM/d^2|
    # tidal tensor GM/d^3 (3 (u.r) u - r) for |r| << d
    r = np.array([3.0e6, 5.0e6, -2.0e6])
    a = oracle.sources_accel(1, 0.0, tb, r, v)
    tidal = gm / d ** 3 * (3.0 * (u @ r) * u - r)
    assert np.allclose(a, tidal, rtol=3 * np.linalg.norm(r) / d * 10, atol=0)
    # the Moon bit uses the second position and the Moon's GM
    tbm = np.concatenate([np.zeros(3), s / 300.0])
    am = oracle.sources_accel(2, 0.0, tbm, r, v)
    dm = d / 300.0
    tidal_m = 4.9028e12 / dm ** 3 * (3.0 * (u @ r) * u - r)
    assert np.allclose(am, tidal_m, rtol=0.05, atol=0)


def test_atmosphere_table(oracle):
    base = [0, 25, 30, 40, 50, 60, 70, 80, 90, 100, 110, 120, 130, 140, 150, 180, 200, 250, 300,
            350, 400, 450, 500, 600, 700, 800, 900, 1000]
    rho = [oracle.density(h * 1e3) for h in base]
    assert rho[0] == 1.225 and rho[-1] == 3.019e-15
    for i in range(1, len(base)):
        below = oracle.density(base[i] * 1e3 - 1e-3)  # 1 mm under the band base
        assert abs(below / rho[i] - 1.0) < 2e-3, base[i]
    hs = np.linspace(-1e3, 1.5e6, 3001)
    d = np.array([oracle.density(h) for h in hs])
    assert np.all(np.diff(d) < 0)


def test_drag(oracle):
    w = 7.292115e-5
    r = np.array([6778e3, 120e3, -40e3])
    v_corot = np.array([-w * r[1], w * r[0], 0.0])
    assert np.array_equal(oracle.sources_accel(4, 0.01, np.zeros(6), r, v_corot), np.zeros(3))
    v = np.array([-150.0, 7600.0, 900.0])
    bc = 0.012
    a = oracle.sources_accel(4, bc, np.zeros(6), r, v)
    vr = v - np.cross([0, 0, w], r)
    rho = oracle.density(np.linalg.norm(r) - 6378137.0)
    assert np.allclose(a, -0.5 * bc * rho * np.linalg.norm(vr) * vr, rtol=1e-14, atol=0)
    assert a @ vr < 0


def test_propagation_with_sources(po):
    o = po.Oracle(degree=8)
    try:
        r0 = np.array([[6700e3, 0.0, 0.0]])
        v0 = np.array([[0.0, 7713.0 * np.cos(0.9), 7713.0 * np.sin(0.9)]])
        r_g, v_g = o.propagate(po.GRAVITY, r0, v0, 30.0, 240)
        o.set_sources(1 | 2, 0.0)
        r_tb, _ = o.propagate(po.GRAVITY, r0, v0, 30.0, 240)
        o.set_sources(4, 0.02)
        r_d, v_d = o.propagate(po.GRAVITY, r0, v0, 30.0, 240)
    finally:
        o.set_sources(0, 0.0)
    # 2 h: Sun and Moon move a LEO by metres, drag at ~320 km by tens of metres along track,
    # and drag removes energy
    assert 0.1 < np.linalg.norm(r_tb - r_g) < 100.0
    assert np.linalg.norm(r_d - r_g) > 1.0
    gm = o.gm
    energy = lambda r, v: 0.5 * np.sum(v * v) - gm / np.linalg.norm(r)
    assert energy(r_d[0], v_d[0]) < energy(r_g[0], v_g[0])
