"""Pins the oracle's time/frame chain (ref_sys_trans.cu, time_points.cu:26-207, types.cuh).

SOFA itself is not available offline; the chain is pinned by published SOFA test-suite known
answers (fundamental arguments, precession angles, nutation, sidereal time), by closed forms
(calendar dates, NTP epochs) and by invariants of the rotation matrices.
"""
import numpy as np


def test_ntp_to_mjd(oracle, golden):
    for ntp, mjd in golden["ntp_to_mjd"]["cases"]:
        assert abs(oracle.ntp_to_mjd(ntp) - mjd) < 1e-10


def test_calendar_to_mjd_closed_forms(oracle, golden):
    assert oracle.calendar_to_mjd(2000, 1, 1, 12, 0, 0.0) == 51544.5
    assert oracle.calendar_to_mjd(2003, 6, 1, 0, 0, 0.0) == golden["sofa"]["cal2jd_2003_06_01"]
    assert oracle.calendar_to_mjd(2025, 1, 1, 0, 0, 0.0) == 60676.0
    assert oracle.calendar_to_mjd(1858, 11, 17, 0, 0, 0.0) == 0.0  # MJD origin
    # seconds beyond a day roll over (the time nodes add offsets to the seconds field)
    assert abs(oracle.calendar_to_mjd(2025, 1, 1, 0, 0, 86400.0 * 3) - 60679.0) < 1e-9
    assert oracle.calendar_to_mjd(1582, 10, 10, 0, 0, 0.0) == -1.0  # no such date


def test_fundamental_arguments_vs_sofa(oracle, golden):
    s = golden["sofa"]
    F = oracle.ftable(s["fa_t"])
    assert np.abs(F - np.array(s["fa"])).max() < 1e-12


def test_precession_angles_vs_sofa_pfw06(oracle, golden):
    s = golden["sofa"]
    t = (s["pfw06_mjd_tt"] - 51544.5) / 36525.0
    eps, gam, phi, psi = oracle.precession(t)
    p = s["pfw06"]
    assert abs(eps - p["epsa"]) < 1e-15
    assert abs(gam - p["gamb"]) < 1e-16
    assert abs(phi - p["phib"]) < 1e-15
    assert abs(psi - p["psib"]) < 1e-15


def test_nutation_vs_sofa_nut06a(oracle70, golden):
    s = golden["sofa"]
    t = (s["nut06a_mjd_tt"] - 51544.5) / 36525.0
    dpsi, deps = oracle70.nutation(t)
    # IERS Table 5.3a/b series versus SOFA's nut06a: same theory, microarcsecond-level
    # differences in argument expressions and truncation.
    assert abs(dpsi - s["nut06a"][0]) < 1e-11
    assert abs(deps - s["nut06a"][1]) < 1e-11


def test_gast_vs_sofa_gst06a(oracle70, golden):
    s = golden["sofa"]
    mjd = s["gst06a_mjd"]
    t = (mjd - 51544.5) / 36525.0
    F = oracle70.ftable(t)
    dpsi, _ = oracle70.nutation(t)
    eps = oracle70.precession(t)[0]
    g = oracle70.gast(mjd, mjd, dpsi, eps, F)
    assert abs(g - s["gst06a"]) < 1e-10
    # GMST alone (dpsi = 0 and no complementary terms is not exposed); the equation of the
    # equinoxes is ~1e-5 rad, so GAST - GMST06 must be of that size and sign
    assert abs((g - s["gmst06"]) - (-0.8834195e-5)) < 1e-9


def test_rotation_matrices_sofa_convention(oracle):
    a = 0.45
    c, s = np.cos(a), np.sin(a)
    # column-major storage; SOFA iauRx: [[1,0,0],[0,c,s],[0,-s,c]]
    Rx = oracle.rot(0, a).reshape(3, 3).T
    assert np.allclose(Rx, [[1, 0, 0], [0, c, s], [0, -s, c]], atol=0)
    Ry = oracle.rot(1, a).reshape(3, 3).T
    assert np.allclose(Ry, [[c, 0, -s], [0, 1, 0], [s, 0, c]], atol=0)
    Rz = oracle.rot(2, a).reshape(3, 3).T
    assert np.allclose(Rz, [[c, s, 0], [-s, c, 0], [0, 0, 1]], atol=0)


def test_eop_interpolation_and_leap_seconds(oracle70):
    rec = oracle70.eop(60676.0)  # exact row of data/IAU2000A_EOP.txt (2025-01-01)
    arcs = 3600.0 * 180.0 / np.pi
    assert rec[0] == 60676.0
    assert abs(rec[1] - 0.144063 / arcs) < 1e-20
    assert abs(rec[2] - 0.305108 / arcs) < 1e-20
    assert rec[3] == 0.0462673
    mid = oracle70.eop(60676.25)
    nxt = oracle70.eop(60677.0)
    assert abs(mid[3] - (0.75 * rec[3] + 0.25 * nxt[3])) < 1e-15
    # clamping outside the table (types.cuh:86-92)
    assert oracle70.eop(1000.0)[0] == 41684.0
    assert oracle70.leap_seconds(60676.0) == 37.0
    assert oracle70.leap_seconds(41318.0) == 10.0


def _nodes_2000_2026(oracle70):
    """Time nodes on several dates 2000-2026 (inside the EOP table), 7 h apart within a date."""
    out = []
    for ymd in [(2000, 1, 1), (2006, 7, 15), (2013, 3, 20), (2019, 11, 2), (2025, 1, 1),
                (2026, 6, 30)]:
        offs = np.arange(0, 86400.0, 7 * 3600.0)
        M, mjd = oracle70.time_nodes(ymd + (0, 0), 0.0, offs)
        out += list(zip(M, mjd))
    return out


def test_time_node_x_axis_sits_at_earth_rotation_angle(oracle70):
    # The equinox-based chain Q Rz(-GAST) W must equal the CIO-based Q' Rz(-ERA) W: the image
    # of the ITRS x axis lies at the Earth Rotation Angle (IERS Conventions 2010 eq. 5.15) in
    # the GCRS equator, up to the CIO locator s and second-order CIP terms (< 1e-5 rad). A
    # GAST sign error, a wrong UT1, a missing equation of the equinoxes or a reversed product
    # order each move it by orders of magnitude more.
    for m, mjd_utc in _nodes_2000_2026(oracle70):
        A = m.reshape(3, 3).T
        mjd_ut1 = mjd_utc + oracle70.eop(mjd_utc)[3] / 86400.0
        du = mjd_ut1 - 51544.5
        era = 2 * np.pi * ((0.7790572732640 + 0.00273781191135448 * du + du % 1.0) % 1.0)
        ang = np.arctan2(A[1, 0], A[0, 0])
        d = (ang - era + np.pi) % (2 * np.pi) - np.pi
        assert abs(d) < 1e-5, (mjd_utc, d)


def test_time_node_pole_is_the_cip(oracle70):
    # The image of the ITRS z axis is the CIP (plus polar motion, ~1e-6 rad): its GCRS X, Y
    # follow the IAU 2006/2000A series (IERS Conventions 2010 eq. 5.16), here the polynomial
    # part and the two 18.6-year nutation terms, good to about 1 arcsec.
    asec = np.pi / 180.0 / 3600.0
    for m, mjd_utc in _nodes_2000_2026(oracle70):
        A = m.reshape(3, 3).T
        mjd_tt = mjd_utc + (oracle70.leap_seconds(mjd_utc) + 32.184) / 86400.0
        t = (mjd_tt - 51544.5) / 36525.0
        om = (450160.398036 - 6962890.5431 * t) * asec  # mean longitude of the lunar node
        X = (-0.016617 + 2004.191898 * t - 0.4297829 * t**2 - 6.8443 * np.sin(om)) * asec
        Y = (-0.006951 - 0.025896 * t - 22.4072747 * t**2 + 9.2052 * np.cos(om)) * asec
        assert abs(A[0, 2] - X) < 1e-5 and abs(A[1, 2] - Y) < 1e-5, (mjd_utc, A[0, 2] - X,
                                                                     A[1, 2] - Y)


def test_time_node_matrices_are_rotations(oracle70):
    offs = np.arange(0, 86400.0 * 2, 3600.0)
    M, mjd = oracle70.time_nodes((2025, 1, 1, 0, 0), 0.0, offs)
    for m in M:
        A = m.reshape(3, 3).T
        assert np.abs(A @ A.T - np.eye(3)).max() < 1e-15
        assert abs(np.linalg.det(A) - 1.0) < 1e-15
    assert np.allclose(mjd, 60676.0 + offs / 86400.0, rtol=0, atol=1e-9)
    # Earth rotation: consecutive hourly nodes rotate by ~15.04 deg about the z axis
    A0 = M[0].reshape(3, 3).T
    A1 = M[1].reshape(3, 3).T
    R = A1 @ A0.T
    ang = np.degrees(np.arctan2(R[1, 0], R[0, 0]))
    assert abs(ang - 15.041) < 0.01
