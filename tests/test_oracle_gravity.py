"""Pins the oracle's Gottlieb gravity (gravity_source_device.cu:63-239, gravity_source.cu:168-262).

* golden vectors of the reference (fm_gravity_source_test.cu:79-107) at the reference's 1e-10;
* a = grad U against an independent long-double textbook potential (normalized associated
  Legendre functions, cos/sin m*lambda directly), 5-point central differences;
* degree <= 1 reduces to the point mass; degree 2 zonal reduces to the closed-form J2 field.
"""
import numpy as np
import pytest


def test_golden_vectors_70x70(oracle70, golden):
    g = golden["gravity_70x70"]
    r = np.array(g["positions"]).reshape(10, 3)
    a = oracle70.gravity_itrs(r)
    assert np.abs(a - np.array(g["accelerations"]).reshape(10, 3)).max() < g["tolerance"]


def test_batched_equals_one_by_one(oracle70, golden):
    r = np.array(golden["gravity_70x70"]["positions"]).reshape(10, 3)
    a = oracle70.gravity_itrs(r)
    for i in range(10):
        assert np.array_equal(oracle70.gravity_itrs(r[i:i + 1])[0], a[i])


def _grad_fd(o, r, d=20.0):
    g = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = d
        U = o.potential_noncentral
        g[k] = (-U(r + 2 * e) + 8 * U(r + e) - 8 * U(r - e) + U(r - 2 * e)) / (12 * d)
    return g


def test_acceleration_is_gradient_of_independent_potential(oracle70, golden):
    r = np.array(golden["gravity_70x70"]["positions"]).reshape(10, 3)
    rng = np.random.default_rng(7)
    extra = rng.normal(size=(6, 3))
    extra = extra / np.linalg.norm(extra, axis=1)[:, None] * rng.uniform(6.6e6, 4.3e7, (6, 1))
    pts = np.vstack([r, extra])
    a = oracle70.gravity_itrs(pts)
    gm = oracle70.gm
    for p, ai in zip(pts, a):
        central = -gm * p / np.linalg.norm(p) ** 3
        assert np.abs(_grad_fd(oracle70, p) + central - ai).max() < 5e-13


@pytest.mark.parametrize("deg", [0, 1])
def test_low_degree_is_point_mass(po, deg):
    o = po.Oracle(degree=deg)
    r = np.array([[7.0e6, 1.0e6, -2.0e6], [-3e6, 4e6, 5e6]])
    a = o.gravity_itrs(r)
    exp = -o.gm * r / np.linalg.norm(r, axis=1)[:, None] ** 3
    assert np.allclose(a, exp, rtol=1e-15, atol=0)


def test_degree2_zonal_matches_closed_form_j2(po):
    o = po.Oracle(degree=2)
    gm, R = o.gm, o.radius
    c20 = -0.484169548456e-03  # data/JGM3.gfc, gfc 2 0
    j2 = -np.sqrt(5.0) * c20
    r = np.array([7.1e6, -1.3e6, 2.2e6])
    x, y, z = r
    rn = np.linalg.norm(r)
    f = 1.5 * j2 * gm * R ** 2 / rn ** 5
    aj2 = np.array([x * (5 * z * z / rn ** 2 - 1), y * (5 * z * z / rn ** 2 - 1),
                    z * (5 * z * z / rn ** 2 - 3)]) * f
    central = -gm * r / rn ** 3
    a = o.gravity_itrs(r[None])[0]
    # C21, S21, C22, S22 are also present at degree 2; check the zonal part through a
    # z-axis point and an equatorial point symmetric under the tesseral terms' sign flip.
    pole = np.array([[0.0, 0.0, 7.0e6], [0.0, 0.0, -7.0e6]])
    ap = o.gravity_itrs(pole)
    zp = 7.0e6
    exp_pole = -gm / zp ** 2 * (1 - 3.0 * j2 * (R / zp) ** 2)
    assert abs(ap[0, 2] - exp_pole) < 1e-13 * abs(exp_pole)
    assert abs(ap[1, 2] + exp_pole) < 1e-13 * abs(exp_pole)
    # full field minus the tesseral closed form equals the J2 closed form
    c21, s21 = -0.186987640000e-09, 0.119528010000e-08  # data/JGM3.gfc
    c22, s22 = 0.243926074866e-05, -0.140026639759e-05
    U = lambda p: _tess_potential(p, gm, R, c21, s21, c22, s22)
    at = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1.0
        at[k] = (U(r + e) - U(r - e)) / 2.0
    assert np.abs(a - central - aj2 - at).max() < 1e-9


def _tess_potential(p, gm, R, c21, s21, c22, s22):
    x, y, z = p
    rn = np.sqrt(x * x + y * y + z * z)
    sphi = z / rn
    cphi = np.sqrt(x * x + y * y) / rn
    lam = np.arctan2(y, x)
    p21 = np.sqrt(15.0) * sphi * cphi
    p22 = np.sqrt(15.0) / 2.0 * cphi ** 2
    return gm / rn * (R / rn) ** 2 * (p21 * (c21 * np.cos(lam) + s21 * np.sin(lam))
                                      + p22 * (c22 * np.cos(2 * lam) + s22 * np.sin(2 * lam)))


def test_degree_above_file_maximum_is_an_error(po):
    with pytest.raises(RuntimeError):
        po.Oracle(degree=71)
