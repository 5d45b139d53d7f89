"""Seeded synthetic inputs for the propagation and conjunction workloads.

Orbital elements are drawn from the stated distributions and converted to FP64 Cartesian state
vectors with Kepler's equation (Newton iteration) and GM = 3.986004415e14 (data/JGM3.gfc);
the result is taken as GCRS. Epoch 2025-01-01T00:00:00 UTC.
"""
import numpy as np

GM = 3.986004415e14
EPOCH = (2025, 1, 1, 0, 0, 0.0)


def elements_to_cartesian(a, e, i, raan, argp, M):
    """Classical elements (m, -, rad) to position/velocity arrays of shape (n, 3)."""
    a, e, i, raan, argp, M = (np.asarray(x, dtype=np.float64) for x in (a, e, i, raan, argp, M))
    E = M.copy()
    for _ in range(50):
        E = E - (E - e * np.sin(E) - M) / (1.0 - e * np.cos(E))
    cE, sE = np.cos(E), np.sin(E)
    fac = np.sqrt(1.0 - e * e)
    xp = a * (cE - e)
    yp = a * fac * sE
    rr = a * (1.0 - e * cE)
    vx = -np.sqrt(GM * a) / rr * sE
    vy = np.sqrt(GM * a) / rr * fac * cE
    co, so = np.cos(raan), np.sin(raan)
    cw, sw = np.cos(argp), np.sin(argp)
    ci, si = np.cos(i), np.sin(i)
    P = np.stack([co * cw - so * sw * ci, so * cw + co * sw * ci, sw * si], axis=-1)
    Q = np.stack([-co * sw - so * cw * ci, -so * sw + co * cw * ci, cw * si], axis=-1)
    r = xp[:, None] * P + yp[:, None] * Q
    v = vx[:, None] * P + vy[:, None] * Q
    return r, v


def leo_elements(rng, n):
    """a U(6678, 7378) km, e U(0, 0.02), i U(0, 98) deg, angles U(0, 360) deg."""
    return (rng.uniform(6678e3, 7378e3, n), rng.uniform(0.0, 0.02, n),
            np.radians(rng.uniform(0.0, 98.0, n)), np.radians(rng.uniform(0.0, 360.0, n)),
            np.radians(rng.uniform(0.0, 360.0, n)), np.radians(rng.uniform(0.0, 360.0, n)))


def leo(n, seed=0):
    rng = np.random.default_rng(seed)
    return elements_to_cartesian(*leo_elements(rng, n))


def config0(seed=2025):
    """64 LEO objects, 6 h at 60 s (360 steps), degree 20."""
    r, v = leo(64, seed)
    return dict(r=r, v=v, step=np.full(64, 60.0), steps=np.full(64, 360, np.int64), degree=20)


def mixed(n, seed=2026):
    """Config 1 population: 80% LEO (30 s), 15% MEO and 5% GEO (300 s); spans U(6, 24) h
    rounded to whole steps."""
    rng = np.random.default_rng(seed)
    n_leo = int(round(0.80 * n))
    n_meo = int(round(0.15 * n))
    n_geo = n - n_leo - n_meo
    el = [np.concatenate(x) for x in zip(
        leo_elements(rng, n_leo),
        (rng.uniform(25500e3, 29600e3, n_meo), rng.uniform(0.0, 0.02, n_meo),
         np.radians(rng.uniform(50.0, 65.0, n_meo)), np.radians(rng.uniform(0, 360, n_meo)),
         np.radians(rng.uniform(0, 360, n_meo)), np.radians(rng.uniform(0, 360, n_meo))),
        (rng.uniform(42164e3 - 50e3, 42164e3 + 50e3, n_geo), rng.uniform(0.0, 0.001, n_geo),
         np.radians(rng.uniform(0.0, 1.0, n_geo)), np.radians(rng.uniform(0, 360, n_geo)),
         np.radians(rng.uniform(0, 360, n_geo)), np.radians(rng.uniform(0, 360, n_geo))))]
    r, v = elements_to_cartesian(*el)
    step = np.concatenate([np.full(n_leo, 30.0), np.full(n_meo + n_geo, 300.0)])
    span = rng.uniform(6 * 3600.0, 24 * 3600.0, n)
    steps = np.maximum(np.rint(span / step).astype(np.int64), 1)
    return dict(r=r, v=v, step=step, steps=steps, degree=70)


def config1(n=100_000, seed=2026):
    return mixed(n, seed)


def config2(n=50_000, days=7.0, seed=2027):
    """LEO, 60 s step, 7 days (10080 steps), degree 70 (split-precision comparison)."""
    r, v = leo(n, seed)
    steps = int(round(days * 86400.0 / 60.0))
    return dict(r=r, v=v, step=np.full(n, 60.0), steps=np.full(n, steps, np.int64), degree=70)


def config3(n=10_000, seed=2028):
    """LEO, 60 s step, 24 h (1440 steps), degree 20 (adjoint)."""
    r, v = leo(n, seed)
    return dict(r=r, v=v, step=np.full(n, 60.0), steps=np.full(n, 1440, np.int64), degree=20)


def config4(n_random=18_000, n_pairs=1_000, n_setA=2_000, days=3.0, seed=2029):
    """Conjunction scenario: n_random LEO objects plus n_pairs forced pairs whose second member
    copies the first's elements with da U(-2, 2) km, di U(-0.05, 0.05) deg, dM U(-0.5, 0.5) deg.
    The first members come first so that they fall in set 0 (the first n_setA objects); the
    second members are placed in set 1."""
    rng = np.random.default_rng(seed)
    first = leo_elements(rng, n_pairs)
    second = list(first)
    second[0] = first[0] + rng.uniform(-2e3, 2e3, n_pairs)
    second[2] = first[2] + np.radians(rng.uniform(-0.05, 0.05, n_pairs))
    second[5] = first[5] + np.radians(rng.uniform(-0.5, 0.5, n_pairs))
    rand = leo_elements(rng, n_random)
    n_fill_a = n_setA - n_pairs
    el = [np.concatenate([f, rnd[:n_fill_a], rnd[n_fill_a:], s])
          for f, rnd, s in zip(first, rand, second)]
    r, v = elements_to_cartesian(*el)
    n = len(r)
    sets = np.zeros(n, np.uint32)
    sets[n_setA:] = 1
    steps = int(round(days * 86400.0 / 60.0))
    return dict(r=r, v=v, sets=sets, step=60.0, steps=steps, degree=20)
