"""Pins the oracle's conjunction path (partition_parameters.cuh, interpolation.cuh,
device.cu:56-173): the reference's partition known answers, Hermite exactness on polynomials,
the secant root finder and the analytic straight-line encounters of closest_approach_test.cu.
"""
import numpy as np
import pytest


def test_partition_kat(oracle, golden):
    k = golden["partition_kat"]
    p = oracle.partition(k["set_ids"])
    assert p["num_sets"] == 4
    assert p["set_offsets"] == k["set_offsets"]
    assert p["num_higher"] == k["num_higher"]
    assert p["map_offsets"] == k["map_offsets"]
    for bad in k["invalid"]:
        assert oracle.partition(bad) is None
    for a, b, idx in k["idx_cases"]:
        assert oracle.partition_idx(k["pairs_ids"], a, b) == idx
    for idx, a, b in k["pair_cases"]:
        assert oracle.partition_pair(k["pairs_ids"], idx) == (True, a, b)
    one = oracle.partition(k["single_set"])
    assert one["num_sets"] == 1 and one["num_pairs"] == 0
    for a, b in k["invalid_pairs"]:
        assert oracle.partition_idx(k["set_ids"], a, b) == -1
    assert oracle.partition_pair(k["set_ids"], k["invalid_pair_index"])[0] is False


def test_partition_enumerates_every_cross_set_pair_once(oracle):
    ids = [0, 0, 1, 1, 1, 2, 3, 3]
    p = oracle.partition(ids)
    seen = set()
    for i in range(p["num_pairs"]):
        ok, a, b = oracle.partition_pair(ids, i)
        assert ok and ids[a] < ids[b]
        assert oracle.partition_idx(ids, a, b) == i == oracle.partition_idx(ids, b, a)
        seen.add((a, b))
    brute = {(a, b) for a in range(8) for b in range(8) if ids[a] < ids[b]}
    assert seen == brute


@pytest.mark.parametrize("coef", [[5.0], [1.0, 2.0], [0.0, 0.0, 1.0], [0.0, -1.0, 0.0, 1.0],
                                  [0.3, -1.2, 0.7, 0.05, -0.02, 0.004, 1e-3, -2e-4, 3e-5]])
def test_hermite_reproduces_polynomials_up_to_degree_8(oracle, coef):
    P = np.polynomial.Polynomial(coef)
    dP, d2P = P.deriv(1), P.deriv(2)
    dt = 0.7
    q = []
    for j in range(3):
        t = j * dt
        q += [P(t), dP(t), d2P(t)]
    c = oracle.hermite(q, dt)
    for t in np.linspace(-0.5, 2.0, 11):
        assert abs(oracle.hermite_eval(c, dt, t) - P(t)) < 1e-11
        assert abs(oracle.hermite_eval(c, dt, t, deriv=True) - dP(t)) < 1e-10


def test_hermite_coefficients_kat(oracle):
    # closest_approach_utils_test.cu:199-236: f = t^3 - t at 0, 1, 2
    c = oracle.hermite([0.0, -1.0, 0.0, 0.0, 2.0, 6.0, 6.0, 11.0, 12.0], 1.0)
    assert np.allclose(c, [0, -1, 0, 1, 0, 0, 0, 0, 0], atol=1e-15)
    assert oracle.hermite_eval(c, 1.0, 0.5) == -0.375
    assert oracle.hermite_eval(c, 1.0, 0.5, deriv=True) == -0.25


def test_find_root_kats(oracle):
    # closest_approach_utils_test.cu:271-296: the secant from (x0, x1)
    assert abs(oracle.find_root_poly([-6.0, 3.0], 0.0, 1.0) - 2.0) < 1e-7
    assert abs(oracle.find_root_poly([-9.0, 0.0, 1.0], 1.0, 5.0) - 3.0) < 1e-7
    assert abs(oracle.find_root_poly([-9.0, 0.0, 1.0], -1.0, -5.0) + 3.0) < 1e-7
    # f(-2) == f(2): a flat secant returns x1 rather than dividing by zero
    assert oracle.find_root_poly([0.0, 0.0, 1.0], -2.0, 2.0) == 2.0


# closest_approach_utils_test.cu:298-347: A(t) = (t, 2, 0), B(t) = (t, t^2/2 - 3t, 0), given as
# Hermite data (p, v, a) at t = 0, 1, 2
_QA = [[0, 1, 0, 1, 1, 0, 2, 1, 0], [2, 0, 0, 2, 0, 0, 2, 0, 0], [0] * 9]
_QB = [[0, 1, 0, 1, 1, 0, 2, 1, 0], [0, -3, 1, -2.5, -2, 1, -4, -1, 1], [0] * 9]


def test_approach_function_kat(oracle):
    # f(t) = (A - B)(t) . (A - B)'(t) = (-t^2/2 + 3t + 2)(3 - t)
    for t, want in ((0.0, 6.0), (1.0, 9.0), (2.0, 6.0), (3.0, 0.0), (4.0, -6.0)):
        got = oracle.approach_func(_QA, _QB, 1.0, t)
        assert abs(got - want) <= 4 * np.spacing(max(abs(want), 1.0))


def test_root_of_approach_function(oracle):
    tol = 1e-12
    assert abs(oracle.approach_root(_QA, _QB, 1.0, 2.0, 10.0, tol) - 3.0) < tol


def _run_phases(oracle, po, case):
    """Multi-phase ClosestApproach::Propagate: every phase starts a new startup from the
    current state and appends its samples; screening and refinement see all samples."""
    r = np.array(case["pos"], float)
    v = np.array(case["vel"], float)
    trajs = []
    results = []
    for n in case["phases"]:
        rr, vv, traj = oracle.propagate(po.NONE, r, v, case["h"], n, trajectory=True)
        trajs.append(traj)
        r, v = rr, vv
        full = np.concatenate(trajs)
        mind, mink = oracle.screen(full, case["sets"])
        res = {}
        for q in range(len(mind)):
            ok, a, b = oracle.partition_pair(case["sets"], q)
            res[(a, b)] = oracle.refine(full, case["h"], a, b, int(mink[q]), mind[q])
        results.append(res)
    return results


def test_analytic_encounters(oracle, po, golden):
    for case in golden["encounters"]["cases"]:
        res = _run_phases(oracle, po, case)
        tol = case["tol"]
        if "expect_phases" in case:
            for out, (dist, tim) in zip(res, case["expect_phases"]):
                ap = out[(0, 1)]
                assert abs(ap["distance"] - dist) <= tol, case["name"]
                if tim is not None:
                    assert abs(ap["time"] - tim) <= tol, case["name"]
            continue
        for a, b, dist, tim in case["expect"]:
            ap = res[-1][(a, b)]
            if tol == 0:
                assert ap["distance"] == dist, case["name"]
            else:
                assert abs(ap["distance"] - dist) <= tol, case["name"]
            if tim is not None:
                assert abs(ap["time"] - tim) <= tol, case["name"]
            else:
                assert ap["time"] >= 0.0


def test_screen_first_occurrence_wins(oracle):
    # two equal minima: the earlier step is kept (strict '<', device.cu:76)
    traj = np.zeros((5, 2, 9))
    d = [3.0, 1.0, 2.0, 1.0, 4.0]
    for k in range(5):
        traj[k, 1, 0] = d[k]
    mind, mink = oracle.screen(traj, [0, 1])
    assert mind[0] == 1.0 and mink[0] == 1
