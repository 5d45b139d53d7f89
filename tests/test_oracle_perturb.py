"""Pins the oracle's Monte-Carlo population (PerturbationGenerator, perturbation_generator.cu:21-79).

The reference draws its normals with cuRAND; the product uses Philox4x32-10 with Box-Muller.
Pinned here by the published Philox known-answer vectors, the Box-Muller map evaluated from the
raw Philox words with numpy, distribution tests, and the generator's structural properties
(positions copied, strength 0 gives identical satellites, entry i depends only on (seed, i)).
"""
import json
import os

import numpy as np
import pytest
from scipy import stats

HERE = os.path.dirname(os.path.abspath(__file__))
R0 = np.array([-3884872.0, -5749111.0, 14723.0])   # propagator_test.cu:40
V0 = np.array([-795.304, 573.772, 7518.179])       # propagator_test.cu:41-45


def _kat():
    with open(os.path.join(HERE, "golden", "philox4x32_10_kat.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _kat())
def test_philox_known_answers(oracle, case):
    h = lambda words: [int(w, 16) for w in words]
    out = oracle.philox4x32_10(h(case["ctr"]), h(case["key"]))
    assert list(out) == h(case["out"])


def test_box_muller_from_philox_words(oracle):
    # z = sqrt(-2 ln u1) (cos, sin)(2 pi u2) with u = (53 bits + 1/2) 2^-53 of a word pair;
    # normals 0, 1 from counter block (i, 0), normal 2 the cosine of block (i, 1)
    seed = 0x0000000500000007
    key = [seed & 0xFFFFFFFF, seed >> 32]
    n = 40
    _, v = oracle.perturbed(R0, np.zeros(3), n, 1.0, seed)

    def unit(hi, lo):
        return ((int(hi) << 21 | int(lo) >> 11) + 0.5) / 2.0 ** 53

    for i in range(n):
        z = []
        for blk in (0, 1):
            w = oracle.philox4x32_10([i & 0xFFFFFFFF, i >> 32, blk, 0], key)
            rad = np.sqrt(-2.0 * np.log(unit(w[0], w[1])))
            ang = 2.0 * np.pi * unit(w[2], w[3])
            z += [rad * np.cos(ang), rad * np.sin(ang)]
        assert np.allclose(v[i], z[:3], rtol=1e-14, atol=1e-15)


def test_distribution_is_standard_normal(oracle):
    n = 200_000
    for seed in (1234, 99):
        _, v = oracle.perturbed(R0, V0, n, 2.5, seed)
        z = (v - V0) / 2.5
        for c in range(3):
            assert abs(z[:, c].mean()) < 5.0 / np.sqrt(n)
            assert abs(z[:, c].var() - 1.0) < 5.0 * np.sqrt(2.0 / n)
            assert stats.kstest(z[:, c], "norm").pvalue > 1e-3
        cc = np.corrcoef(z.T)
        assert np.abs(cc - np.eye(3)).max() < 5.0 / np.sqrt(n)


def test_structure(oracle):
    r, v = oracle.perturbed(R0, V0, 64, 3.0, 1234)
    assert np.array_equal(r, np.tile(R0, (64, 1)))  # positions are the epoch position
    r0, v0 = oracle.perturbed(R0, V0, 16, 0.0, 1234)
    assert np.array_equal(v0, np.tile(V0, (16, 1)))  # strength 0: identical satellites
    _, v_short = oracle.perturbed(R0, V0, 10, 3.0, 1234)
    assert np.array_equal(v_short, v[:10])           # entry i depends only on (seed, i)
    _, v_other = oracle.perturbed(R0, V0, 10, 3.0, 1235)
    assert not np.any(v_other == v_short)
