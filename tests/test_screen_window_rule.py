"""The screen kernel's deferred decision (opengpop_amd/csrc/conjunction.hip, screen_kernel) against
the reference's per-sample rule (closest_approach/src/device.cu:56-91: a sample is taken when
its distance, sqrt(d2), is below the running minimum; the first occurrence wins), both applied to
the same squared-distance sequences. The sequences are built to be tie-heavy: values that repeat
exactly or move by a few ulps, so that roots tie or nearly tie with the running minimum, plus
NaN and inf samples. A CPU model of the kernel's window arithmetic, not of its code.
"""
import numpy as np
import pytest

SHRINK = 1.0 - 2.0**-50


def per_sample(d2):
    best = np.full(d2.shape[1], np.finfo(np.float64).max)
    bk = np.full(d2.shape[1], -1)
    with np.errstate(invalid="ignore"):
        for t in range(d2.shape[0]):
            d = np.sqrt(d2[t])
            take = d < best
            best = np.where(take, d, best)
            bk = np.where(take, t, bk)
    return best, bk


def windowed(d2, span):
    nsteps, npairs = d2.shape
    mn = np.full(npairs, np.inf)
    sk = np.full(npairs, -1)
    with np.errstate(invalid="ignore"):
        for w0 in range(0, nsteps, span):
            wm = np.fmin.reduce(d2[w0:w0 + span], axis=0, initial=np.inf)  # NaN never taken
            drop = wm < mn * SHRINK
            unsure = (wm < mn) & ~drop
            exact = unsure & (np.sqrt(wm) < np.sqrt(mn))
            sk = np.where(drop | exact, w0 // span, sk)
            mn = np.fmin(mn, wm)
        d = np.where(sk < 0, np.finfo(np.float64).max, np.sqrt(mn))
        k = np.full(npairs, -1)
        for p in np.nonzero(sk >= 0)[0]:
            w = d2[sk[p] * span:(sk[p] + 1) * span, p]
            k[p] = sk[p] * span + int(np.nonzero(np.sqrt(w) == d[p])[0][0])
    return d, k


def tie_heavy(nsteps, npairs, seed):
    rng = np.random.default_rng(seed)
    base = rng.uniform(1e6, 1e14, npairs)
    # slow drift plus ulp-level jitter, exact repeats and occasional real drops
    ulps = rng.integers(-4, 5, (nsteps, npairs)).astype(np.float64)
    ulps[rng.random((nsteps, npairs)) < 0.3] = 0.0
    d2 = base * (1.0 + ulps * 2.0**-52)
    d2 *= np.where(rng.random((nsteps, npairs)) < 0.01, rng.uniform(0.5, 1.0, (nsteps, npairs)), 1.0)
    d2 = np.minimum.accumulate(d2, axis=0) * (1.0 + rng.integers(0, 3, (nsteps, npairs)) * 2.0**-52)
    d2[rng.random((nsteps, npairs)) < 0.002] = np.nan
    d2[rng.random((nsteps, npairs)) < 0.002] = np.inf
    d2[:, :5] = np.nan          # never taken: DBL_MAX and step -1
    d2[:, 5:10] = 42.0          # exact ties at every sample: the first wins
    d2[:, 10:15] = 0.0
    return d2


@pytest.mark.parametrize("span", [8, 16, 32])
@pytest.mark.parametrize("nsteps", [1, 7, 16, 203])
def test_window_rule_equals_per_sample_rule(span, nsteps):
    d2 = tie_heavy(nsteps, 4000, seed=nsteps * 100 + span)
    ref_d, ref_k = per_sample(d2)
    d, k = windowed(d2, span)
    assert np.array_equal(d, ref_d) and np.array_equal(k, ref_k)
    assert np.all(ref_k[:5] == -1) and np.all(ref_k[5:15] == 0)


def test_sequences_are_tie_heavy():
    # most pairs meet a sample whose root equals the running minimum's without being taken
    d2 = tie_heavy(203, 4000, seed=1)
    r = np.sqrt(d2)
    run = np.fmin.accumulate(np.where(np.isnan(r), np.inf, r), axis=0)
    ties = (r[1:] == run[:-1]).sum(0)
    assert (ties > 0).mean() > 0.5
    near = (d2[1:] < np.fmin.accumulate(np.nan_to_num(d2, nan=np.inf), axis=0)[:-1]) & \
        (np.sqrt(d2[1:]) == run[:-1])
    assert near.sum() > 1000  # smaller d2 with an equal root: the case the exact branch decides
