"""The oracle's blocked screen (orc_screen_blocked, used for the full-size conjunction check) against
its per-pair screen (orc_screen, the restatement of UpdateClosestApproachKernel,
closest_approach/src/device.cu:56-91): identical minima and steps, bit for bit, on tie-heavy,
NaN-carrying and ragged inputs."""
import numpy as np


def _ties(rng, nsteps=77, n=150):
    # static objects repeat their distances exactly at every sample (the first must win); the
    # moving ones drift by ~1 ulp per step, so roots tie or nearly tie with the running minimum
    r = rng.normal(size=(n, 3)) * 7e6
    traj = np.repeat(r[None], nsteps, axis=0)
    moving = rng.random(n) < 0.6
    drift = np.arange(nsteps)[:, None, None] * np.array([1e-9, -2e-9, 1e-9])
    traj[:, moving] += drift * np.abs(r[None, moving]) * 1e-7
    traj[:, 5:10] = traj[:, 0:5]  # distance 0 pairs across sets
    return np.concatenate([traj, np.zeros_like(traj)], axis=2)  # 6 components per sample


def test_blocked_equals_per_pair_on_ties(oracle):
    rng = np.random.default_rng(5)
    traj = _ties(rng)
    sets = np.repeat(np.arange(3, dtype=np.uint32), [5, 70, 75])
    ref_d, ref_k = oracle.screen(traj, sets)
    d, k = oracle.screen_blocked(traj, sets)
    assert np.array_equal(d, ref_d)
    assert np.array_equal(k, ref_k)
    assert np.any(ref_k == 0) and np.any(ref_d == 0.0)


def test_blocked_equals_per_pair_random_with_nans(oracle):
    rng = np.random.default_rng(9)
    nsteps, n = 131, 257  # partial tiles in both dimensions
    traj = rng.normal(size=(nsteps, n, 3)) * 1e4
    traj[rng.random((nsteps, n)) < 0.02] = np.nan
    traj[:, 3] = np.nan  # an object that never gives a sample
    sets = np.repeat(np.arange(4, dtype=np.uint32), [33, 64, 100, 60])
    ref_d, ref_k = oracle.screen(traj, sets)
    d, k = oracle.screen_blocked(traj, sets, threads=3)
    assert np.array_equal(d, ref_d)
    assert np.array_equal(k, ref_k)
    assert np.any(ref_k == -1) and np.any(ref_d == np.finfo(float).max)


def test_blocked_threshold_steps_only_below(oracle):
    rng = np.random.default_rng(2)
    traj = rng.normal(size=(60, 90, 3)) * 1e4
    sets = np.repeat(np.arange(2, dtype=np.uint32), [30, 60])
    ref_d, ref_k = oracle.screen(traj, sets)
    thr = float(np.median(ref_d))
    d, k = oracle.screen_blocked(traj, sets, threshold=thr)
    assert np.array_equal(d, ref_d)
    below = ref_d < thr
    assert np.array_equal(k[below], ref_k[below])
    assert np.all(k[~below] == -1)
