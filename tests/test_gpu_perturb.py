"""GPU parity of the Monte-Carlo population (PerturbationGenerator, perturbation_generator.cu:21-79)
against the oracle, and the reference's own uses of the generator: identical satellites
(strength 0) must propagate and accelerate bit-identically (propagator_test.cu:38-80,
force_model_test.cu:38-90).
"""
import numpy as np
import pytest
import torch

import opengpop_amd as ogp

pytestmark = pytest.mark.gpu

R0 = np.array([-3884872.0, -5749111.0, 14723.0])   # propagator_test.cu:40
V0 = np.array([-795.304, 573.772, 7518.179])       # propagator_test.cu:41-45
EPOCH = (2015, 5, 4, 22, 22, 2.0)                   # propagator_test.cu:52-57


@pytest.fixture(scope="module")
def ctx():
    with ogp.Context(degree=20, epoch=EPOCH) as c:
        yield c


def test_matches_oracle(ctx, po):
    o = po.Oracle(degree=None)
    for n, strength, seed in ((1, 1.0, 1234), (1000, 3.5, 1234), (777, 0.01, 2 ** 40 + 3)):
        r, v = ctx.perturbed_satellites(R0, V0, n, strength, seed)
        ro, vo = o.perturbed(R0, V0, n, strength, seed)
        assert np.array_equal(r, ro)
        # the normals go through log, sqrt, cos and sin of two math libraries (a few ulps of z),
        # and v = v0 + z * strength rounds once more: allow 1e-14 * strength plus 2 ulps of v
        tol = 1e-14 * strength + 2.0 * np.spacing(np.abs(vo))
        assert np.all(np.abs(v - vo) <= tol), (n, seed, np.max(np.abs(v - vo) / tol))


def test_appended_object_and_device_layout(ctx):
    n = 300
    obj_r = np.array([7.0e6, -1.0, 2.0])
    obj_v = np.array([1.0, 7.5e3, -3.0])
    r, v = ctx.perturbed_satellites(R0, V0, n, 2.0, 77, obj_r, obj_v)
    assert r.shape == (n + 1, 3)
    assert np.array_equal(r[n], obj_r) and np.array_equal(v[n], obj_v)
    r_only, v_only = ctx.perturbed_satellites(R0, V0, n, 2.0, 77, obj_r, None)
    assert r_only.shape == (n, 3)  # appended only when both arrays are given
    assert np.array_equal(v_only, v[:n])
    # device SoA output with a larger stride: same numbers, layout x[stride], y[stride], z[stride]
    stride = n + 5
    dev = torch.device("cuda:0")
    rd = torch.full((3, stride), -7.0, dtype=torch.float64, device=dev)
    vd = torch.full((3, stride), -7.0, dtype=torch.float64, device=dev)
    ctx.perturbed_satellites_device(R0, V0, n, 2.0, rd, vd, stride=stride, seed=77,
                                    obj_r=obj_r, obj_v=obj_v)
    rh, vh = rd.cpu().numpy(), vd.cpu().numpy()
    assert np.array_equal(rh[:, :n + 1].T, r) and np.array_equal(vh[:, :n + 1].T, v)
    assert np.all(rh[:, n + 1:] == -7.0) and np.all(vh[:, n + 1:] == -7.0)


def test_identical_satellites_propagate_identically(ctx):
    # strength 0 gives 8 copies of the epoch state; 10 steps of 30 s (propagator_test.cu:46-80)
    r0, v0 = ctx.perturbed_satellites(R0, V0, 8, 0.0)
    assert np.array_equal(r0, np.tile(R0, (8, 1))) and np.array_equal(v0, np.tile(V0, (8, 1)))
    p = ogp.GpuPropagator(ctx, r0, v0, 30.0, 10)
    for _ in range(10):
        r, v, a = p.step()
        for arr in (r, v, a):
            assert np.all(arr == arr[0])
    # force_model_test.cu:63-90: column 0 of the [obj][9][3] acceleration matrix at propagation
    # node 0 is identical for every satellite
    out = ctx.accel(r0, 30.0, 0, out_stride=27)
    a = np.array([out[27 * s:27 * s + 3] for s in range(8)])
    assert np.all(np.isfinite(a)) and np.all(a == a[0])


def test_population_on_device_feeds_the_propagator(ctx):
    # a 4,096-satellite cloud generated in HBM and propagated without leaving the device
    n = 4096
    dev = torch.device("cuda:0")
    rd = torch.empty((3, n), dtype=torch.float64, device=dev)
    vd = torch.empty((3, n), dtype=torch.float64, device=dev)
    ctx.perturbed_satellites_device(R0, V0, n, 1.0, rd, vd)
    ro = torch.empty_like(rd)
    vo = torch.empty_like(vd)
    ctx.propagate_device(rd, vd, n, 30.0, 40, ro, vo)
    r_host, v_host = ctx.perturbed_satellites(R0, V0, n, 1.0)
    rb, vb = ctx.propagate(r_host, v_host, 30.0, 40)
    assert np.array_equal(ro.cpu().numpy().T, rb) and np.array_equal(vo.cpu().numpy().T, vb)
    # the velocity spread of 1 m/s grows into an along-track spread of kilometres
    spread = np.linalg.norm(rb - rb.mean(0), axis=1)
    assert 1e2 < np.median(spread) < 1e5


def test_statistics_at_scale(ctx):
    n = 1_000_000
    _, v = ctx.perturbed_satellites(R0, V0, n, 4.0, 2026)
    z = (v - V0) / 4.0
    for c in range(3):
        assert abs(z[:, c].mean()) < 5.0 / np.sqrt(n)
        assert abs(z[:, c].var() - 1.0) < 5.0 * np.sqrt(2.0 / n)


def test_errors(ctx):
    with pytest.raises(ogp.OgpError):
        ctx.perturbed_satellites(R0, V0, -1, 1.0)
    with pytest.raises(ogp.OgpError):
        ctx.perturbed_satellites(R0, V0, 4, float("nan"))
    dev = torch.device("cuda:0")
    rd = torch.empty((3, 4), dtype=torch.float64, device=dev)
    with pytest.raises(ogp.OgpError):
        ctx.perturbed_satellites_device(R0, V0, 8, 1.0, rd, rd, stride=4)
