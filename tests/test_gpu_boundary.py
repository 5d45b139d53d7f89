"""The device-pointer boundary and the stepwise propagators (GpuPropagator, gpu_propagator.cu:28-154;
IForceModel, iforce_model.cuh:19-33; GpuStartup, gpu_startup.cu:55-68; Predictor,
prediction.cu:49-74), and the per-object failure flags.

Parity: stepwise propagators equal the bulk propagation bit for bit whatever other calls of the
context are interleaved; device forms equal their host forms bit for bit; the generation
acceleration equals the propagation acceleration at the same time node and the oracle's frame
and gravity at a fractional stage; the predictor equals the reference's known answer.
"""

import numpy as np
import pytest

import opengpop_amd as ogp
from opengpop_amd import _capi as C
from opengpop_amd import scenarios

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
DEV = torch.device("cuda")


@pytest.fixture(scope="module")
def ctx20():
    with ogp.Context(degree=20) as c:
        yield c


def _aos(a):
    return torch.tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)


def test_propagators_keep_their_own_state(ctx20):
    # two propagators of different sizes and the context's own, with every other kind of call
    # of the context interleaved between their steps: each equals the bulk propagation
    d = scenarios.config0()
    a = ogp.GpuPropagator(ctx20, d["r"][:16], d["v"][:16], 60.0, 30)
    b = ogp.GpuPropagator(ctx20, d["r"][16:56], d["v"][16:56], 60.0, 30)
    L = ogp.lib()
    r0, v0 = np.ascontiguousarray(d["r"][56:64]), np.ascontiguousarray(d["v"][56:64])
    assert L.ogp_prop_init(ctx20.h, r0.ctypes.data, v0.ctypes.data, 8, 60.0, 30) == 0
    own = np.zeros((8, 3)), np.zeros((8, 3)), np.zeros((8, 3))
    for k in range(30):
        ra, va, _ = a.step()
        rb, vb, _ = b.step()
        assert L.ogp_prop_step(ctx20.h, *[o.ctypes.data for o in own]) == 0
        if k % 7 == 0:  # the calls that reuse the context's shared buffers
            ctx20.propagate(d["r"], d["v"], 60.0, 9)
            ctx20.startup(d["r"][:40], d["v"][:40], 60.0)
            ca = ogp.ClosestApproach(ctx20, 60.0)
            ca.initialize(d["r"], d["v"], np.repeat([0, 1], 32).astype(np.uint32))
            ca.propagate(6)
    for (r, v), lo, hi in (((ra, va), 0, 16), ((rb, vb), 16, 56), (own[:2], 56, 64)):
        rr, vv = ctx20.propagate(d["r"][lo:hi], d["v"][lo:hi], 60.0, 30)
        assert np.array_equal(r, rr) and np.array_equal(v, vv)
    assert L.ogp_prop_step(ctx20.h, *[o.ctypes.data for o in own]) == C.ERR_STATE  # max_steps
    a.close()
    b.close()


def test_device_propagator_advances_caller_buffers(ctx20):
    d = scenarios.config0()
    n = 24
    R, V = _aos(d["r"][:n]), _aos(d["v"][:n])
    acc = torch.zeros((n, 3), dtype=torch.float64, device=DEV)
    torch.cuda.synchronize()
    pd = ogp.GpuPropagator(ctx20, R, V, 60.0, 12, n=n, device=True)
    ph = ogp.GpuPropagator(ctx20, d["r"][:n], d["v"][:n], 60.0, 12)
    for _ in range(12):
        rd, vd, ad = pd.step(d_acc=acc)
        rh, vh, ah = ph.step()
        torch.cuda.synchronize()
        assert np.array_equal(R.cpu().numpy(), rh) and np.array_equal(V.cpu().numpy(), vh)
        assert np.array_equal(acc.cpu().numpy(), ah) and np.array_equal(ad, ah)
        assert np.array_equal(rd, rh) and np.array_equal(vd, vh)
    cr, cv = pd.copy_state_to_host()
    assert np.array_equal(cr, rh) and np.array_equal(cv, vh)
    rb, vb = ctx20.propagate(d["r"][:n], d["v"][:n], 60.0, 12)
    assert np.array_equal(rh, rb) and np.array_equal(vh, vb)
    pd.close()
    ph.close()


def test_propagator_outlived_by_nothing():
    # a context destroyed before its propagator: later calls report a state error
    d = scenarios.config0()
    c = ogp.Context(degree=20)
    p = ogp.GpuPropagator(c, d["r"][:4], d["v"][:4], 60.0, 5)
    p.step()
    c.close()
    r = np.zeros((4, 3))
    assert ogp.lib().ogp_propagator_step(p.h, r.ctypes.data, None, None) == C.ERR_STATE
    p.close()


def test_accel_generation_equals_propagation_node_and_oracle(ctx20, po):
    # stage k of startup step s sits at (s - 1 + 4 + c_k) h - 4 h from the epoch for s > 0 and
    # at the reference's mirrored (s + 4 + c_k) h - 4 h for s < 0: with c_k = 0 or 1 that is a
    # propagation node, computed here with the same time-node arithmetic
    n = 40
    r, v = scenarios.leo(n, seed=8)
    Rd, Vd = _aos(r), _aos(v)
    K = torch.zeros((n, 10, 3), dtype=torch.float64, device=DEV)
    A = torch.zeros((n, 3), dtype=torch.float64, device=DEV)
    for s, k, node in ((1, 0, 4), (2, 0, 5), (1, 9, 5), (-1, 0, 3), (-4, 9, 1), (4, 7, 8)):
        ctx20.accel_generation(Rd, Vd, n, 60.0, s, k, K)
        ctx20.accel_device(Rd, Vd, n, 60.0, node, A)
        torch.cuda.synchronize()
        assert np.array_equal(K[:, k].cpu().numpy(), A.cpu().numpy()), (s, k)
    # a fractional stage against the oracle's frame and gravity: c_3 = 1/3 of step 2
    ctx20.accel_generation(Rd, Vd, n, 60.0, 2, 3, K)
    torch.cuda.synchronize()
    o = po.Oracle(degree=20)
    M, _ = o.time_nodes(scenarios.EPOCH, -4 * 60.0, [(5 + 1.0 / 3.0) * 60.0])
    Mm = M[0].reshape(3, 3).T  # column-major
    ref = (Mm @ o.gravity_itrs(r @ Mm).T).T
    got = K[:, 3].cpu().numpy()
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13
    with pytest.raises(ogp.OgpError):
        ctx20.accel_generation(Rd, Vd, n, 60.0, 0, 0, K)
    with pytest.raises(ogp.OgpError):
        ctx20.accel_generation(Rd, Vd, n, 60.0, 1, 10, K)


def test_startup_device_equals_host(ctx20):
    d = scenarios.config0()
    n = 20
    host = ctx20.startup(d["r"][:n], d["v"][:n], 60.0)
    outs = {k: torch.zeros((n, 9, 3), dtype=torch.float64, device=DEV)
            for k in ("pos", "vel", "s", "S", "acc")}
    ctx20.startup_device(_aos(d["r"][:n]), _aos(d["v"][:n]), n, 60.0, d_pos=outs["pos"],
                         d_vel=outs["vel"], d_s=outs["s"], d_S=outs["S"], d_acc=outs["acc"])
    torch.cuda.synchronize()
    for k, t in outs.items():
        assert np.array_equal(t.cpu().numpy(), host[k]), k
    bad = d["r"][:n].copy()
    bad[3, 1] = np.nan
    with pytest.raises(ogp.OgpError) as e:
        ctx20.startup_device(_aos(bad), _aos(d["v"][:n]), n, 60.0, d_pos=outs["pos"])
    assert e.value.code == C.ERR_NONFINITE


def test_predict_device_kat_and_history(ctx20, golden):
    # Predictor::Predict: the predicted state is the reference's known answer
    # (integrator_prediction_test.cu), the history moves by one column and column 8 is the
    # acceleration at node step + 4 of the predicted state
    k = golden["predictor_kat"]
    acc = np.asarray(k["acc"], dtype=np.float64).reshape(-1, 9, 3)
    n = len(acc)
    s = np.asarray(k["s"], dtype=np.float64).reshape(n, 3)
    S = np.asarray(k["S"], dtype=np.float64).reshape(n, 3)
    Rd = torch.zeros((n, 3), dtype=torch.float64, device=DEV)
    Vd = torch.zeros_like(Rd)
    Ad, sd, Sd = _aos(acc), _aos(s), _aos(S)
    ctx20.predict_device(Rd, Vd, Ad, sd, Sd, n, k["h"], 5)
    torch.cuda.synchronize()
    r_kat, v_kat, S_kat = ctx20.kat_predict(acc.reshape(n, 27), s, S, k["h"])
    assert np.array_equal(Rd.cpu().numpy(), r_kat) and np.array_equal(Vd.cpu().numpy(), v_kat)
    assert np.array_equal(Sd.cpu().numpy(), S_kat)
    assert np.allclose(Rd.cpu().numpy(), np.full((n, 3), k["r"]), rtol=1e-7, atol=0)
    got = Ad.cpu().numpy()
    assert np.array_equal(got[:, :8], acc[:, 1:])
    # (the known answer's state sits centimetres from the geocentre, where the field is huge;
    # the comparison is of bits, whatever they are)
    a9 = torch.zeros((n, 3), dtype=torch.float64, device=DEV)
    ctx20.accel_device(Rd, Vd, n, k["h"], 9, a9)
    torch.cuda.synchronize()
    assert np.array_equal(got[:, 8], a9.cpu().numpy(), equal_nan=True)


def _with_bad_objects(n, seed=12):
    r, v = scenarios.leo(n, seed=seed)
    bad_in = np.array([5, n // 2, n - 1])
    r[bad_in[0], 0] = np.nan
    v[bad_in[1], 2] = np.inf
    r[bad_in[2]] = [np.nan, 1.0, -np.inf]
    at_origin = 777  # finite, but the gravity of r = 0 is not: the result turns non-finite
    r[at_origin] = 0.0
    v[at_origin] = 0.0
    return r, v, bad_in, at_origin


def test_failure_flags_host_and_device():
    n = 100_000
    r, v, bad_in, at_origin = _with_bad_objects(n)
    good = np.setdiff1d(np.arange(n), np.r_[bad_in, at_origin])
    with ogp.Context(degree=20) as c:
        ro, vo, st = c.propagate_checked(r, v, 60.0, 12)
        rc, vc = c.propagate(r[good], v[good], 60.0, 12)  # the clean run
        R0 = torch.tensor(r.T.copy(), device=DEV)
        V0 = torch.tensor(v.T.copy(), device=DEV)
        Ro, Vo = torch.empty_like(R0), torch.empty_like(V0)
        torch.cuda.synchronize()
        std = c.propagate_device(R0, V0, n, 60.0, 12, Ro, Vo, status=True)
        with pytest.raises(ogp.OgpError) as e:
            c.propagate_device(R0, V0, n, 60.0, 12, Ro, Vo)
        assert e.value.code == C.ERR_NONFINITE and str(bad_in[0]) in str(e.value)
        with pytest.raises(ogp.OgpError):
            c.propagate(r, v, 60.0, 12)
    for flags in (st, std):
        assert set(np.nonzero(flags == ogp.OBJECT_NONFINITE_INPUT)[0]) == set(bad_in)
        assert list(np.nonzero(flags == ogp.OBJECT_NONFINITE_RESULT)[0]) == [at_origin]
        assert (flags == ogp.OBJECT_OK).sum() == n - 4
    assert np.array_equal(ro[good], rc) and np.array_equal(vo[good], vc)
    dr, dv = Ro.cpu().numpy().T, Vo.cpu().numpy().T
    assert np.array_equal(dr[good], rc) and np.array_equal(dv[good], vc)
    # skipped objects keep their input
    assert np.array_equal(ro[bad_in], r[bad_in], equal_nan=True)
    assert np.array_equal(dr[bad_in], r[bad_in], equal_nan=True)
