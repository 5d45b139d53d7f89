"""Derivatives of the propagation (no counterpart in the reference): the gravity Jacobian, the
state-transition matrix from the variational equations (forward mode) and the discrete adjoint
of the GJ8 steps (reverse mode), checked against each other, against central finite
differences of the FP64 oracle's propagation, against the point-mass closed form, and through a
torch.autograd.Function wrapper.
"""
import numpy as np
import pytest

import opengpop_amd as ogp
from opengpop_amd import scenarios

pytestmark = pytest.mark.gpu


def _rand_points(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return d * rng.uniform(6.6e6, 4.3e7, (n, 1))


@pytest.fixture(scope="module")
def ctx20():
    with ogp.Context(degree=20) as c:
        yield c


@pytest.mark.parametrize("degree", [20, 70])
def test_gravity_jacobian_value_symmetry_laplace(degree):
    r = _rand_points(512, 3)
    with ogp.Context(degree=degree) as c:
        a, G = c.gravity_jacobian_itrs(r)
        assert np.array_equal(a, c.gravity_itrs(r))  # same arithmetic on the value part
    nG = np.linalg.norm(G, axis=(1, 2))
    assert np.all(np.linalg.norm(G - G.transpose(0, 2, 1), axis=(1, 2)) <= 1e-14 * nG)
    assert np.all(np.abs(np.trace(G, axis1=1, axis2=2)) <= 1e-13 * nG)  # div a = -lap U = 0


def test_gravity_jacobian_matches_oracle_differences(oracle70):
    # central differences of the oracle's acceleration (an independent restatement)
    r = _rand_points(64, 4)
    with ogp.Context(degree=70) as c:
        _, G = c.gravity_jacobian_itrs(r)
    hstep = 1.0
    for j in range(3):
        dr = np.zeros(3)
        dr[j] = hstep
        fd = (oracle70.gravity_itrs(r + dr) - oracle70.gravity_itrs(r - dr)) / (2 * hstep)
        err = np.linalg.norm(G[:, :, j] - fd, axis=1) / np.linalg.norm(G, axis=(1, 2))
        assert err.max() < 1e-8


def _near_pole_points(n, seed):
    # |(x, y)| / r from 1e-9 to 1e-1: the columns m >= 2 of the gradient are divided by cos(phi)
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi, n)
    c = 10.0 ** rng.uniform(-9, -1, n)
    u = np.stack([c * np.cos(ang), c * np.sin(ang),
                  np.where(rng.random(n) < 0.5, -1.0, 1.0) * np.sqrt(1 - c * c)], 1)
    return u * rng.uniform(6.6e6, 4.3e7, (n, 1))


@pytest.mark.parametrize("degree", [20, 70])
def test_gravity_jacobian_matches_richardson_differences(degree):
    # the gradient against differences of the same library's acceleration, Richardson-extrapolated
    # (error O(h^4)) with h = 3e-5 r: truncation ~1e-18 and rounding ~5e-12 of |G|, while the
    # terms of each degree up to 70 add more than ~1e-7 of |G| at LEO radii, so a wrong radial
    # weight or second-derivative coefficient of any degree cannot pass
    r = np.vstack([_rand_points(384, 6), _near_pole_points(128, 7)])
    h = 3e-5 * np.linalg.norm(r, axis=1)[:, None]
    with ogp.Context(degree=degree) as c:
        _, G = c.gravity_jacobian_itrs(r)
        fd = np.empty_like(G)
        for j in range(3):
            e = np.zeros(3)
            e[j] = 1.0
            d1 = (c.gravity_itrs(r + h * e) - c.gravity_itrs(r - h * e)) / (2 * h)
            d2 = (c.gravity_itrs(r + 2 * h * e) - c.gravity_itrs(r - 2 * h * e)) / (4 * h)
            fd[:, :, j] = (4 * d1 - d2) / 3
    err = np.linalg.norm(G - fd, axis=(1, 2)) / np.linalg.norm(G, axis=(1, 2))
    assert err.max() < 3e-11, err.max()


def test_gravity_jacobian_point_mass_closed_form():
    r = _rand_points(128, 5)
    with ogp.Context(degree=1) as c:
        _, G = c.gravity_jacobian_itrs(r)
        mu = 3.986004415e14
    rn = np.linalg.norm(r, axis=1)
    u = r / rn[:, None]
    ref = mu / rn[:, None, None] ** 3 * (3 * u[:, :, None] * u[:, None, :] - np.eye(3))
    assert np.abs(G - ref).max() <= 1e-14 * np.abs(ref).max()


def _rel(A, B):
    return np.linalg.norm(A - B, axis=(1, 2)) / np.linalg.norm(B, axis=(1, 2))


def test_adjoint_matches_stm(ctx20):
    # 24 h at 60 s, degree 20 (the differentiation workload, 16 objects)
    r0, v0 = scenarios.leo(16, seed=31)
    r1, v1, Phi = ctx20.stm(r0, v0, 60.0, 1440)
    r2, v2, J = ctx20.adjoint_jacobian(r0, v0, 60.0, 1440)
    rp, vp = ctx20.propagate(r0, v0, 60.0, 1440)
    # the state lane of the variational run and the adjoint's forward pass are the propagation
    assert np.array_equal(r1, rp) and np.array_equal(v1, vp)
    assert np.array_equal(r2, rp) and np.array_equal(v2, vp)
    assert _rel(J, Phi).max() < 1e-9
    # symplectic structure of the exact flow, to the integrator's accuracy
    Jm = np.block([[np.zeros((3, 3)), np.eye(3)], [-np.eye(3), np.zeros((3, 3))]])
    for P in Phi[:4]:
        assert np.abs(P.T @ Jm @ P - Jm).max() < 1e-6 * np.abs(P).max() ** 2


def test_adjoint_uneven_spans_and_checkpoint_interval(ctx20):
    r0, v0 = scenarios.leo(12, seed=32)
    steps = np.array([0, 1, 3, 4, 5, 6, 12, 13, 70, 133, 200, 64 + 5])
    _, _, Phi = ctx20.stm(r0, v0, 60.0, steps)
    _, _, J7 = ctx20.adjoint_jacobian(r0, v0, 60.0, steps, checkpoint_steps=7)
    _, _, J64 = ctx20.adjoint_jacobian(r0, v0, 60.0, steps, checkpoint_steps=64)
    assert np.array_equal(J7, J64)  # recomputation from checkpoints is deterministic
    # automatic interval: the forward pass records the whole span (no checkpoints, no re-run)
    ra, va, Ja = ctx20.adjoint_jacobian(r0, v0, 60.0, steps)
    rp, vp = ctx20.propagate(r0, v0, 60.0, steps)
    assert np.array_equal(Ja, J7)
    assert np.array_equal(ra, rp) and np.array_equal(va, vp)
    assert np.array_equal(J7[0], np.eye(6))
    assert _rel(J7, Phi).max() < 1e-10


def test_derivative_memory_budget_segments():
    # the STM records and consumes the Jacobians segment by segment when the span's do not fit
    # the budget, and the adjoint falls back to checkpoints: the results are unchanged, bit for
    # bit (the same arithmetic, with the state passing through memory between segments)
    r0, v0 = scenarios.leo(12, seed=36)
    steps = np.array([0, 1, 3, 4, 5, 6, 12, 13, 70, 133, 200, 69])
    with ogp.Context(degree=20) as c:
        r1, v1, P1 = c.stm(r0, v0, 60.0, steps)
        c.set_derivative_memory(7 * 6 * 12 * 8)  # 7 steps x 6 entries x 12 objects
        r2, v2, P2 = c.stm(r0, v0, 60.0, steps)
        _, _, Ja = c.adjoint_jacobian(r0, v0, 60.0, steps)
        _, _, J64 = c.adjoint_jacobian(r0, v0, 60.0, steps, checkpoint_steps=64)
        with pytest.raises(ogp.OgpError):
            c.set_derivative_memory(-1)
        c.set_derivative_memory(0)
        _, _, J = c.adjoint_jacobian(r0, v0, 60.0, steps)
        rp, vp = c.propagate(r0, v0, 60.0, steps)
    assert np.array_equal(P1, P2) and np.array_equal(r1, r2) and np.array_equal(v1, v2)
    assert np.array_equal(r1, rp) and np.array_equal(v1, vp)
    assert np.array_equal(Ja, J64) and np.array_equal(J, J64)
    assert np.array_equal(P1[0], np.eye(6))
    assert _rel(P1[1:], J[1:]).max() < 1e-10


def _oracle_fd(po, r0, v0, n_steps, hr, hv):
    o = po.Oracle(degree=20)
    fd = np.zeros((len(r0), 6, 6))
    x0 = np.hstack([r0, v0])
    for col in range(6):
        hstep = hr if col < 3 else hv
        d = np.zeros_like(x0)
        d[:, col] = hstep
        rp, vp = o.propagate(po.GRAVITY, (x0 + d)[:, :3], (x0 + d)[:, 3:], 60.0, n_steps)
        rm, vm = o.propagate(po.GRAVITY, (x0 - d)[:, :3], (x0 - d)[:, 3:], 60.0, n_steps)
        fd[:, :, col] = np.hstack([rp - rm, vp - vm]) / (2 * hstep)
    return fd


def test_adjoint_matches_oracle_finite_differences(ctx20, po):
    # central differences of the oracle's propagation over 24 h. With h = 1e-3 m and 1e-6 m/s
    # the quotients carry ~1e-5 of rounding noise (they differ by that much from quotients with
    # 1000x larger steps); with h = 0.1 m and 1e-4 m/s both rounding and truncation stay below
    # 1e-7, which pins the derivative tightly.
    r0, v0 = scenarios.leo(3, seed=33)
    _, _, J = ctx20.adjoint_jacobian(r0, v0, 60.0, 1440)
    assert _rel(J, _oracle_fd(po, r0, v0, 1440, 1e-3, 1e-6)).max() < 2e-5
    assert _rel(J, _oracle_fd(po, r0, v0, 1440, 0.1, 1e-4)).max() < 1e-6


def test_vjp_matches_jacobian_transpose(ctx20):
    r0, v0 = scenarios.leo(20, seed=34)
    steps = np.full(20, 300)
    steps[:3] = [2, 5, 64]
    _, _, J = ctx20.adjoint_jacobian(r0, v0, 60.0, steps)
    rng = np.random.default_rng(8)
    gr, gv = rng.normal(size=(20, 3)), rng.normal(size=(20, 3))
    a, b = ctx20.adjoint_vjp(r0, v0, 60.0, steps, gr, gv)
    ref = np.einsum("nij,ni->nj", J, np.hstack([gr, gv]))
    assert np.allclose(np.hstack([a, b]), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_torch_autograd_wrapper(ctx20):
    torch = pytest.importorskip("torch")

    class Propagate(torch.autograd.Function):
        """Final state of a propagation; the backward pass is the library's adjoint."""

        @staticmethod
        def forward(fctx, r0, v0, step, steps):
            r, v = ctx20.propagate(r0.detach().numpy(), v0.detach().numpy(), step, steps)
            fctx.save_for_backward(r0, v0)
            fctx.args = (step, steps)
            return torch.from_numpy(r), torch.from_numpy(v)

        @staticmethod
        def backward(fctx, gr, gv):
            r0, v0 = fctx.saved_tensors
            step, steps = fctx.args
            a, b = ctx20.adjoint_vjp(r0.detach().numpy(), v0.detach().numpy(), step, steps,
                                     gr.numpy(), gv.numpy())
            return torch.from_numpy(a), torch.from_numpy(b), None, None

    rn, vn = scenarios.leo(5, seed=35)
    r0 = torch.tensor(rn, requires_grad=True)
    v0 = torch.tensor(vn, requires_grad=True)
    r, v = Propagate.apply(r0, v0, 60.0, 500)
    target = torch.tensor(rn[::-1].copy())
    loss = ((r - target) ** 2).sum() + 1e3 * (v ** 2).sum()
    loss.backward()
    _, _, Phi = ctx20.stm(rn, vn, 60.0, 500)
    g = np.hstack([2 * (r.detach().numpy() - target.numpy()), 2e3 * v.detach().numpy()])
    ref = np.einsum("nij,ni->nj", Phi, g)
    got = np.hstack([r0.grad.numpy(), v0.grad.numpy()])
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_derivatives_need_fp64_gravity():
    r0, v0 = scenarios.leo(2, seed=1)
    with ogp.Context(degree=20, precision="split") as c:
        with pytest.raises(ogp.OgpError):
            c.adjoint_jacobian(r0, v0, 60.0, 10)
    with ogp.Context(force=ogp.FORCE_NONE) as c:
        with pytest.raises(ogp.OgpError):
            c.stm(r0, v0, 60.0, 10)
