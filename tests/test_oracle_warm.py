"""CPU checks of the oracles of the learned warm start: the NumPy network restatement
(oracle/mlp_oracle.py) against finite differences, a library BF16 conversion and Adam's closed
form, and the FP64 solver's multiplier warm start against the fixed point it must share with the
plain solver."""
import numpy as np
import pytest

import mlp_oracle as M
import oracle as O
from sbs_amd import scenes


def random_params(rng, scale=1.0):
    p = np.zeros(M.NPARAMS)
    for (w, b, fan) in ((M.W1, M.B1, 12), (M.W2, M.B2, 128), (M.W3, M.B3, 128), (M.W4, M.B4, 128)):
        p[w:b] = rng.normal(0, scale * np.sqrt(2.0 / fan), b - w)
    p[M.B1:M.W2] = rng.normal(0, 0.1, 128)
    p[M.B2:M.W3] = rng.normal(0, 0.1, 128)
    p[M.B3:M.W4] = rng.normal(0, 0.1, 128)
    p[M.B4:] = rng.normal(0, 0.1, 32)
    p[M.W1:M.B1].reshape(16, 128)[12:] = 0.0  # unused inputs
    return p


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 1000) * 1e-30, [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8,
                                                                                   -2.5, 0.0, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(x).astype(np.float32), ref)
    assert M.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0  # tie to even
    assert M.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


@pytest.mark.parametrize("nout", [2, 3])
def test_network_gradient_matches_finite_differences(nout):
    rng = np.random.default_rng(7)
    p = random_params(rng)
    x = rng.normal(0, 1, (9, 12))
    xn = M.normalize(x, np.zeros(12), np.ones(12), rounding=False)
    t = rng.normal(0, 1, (9, nout))
    loss, g = M.loss_grad(p, xn, t, nout, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(M.W1, M.W1 + 12 * 128), 60, replace=False),
                          rng.choice(np.arange(M.W2, M.W4), 120, replace=False),
                          np.arange(M.W4, M.W4 + 128 * 32, 32)[:40] + rng.integers(0, nout),
                          np.arange(M.B1, M.W2)[:20], np.arange(M.B4, M.B4 + nout)])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_grad(pp, xn, t, nout, False)[0] - M.loss_grad(pm, xn, t, nout, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # outputs past nout carry no gradient
    assert np.all(g[M.W4:M.B4].reshape(128, 32)[:, nout:] == 0) and np.all(g[M.B4 + nout:] == 0)


def test_rounded_network_stays_close_to_exact():
    rng = np.random.default_rng(3)
    p = random_params(rng)
    x = rng.normal(0, 1, (64, 12))
    y1, _ = M.forward(p, M.normalize(x, np.zeros(12), np.ones(12), True), True)
    y0, _ = M.forward(p, M.normalize(x, np.zeros(12), np.ones(12), False), False)
    assert np.abs(y1 - y0).max() <= 0.03 * np.abs(y0).max()


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([3.0, -0.5, 1e-3, 0.0])
    p, m, v = M.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    # m^ = g and v^ = g^2 after bias correction: the step is lr * g / (|g| + eps)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0) and p[3] == 0.0


def _beam():
    return scenes.cantilever(6, 2, 2, 0.1, "neohookean", 5e5, 0.45, 1000.0, num_spheres=0, iterations=4, substeps=2)


def test_oracle_warm_start_is_plain_on_the_first_substep():
    # multipliers start at zero, so the warm start applies nothing and keeps nothing
    sc = _beam()
    a, b = O.Simulation(), O.Simulation()
    for s in (a, b):
        sc.apply(s)
        s.set_order(True, 2)
    b.set_warm_start("previous")
    a.step(sc.dt, sc.iterations, 1)
    b.step(sc.dt, sc.iterations, 1)
    assert np.array_equal(a.read_body(0)[0], b.read_body(0)[0])


def nh_energy(x, tets, x0, young, poisson):
    """Stable neo-Hookean energy sum V0 (mu/2 |F|^2 + lambda/2 (det F - 1 - mu/lambda)^2), whose
    constraint pair (C_dev = |F|, C_hyd = det F - 1 - mu/lambda, compliances 1/(mu V0), 1/(lambda V0))
    the solver projects."""
    mu = young / (2 * (1 + poisson))
    la = young * poisson / ((1 + poisson) * (1 - 2 * poisson))
    Dm = np.stack([x0[tets[:, k]] - x0[tets[:, 3]] for k in range(3)], axis=2)
    Ds = np.stack([x[tets[:, k]] - x[tets[:, 3]] for k in range(3)], axis=2)
    F = Ds @ np.linalg.inv(Dm)
    V0 = np.abs(np.linalg.det(Dm)) / 6
    return np.sum(V0 * (0.5 * mu * np.sum(F * F, axis=(1, 2)) + 0.5 * la * (np.linalg.det(F) - 1 - mu / la) ** 2))


def test_oracle_explicit_warm_start_is_an_explicit_elastic_step():
    # with lambda = -C(x~)/alpha~ the warm-start displacement is -dt^2 M^-1 grad E(x~): the tet
    # pair's multipliers times their gradients sum to the elastic force (checked by finite
    # differences of the energy, zero iterations so the step ends right after the warm start)
    sc = scenes.cantilever(4, 2, 2, 0.1, "neohookean", 5e5, 0.45, 1000.0, num_spheres=0, iterations=1, substeps=1)
    b = sc.bodies[0]
    tets = np.asarray(sc.groups[0].conn).reshape(-1, 4)
    s = O.Simulation()
    sc.apply(s)
    s.set_order(True, 2)
    s.set_warm_start("explicit")
    dt = 1.0 / 240.0
    # a deformed start (no elastic equilibrium): stretch the beam along x
    x = b.x0.copy()
    x[:, 0] *= 1.05
    s.set_body_state(0, x, np.zeros_like(x))
    s.step(dt, 0, 1)
    xs = s.read_body(0)[0]
    w = np.where(b.mass > 0, 1.0 / np.where(b.mass > 0, b.mass, 1.0), 0.0)
    xt = x + (w > 0)[:, None] * np.array([0.0, -9.81, 0.0]) * dt * dt  # predicted positions (gravity as acceleration)
    E0 = nh_energy(xt, tets, b.x0, sc.groups[0].young, sc.groups[0].poisson)
    grad = np.zeros_like(xt)
    h = 1e-7
    for i in range(len(xt)):
        for k in range(3):
            xp, xm = xt.copy(), xt.copy()
            xp[i, k] += h
            xm[i, k] -= h
            grad[i, k] = (nh_energy(xp, tets, b.x0, sc.groups[0].young, sc.groups[0].poisson) -
                          nh_energy(xm, tets, b.x0, sc.groups[0].young, sc.groups[0].poisson)) / (2 * h)
    expect = xt - dt * dt * w[:, None] * grad
    assert E0 > 0 and np.abs(xs - xt).max() > 1e-5  # the stretched beam does get pulled back
    assert np.abs(xs - expect).max() <= 1e-6 * np.abs(xs - xt).max()
