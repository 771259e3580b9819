"""CPU checks of the oracles of the learned warm start: the NumPy network restatement
(oracle/mlp_oracle.py) against finite differences, a library BF16 conversion and Adam's closed
form, and the FP64 solver's multiplier warm start against the fixed point it must share with the
plain solver."""
import numpy as np
import pytest

import mlp_oracle as M
import oracle as O
from sbs_amd import scenes


NIN = 39  # used inputs of the warm-start network (a hex cluster's features)


def random_params(rng, scale=1.0):
    p = np.zeros(M.NPARAMS)
    for (w, b, fan) in ((M.W1, M.B1, NIN), (M.W2, M.B2, M.H), (M.W3, M.B3, M.H), (M.W4, M.B4, M.H)):
        p[w:b] = rng.normal(0, scale * np.sqrt(2.0 / fan), b - w)
    p[M.B1:M.W2] = rng.normal(0, 0.1, M.H)
    p[M.B2:M.W3] = rng.normal(0, 0.1, M.H)
    p[M.B3:M.W4] = rng.normal(0, 0.1, M.H)
    p[M.B4:] = rng.normal(0, 0.1, M.OUT)
    p[M.W1:M.B1].reshape(M.IN, M.H)[NIN:] = 0.0  # unused inputs
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


@pytest.mark.parametrize("nout", [2, 12])
def test_network_gradient_matches_finite_differences(nout):
    rng = np.random.default_rng(7)
    p = random_params(rng)
    x = rng.normal(0, 1, (9, NIN))
    xn = M.normalize(x, np.zeros(NIN), np.ones(NIN), rounding=False)
    t = rng.normal(0, 1, (9, nout))
    loss, g = M.loss_grad(p, xn, t, nout, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(M.W1, M.W1 + NIN * M.H), 60, replace=False),
                          rng.choice(np.arange(M.W2, M.W4), 120, replace=False),
                          np.arange(M.W4, M.W4 + M.H * M.OUT, M.OUT)[:40] + rng.integers(0, nout),
                          np.arange(M.B1, M.W2)[:20], np.arange(M.B4, M.B4 + nout)])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_grad(pp, xn, t, nout, False)[0] - M.loss_grad(pm, xn, t, nout, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # outputs past nout carry no gradient
    assert np.all(g[M.W4:M.B4].reshape(M.H, M.OUT)[:, nout:] == 0) and np.all(g[M.B4 + nout:] == 0)


def test_rounded_network_stays_close_to_exact():
    rng = np.random.default_rng(3)
    p = random_params(rng)
    x = rng.normal(0, 1, (64, NIN))
    y1, _ = M.forward(p, M.normalize(x, np.zeros(NIN), np.ones(NIN), True), True)
    y0, _ = M.forward(p, M.normalize(x, np.zeros(NIN), np.ones(NIN), False), False)
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


def _hex_beam():
    return scenes.cantilever(5, 2, 2, 0.1, "neohookean", 5e5, 0.45, 1000.0, num_spheres=2, iterations=3, substeps=2,
                             cluster=True, coloring="index")


def _oracle(sc):
    s = O.Simulation()
    sc.apply(s)
    s.set_order(True, 2)
    return s


def test_oracle_network_matches_the_numpy_restatement():
    rng = np.random.default_rng(11)
    p = random_params(rng)
    mean, inv_std = rng.normal(0, 1, NIN), np.exp(rng.normal(0, 0.5, NIN))
    out_scale = np.exp(rng.normal(0, 0.5, 12)).astype(np.float32)
    x = rng.normal(0, 2, (40, NIN))
    s = _oracle(_hex_beam())
    s.set_network(p.astype(np.float32), mean.astype(np.float32), inv_std.astype(np.float32), out_scale)
    y = s.network_forward(x)
    # the parameters the oracle holds are the FP32 values
    p32 = p.astype(np.float32).astype(np.float64)
    ref, _ = M.forward(p32, M.normalize(x, mean.astype(np.float32), inv_std.astype(np.float32), True), True)
    ref = ref[:, :12] * out_scale.astype(np.float64)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_oracle_cluster_features_at_a_deformed_state():
    sc = _hex_beam()
    s = _oracle(sc)
    b, g = sc.bodies[0], sc.groups[0]
    rng = np.random.default_rng(5)
    x = b.x0 + rng.normal(0, 0.01, b.x0.shape)
    s.set_body_state(0, x, np.zeros_like(x))
    dt = sc.dt / sc.substeps
    f = s.cluster_features(0, dt, len(g.clusters) - 1)
    tets = np.asarray(g.conn).reshape(-1, 6, 4)
    mu = g.young / (2 * (1 + g.poisson))
    la = g.young * g.poisson / ((1 + g.poisson) * (1 - 2 * g.poisson))
    Dm = np.stack([b.x0[tets[..., k]] - b.x0[tets[..., 3]] for k in range(3)], axis=-1)
    Ds = np.stack([x[tets[..., k]] - x[tets[..., 3]] for k in range(3)], axis=-1)
    F = Ds @ np.linalg.inv(Dm)
    V0 = np.abs(np.linalg.det(Dm)) / 6
    atd, ath = 1 / (mu * V0) / dt ** 2, 1 / (la * V0) / dt ** 2
    assert np.allclose(f[:, 24::2][:, :6], np.sqrt(np.sum(F * F, axis=(-2, -1))), rtol=1e-12)
    assert np.allclose(f[:, 25::2][:, :6], np.linalg.det(F) - (1 + mu / la), rtol=1e-10, atol=1e-12)
    assert np.allclose(f[:, 36], np.log(atd[:, 0]), rtol=1e-12) and np.allclose(f[:, 37], np.log(ath[:, 0]), rtol=1e-12)
    assert np.all(f[:, 38] == np.log(len(g.conn)))
    assert np.all(f[:, :24] == 0.0)  # multipliers start at zero


def test_oracle_learned_warm_start_with_a_silent_network_is_the_previous_start():
    # a network whose outputs are all zero predicts no change: the learned start is the previous
    # one, and its history holds the multipliers each substep started from
    sc = _hex_beam()
    a, b = _oracle(sc), _oracle(sc)
    a.set_warm_start("previous", 0.6, 0.5)
    b.set_network(np.zeros(M.NPARAMS, np.float32), np.zeros(NIN, np.float32), np.ones(NIN, np.float32),
                  np.ones(12, np.float32))
    b._L.orc_set_warm_start(b.h, 2, 0.6, 0.5)
    n = len(sc.groups[0].conn)
    for _ in range(3):
        before = a.multipliers(0, n)
        a.step(sc.dt, sc.iterations, 1)
        b.step(sc.dt, sc.iterations, 1)
        assert np.array_equal(a.read_body(0)[0], b.read_body(0)[0])
        mb = b.multipliers(0, n)
        assert np.array_equal(mb[:, :2], a.multipliers(0, n)[:, :2])
        assert np.array_equal(mb[:, 2:], before[:, :2])


def test_oracle_learned_features_carry_the_multipliers():
    # after warm-started substeps the first 24 features are alpha~ times the stored pairs and the
    # pairs one substep before
    sc = _hex_beam()
    s = _oracle(sc)
    s.set_network(np.zeros(M.NPARAMS, np.float32), np.zeros(NIN, np.float32), np.ones(NIN, np.float32),
                  np.ones(12, np.float32))
    s._L.orc_set_warm_start(s.h, 2, 0.6, 0.5)
    for _ in range(2):
        s.step(sc.dt, sc.iterations, 1)
    g = sc.groups[0]
    dt = sc.dt
    f = s.cluster_features(0, dt, len(g.clusters) - 1)
    m = s.multipliers(0, len(g.conn)).reshape(-1, 6, 4)
    b = sc.bodies[0]
    tets = np.asarray(g.conn).reshape(-1, 6, 4)
    mu = g.young / (2 * (1 + g.poisson))
    la = g.young * g.poisson / ((1 + g.poisson) * (1 - 2 * g.poisson))
    Dm = np.stack([b.x0[tets[..., k]] - b.x0[tets[..., 3]] for k in range(3)], axis=-1)
    V0 = np.abs(np.linalg.det(Dm)) / 6
    atd, ath = 1 / (mu * V0) / dt ** 2, 1 / (la * V0) / dt ** 2
    assert np.abs(m[..., 0]).max() > 0 and np.abs(m[..., 2]).max() > 0
    assert np.allclose(f[:, 0:12:2], atd * m[..., 0], rtol=1e-12)
    assert np.allclose(f[:, 1:12:2], ath * m[..., 1], rtol=1e-12)
    assert np.allclose(f[:, 12:24:2], atd * m[..., 2], rtol=1e-12)
    assert np.allclose(f[:, 13:24:2], ath * m[..., 3], rtol=1e-12)


def test_oracle_network_waits_for_carried_multipliers():
    # the network corrects carried-over multipliers: from rest the first substep has none and the
    # second none from the substep before, so both equal the previous start exactly, whatever the
    # network says; from the third substep on a loud network changes the result
    sc = _hex_beam()
    a, b = _oracle(sc), _oracle(sc)
    a.set_warm_start("previous", 0.6, 0.5)
    rng = np.random.default_rng(17)
    b.set_network(random_params(rng).astype(np.float32), np.zeros(NIN, np.float32), np.ones(NIN, np.float32),
                  np.full(12, 1e-3, np.float32))
    b._L.orc_set_warm_start(b.h, 2, 0.6, 0.5)
    for s in range(3):
        a.step(sc.dt, sc.iterations, 1)
        b.step(sc.dt, sc.iterations, 1)
        same = np.array_equal(a.read_body(0)[0], b.read_body(0)[0])
        assert same == (s < 2), s
