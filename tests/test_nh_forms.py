"""The neo-Hookean kernels evaluate the constraint pair from the rest data Q = R R^T and det R
(R = DmInv) instead of F = Ds R: ||F||^2 = sum_mn (e_m . e_n) Q_mn, F r_k = sum_m Q_mk e_m,
cof(F) r_k = det R (e_{k+1} x e_{k+2}) and det F = det R e0 . (e1 x e2). These identities are
checked here in FP64 against F and its cofactor matrix formed directly, on random and inverted tets.
"""
import numpy as np
import pytest


def _cof(F):
    return np.linalg.det(F) * np.linalg.inv(F).T


@pytest.mark.parametrize("seed", range(8))
def test_rest_gram_forms(seed):
    rng = np.random.default_rng(seed)
    Dm = rng.normal(size=(3, 3))
    if seed % 2:  # inverted rest orientation
        Dm[:, 0] *= -1.0
    R = np.linalg.inv(Dm)  # rows r_k
    Q = R @ R.T
    detR = np.linalg.det(R)
    Ds = Dm + 0.3 * rng.normal(size=(3, 3))
    if seed % 4 == 3:  # inverted deformed tet
        Ds[:, 2] *= -1.0
    e = [Ds[:, k] for k in range(3)]
    F = Ds @ R
    gram = np.array([[e[m] @ e[n] for n in range(3)] for m in range(3)])
    assert np.isclose((gram * Q).sum(), np.sum(F * F), rtol=1e-12)
    for k in range(3):
        h_dev = sum(Q[m, k] * e[m] for m in range(3))
        np.testing.assert_allclose(h_dev, F @ R[k], rtol=1e-11, atol=1e-12)
        h_hyd = np.cross(e[(k + 1) % 3], e[(k + 2) % 3])
        np.testing.assert_allclose(detR * h_hyd, _cof(F) @ R[k], rtol=1e-10, atol=1e-12)
    assert np.isclose(detR * (e[0] @ np.cross(e[1], e[2])), np.linalg.det(F), rtol=1e-12)


def test_vertex_four_gradient_is_minus_the_sum():
    """grad_3 C = -(grad_0 + grad_1 + grad_2) with the fourth vertex the edges' origin."""
    rng = np.random.default_rng(11)
    X = rng.normal(size=(4, 3))
    x = X + 0.2 * rng.normal(size=(4, 3))
    Dm = np.stack([X[k] - X[3] for k in range(3)], axis=1)
    R = np.linalg.inv(Dm)

    def c_dev(p):
        F = np.stack([p[k] - p[3] for k in range(3)], axis=1) @ R
        return np.sqrt(np.sum(F * F))

    g = np.zeros((4, 3))
    h = 1e-6
    for v in range(4):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[v, a] += h
            xm[v, a] -= h
            g[v, a] = (c_dev(xp) - c_dev(xm)) / (2 * h)
    e = [x[k] - x[3] for k in range(3)]
    Q = R @ R.T
    C = c_dev(x)
    for k in range(3):
        np.testing.assert_allclose(g[k], sum(Q[m, k] * e[m] for m in range(3)) / C, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(g[3], -(g[0] + g[1] + g[2]), atol=1e-8)
