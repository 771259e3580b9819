"""Pins of the oracle's grid SDFs (the restatement of Discregrid's cubic Lagrange grid and mesh
distance, absent from the reference tree) by what the mathematics fixes: the element's Kronecker
property at its nodes, partition of unity, exact reproduction of cubic polynomials, gradients equal
to finite differences, continuity across cells; mesh distances against an independent closest-point
formulation, signs against winding numbers, a sphere against its analytic distance; and
environment_body_t's domain growth against its closed form. The library's host-side grid helpers
(node positions, extended domain) are checked against the oracle here too."""
import os

import numpy as np
import pytest

import oracle as O
import sbs_amd as S
from helpers import GOLDEN
from sbs_amd import ply

DBL_MAX = np.finfo(np.float64).max
GRIDS = [([-1.0, 0.5, -0.25], [1.0, 2.0, 1.0], [3, 4, 2]), ([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1, 1, 1]),
         ([-2.0, -1.0, 3.0], [0.5, 4.0, 3.3], [5, 2, 7])]


def cubic(P):
    x, y, z = P[..., 0], P[..., 1], P[..., 2]
    v = 1 + 2 * x - y + 0.5 * z + x * y - 3 * y * z + x * x * z + 0.7 * y ** 3 - z ** 3 + x * y * z + 2 * x ** 3
    g = np.stack([2 + y + 2 * x * z + y * z + 6 * x * x, -1 + x - 3 * z + 2.1 * y * y + x * z,
                  0.5 - 3 * y + x * x - 3 * z * z + x * y], -1)
    return v, g


def grid_sim(lo, hi, res, nodes):
    s = O.Simulation()
    s.add_grid_sdf(lo, hi, res, nodes)
    return s


@pytest.mark.parametrize("g", range(len(GRIDS)))
def test_node_layout_and_kronecker_property(g):
    lo, hi, res = GRIDS[g]
    X = O.grid_node_positions(lo, hi, res)
    assert len(X) == O.grid_node_count(res) == S.grid_node_count(res)
    assert np.array_equal(S.grid_node_positions(lo, hi, res), X)
    # every node is a distinct point of the lattice of thirds of the cells
    h = (np.asarray(hi) - lo) / res
    t = (X - lo) / h * 3
    assert np.allclose(t, np.round(t), atol=1e-9) and len(np.unique(np.round(t), axis=0)) == len(X)
    rng = np.random.default_rng(g)
    for l in rng.choice(len(X), min(len(X), 40), replace=False):
        v = np.zeros(len(X))
        v[l] = 1.0
        sd, _ = grid_sim(lo, hi, res, v).evaluate_sdf(0, X)
        assert np.allclose(sd, v, rtol=0, atol=1e-12)


@pytest.mark.parametrize("g", range(len(GRIDS)))
def test_partition_of_unity_and_cubic_reproduction(g):
    lo, hi, res = GRIDS[g]
    X = O.grid_node_positions(lo, hi, res)
    P = np.random.default_rng(10 + g).uniform(lo, hi, (3000, 3))
    sd, gr = grid_sim(lo, hi, res, np.ones(len(X))).evaluate_sdf(0, P)
    assert np.allclose(sd, 1.0, rtol=0, atol=1e-13) and np.allclose(gr, 0.0, rtol=0, atol=1e-11)
    v, _ = cubic(X)
    sd, gr = grid_sim(lo, hi, res, v).evaluate_sdf(0, P)
    ev, eg = cubic(P)
    scale = np.abs(ev).max()
    assert np.abs(sd - ev).max() <= 1e-12 * scale
    assert np.abs(gr - eg).max() <= 1e-10 * scale


def test_gradient_is_the_derivative_and_field_is_continuous():
    lo, hi, res = GRIDS[2]
    X = O.grid_node_positions(lo, hi, res)
    rng = np.random.default_rng(3)
    s = grid_sim(lo, hi, res, rng.normal(size=len(X)))
    h = (np.asarray(hi) - lo) / res
    P = rng.uniform(np.asarray(lo) + h * 0.1, np.asarray(hi) - h * 0.1, (400, 3))
    # keep points away from cell faces so the central differences stay inside one cell
    frac = (P - lo) / h - np.floor((P - lo) / h)
    P = P[np.all((frac > 0.01) & (frac < 0.99), axis=1)]
    _, gr = s.evaluate_sdf(0, P)
    eps = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = eps * h[d]
        fd = (s.evaluate_sdf(0, P + e)[0] - s.evaluate_sdf(0, P - e)[0]) / (2 * e[d])
        assert np.allclose(gr[:, d], fd, rtol=1e-5, atol=1e-5 * np.abs(gr).max())
    # C0 across the interior cell faces of each axis
    for d in range(3):
        for c in range(1, res[d]):
            Q = rng.uniform(lo, hi, (50, 3))
            Q[:, d] = lo[d] + c * h[d]
            a, _ = s.evaluate_sdf(0, Q - np.eye(3)[d] * 1e-12)
            b, _ = s.evaluate_sdf(0, Q + np.eye(3)[d] * 1e-12)
            assert np.allclose(a, b, rtol=0, atol=1e-9)


def test_outside_the_domain_is_dbl_max():
    lo, hi, res = GRIDS[0]
    s = grid_sim(lo, hi, res, np.zeros(O.grid_node_count(res)))
    P = np.array([[lo[0] - 1e-9, 1.0, 0.0], [0.0, hi[1] + 1e-3, 0.0], [0.0, 1.0, lo[2] - 5.0], lo, hi])
    sd, gr = s.evaluate_sdf(0, P)
    assert list(sd[:3]) == [DBL_MAX] * 3 and np.all(gr[:3] == 0.0)
    assert np.all(sd[3:] == 0.0)  # the domain is closed


def icosphere(r, levels):
    t = (1 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(levels):
        mid = {}
        nf = []

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return r * np.array(v), np.array(f, dtype=np.uint32)


def brute_distance(V, F, P):
    """Unsigned distance by a different formulation: the plane projection when it falls inside the
    triangle (barycentric test), otherwise the nearest of the three edge segments."""
    A, B, C = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    out = np.empty(len(P))
    for i, p in enumerate(P):
        n = np.cross(B - A, C - A)
        n /= np.linalg.norm(n, axis=1)[:, None]
        dpl = np.einsum("ij,ij->i", p - A, n)
        q = p - dpl[:, None] * n
        # barycentric signs of q
        s1 = np.einsum("ij,ij->i", np.cross(B - A, q - A), n)
        s2 = np.einsum("ij,ij->i", np.cross(C - B, q - B), n)
        s3 = np.einsum("ij,ij->i", np.cross(A - C, q - C), n)
        inside = (s1 >= 0) & (s2 >= 0) & (s3 >= 0)
        best = np.where(inside, np.abs(dpl), np.inf)

        def seg(a, b):
            ab = b - a
            t = np.clip(np.einsum("ij,ij->i", p - a, ab) / np.einsum("ij,ij->i", ab, ab), 0, 1)
            return np.linalg.norm(p - (a + t[:, None] * ab), axis=1)
        best = np.minimum(best, np.minimum(seg(A, B), np.minimum(seg(B, C), seg(C, A))))
        out[i] = best.min()
    return out


def winding_number(V, F, P):
    """Generalized winding number (solid angles, Van Oosterom and Strackee 1983)."""
    out = np.empty(len(P))
    for i, p in enumerate(P):
        a, b, c = V[F[:, 0]] - p, V[F[:, 1]] - p, V[F[:, 2]] - p
        la, lb, lc = (np.linalg.norm(x, axis=1) for x in (a, b, c))
        num = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = la * lb * lc + np.einsum("ij,ij->i", a, b) * lc + np.einsum("ij,ij->i", b, c) * la + \
            np.einsum("ij,ij->i", c, a) * lb
        out[i] = np.sum(2 * np.arctan2(num, den)) / (4 * np.pi)
    return out


def bunny():
    g = ply.read_ply(os.path.join(GOLDEN, "data", "meshes", "bunny_simplified.ply"), raise_errors=True)
    return g.positions.astype(np.float64).reshape(-1, 3), g.indices.astype(np.uint32).reshape(-1, 3)


@pytest.mark.parametrize("mesh", ["icosphere", "bunny"])
def test_mesh_distance_against_brute_force_and_winding_numbers(mesh):
    V, F = icosphere(0.7, 2) if mesh == "icosphere" else bunny()
    lo, hi = V.min(0), V.max(0)
    P = np.random.default_rng(5).uniform(lo - 0.2 * (hi - lo), hi + 0.2 * (hi - lo), (300, 3))
    sd = O.mesh_signed_distance(V, F, P)
    assert np.allclose(np.abs(sd), brute_distance(V, F, P), rtol=1e-12, atol=1e-13)
    w = winding_number(V, F, P)
    clear = np.abs(np.abs(w) - 0.5) > 0.2  # points well inside or well outside
    assert np.all((sd[clear] < 0) == (np.abs(w[clear]) > 0.5))


def test_sphere_mesh_grid_approximates_the_sphere():
    r = 0.5
    V, F = icosphere(r, 4)
    lo, hi, res = [-0.8] * 3, [0.8] * 3, [12, 12, 12]
    nodes = O.grid_from_mesh(V, F, lo, hi, res)
    s = grid_sim(lo, hi, res, nodes)
    P = np.random.default_rng(8).uniform(-0.75, 0.75, (2000, 3))
    P = P[np.linalg.norm(P, axis=1) > 0.2]  # the distance has a kink at the centre
    sd, gr = s.evaluate_sdf(0, P)
    exact = np.linalg.norm(P, axis=1) - r
    assert np.abs(sd - exact).max() < 2e-3  # faceting (about r (1 - cos of half an edge angle)) and interpolation
    near = np.abs(exact) > 0.05
    cosang = np.einsum("ij,ij->i", gr[near], P[near]) / (np.linalg.norm(gr[near], axis=1) *
                                                         np.linalg.norm(P[near], axis=1))
    assert cosang.min() > 0.99


def test_environment_domain_growth_closed_form():
    # a cube domain already containing the mesh: each vertex adds 1e-3 |diagonal| to max, then 1e-3
    # |diagonal| of the grown box to min, so each side grows by the factor (1 + 1e-3 sqrt 3)^2 per vertex
    a = 2.0
    V = np.random.default_rng(1).uniform(-0.5, 0.5, (25, 3))
    lo0, hi0 = np.full(3, -a / 2), np.full(3, a / 2)
    for f in (O.environment_domain, S.environment_domain):
        lo, hi = f(V, lo0, hi0)
        assert np.allclose(hi - lo, a * (1 + 1e-3 * 3 ** 0.5) ** (2 * len(V)), rtol=1e-12)
        assert np.all(lo < lo0) and np.all(hi > hi0)
    assert all(np.array_equal(x, y) for x, y in zip(O.environment_domain(V, lo0, hi0), S.environment_domain(V, lo0, hi0)))
    # a domain that does not contain the mesh is first extended to its bounding box
    lo, hi = O.environment_domain(V[:1], [5.0, 5.0, 5.0], [6.0, 6.0, 6.0])
    assert np.all(lo < V[0]) and np.all(hi > 6.0)
    # no vertices: the domain as given
    lo, hi = O.environment_domain(np.zeros((0, 3)), lo0, hi0)
    assert np.array_equal(lo, lo0) and np.array_equal(hi, hi0)
