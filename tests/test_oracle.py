"""Pins the FP64 oracle against the reference's fixtures, closed forms, library routines and brute
force (CPU only). Every check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, a transposed operand, a wrong face order) fails it."""
import itertools
import os

import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN, read_ply_tets, tets_conflict_free

RNG = np.random.default_rng(1234)


# ------------------------------------------------------------------ linear algebra

def test_svd3_matches_numpy():
    mats = [RNG.normal(size=(3, 3)) for _ in range(300)]
    mats += [np.eye(3) + 1e-3 * RNG.normal(size=(3, 3)) for _ in range(100)]
    mats += [np.diag([2.0, 2.0, 0.5]), np.diag([1.0, 1.0, 1.0]), np.zeros((3, 3)), np.outer([1, 2, 3], [0, 1, 1.0])]
    for F in mats:
        U, s, V = O.svd3(F)
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-12)
        assert np.allclose(U.T @ U, np.eye(3), atol=1e-12)
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.allclose(s, np.linalg.svd(F, compute_uv=False), atol=1e-12)
        assert s[0] >= s[1] >= s[2] >= 0.0


def test_inverse3_and_determinant():
    for _ in range(100):
        A = RNG.normal(size=(3, 3))
        inv, det = O.inverse3(A)
        assert np.allclose(inv, np.linalg.inv(A), rtol=1e-10, atol=1e-10)
        assert np.isclose(det, np.linalg.det(A), rtol=1e-12)


# ------------------------------------------------------------------ meshes (reference fixtures)

@pytest.mark.parametrize("fixture,dims", [("bar_tet.ply", (12, 4, 4)), ("tet_bar_5x2x2.ply", (5, 2, 2)),
                                          ("cube_tet.ply", (2, 2, 2))])
def test_bar5_reproduces_reference_fixtures(fixture, dims):
    # data/meshes/*.ply hold outputs of get_simple_bar_model; index-exact equality pins the lattice
    # order and both parity splits.
    pos, tets = read_ply_tets(os.path.join(GOLDEN, fixture))
    p, t = O.bar5(*dims)
    assert np.array_equal(p, pos)
    assert np.array_equal(t, tets)


def _tet_volumes(p, t):
    e = p[t[:, :3]] - p[t[:, 3:4]]
    return np.linalg.det(np.transpose(e, (0, 2, 1))) / 6.0


def test_bar5_orientation_and_volume():
    p, t = O.bar5(12, 4, 4)
    V = _tet_volumes(p, t)
    assert np.all(V < 0)
    assert np.isclose(np.abs(V).sum(), 11 * 3 * 3)
    assert set(np.round(np.abs(V) * 6).astype(int)) == {1, 2}


def test_box6_volume_orientation_and_conformity():
    p, t = O.box6(4, 3, 2, 0.25, (1.0, -2.0, 0.5))
    V = _tet_volumes(p, t)
    assert np.all(V < 0)
    assert np.isclose(np.abs(V).sum(), 1.0 * 0.75 * 0.5)
    # conforming: every face is shared by at most 2 tets, interior faces by exactly 2
    faces = {}
    for tet in t:
        for f in itertools.combinations(sorted(tet), 3):
            faces[f] = faces.get(f, 0) + 1
    assert max(faces.values()) == 2
    boundary = [f for f, c in faces.items() if c == 1]
    # boundary of a 4x3x2 box: 2 triangles per boundary quad
    assert len(boundary) == 2 * 2 * (4 * 3 + 4 * 2 + 3 * 2)


def _brute_boundary(nv, tets):
    """Reference semantics restated from scratch: faces numbered in first-insertion order over
    tets (faces_copy order v1v2v4, v2v3v4, v3v1v4, v1v3v2), boundary = faces not shared by
    exactly two tets, surface vertices numbered by first appearance over boundary faces."""
    order, count, first = [], {}, {}
    for tet in tets:
        v1, v2, v3, v4 = [int(x) for x in tet]
        for f in ((v1, v2, v4), (v2, v3, v4), (v3, v1, v4), (v1, v3, v2)):
            k = tuple(sorted(f))
            if k not in count:
                count[k] = 0
                first[k] = f
                order.append(k)
            count[k] += 1
    s2t, tris, index = [], [], {}
    for k in order:
        if count[k] == 2:
            continue
        tri = []
        for v in first[k]:
            if v not in index:
                index[v] = len(s2t)
                s2t.append(v)
            tri.append(index[v])
        tris.append(tri)
    return np.array(s2t), np.array(tris).reshape(-1, 3)


@pytest.mark.parametrize("fixture", ["tetrahedron.ply", "2tets.ply", "3tets.ply", "cube_tet.ply", "bar_tet.ply"])
def test_boundary_matches_brute_force(fixture):
    pos, tets = read_ply_tets(os.path.join(GOLDEN, fixture))
    s2t, tris, _ = O.boundary(len(pos), tets)
    bs, bt = _brute_boundary(len(pos), tets)
    assert np.array_equal(s2t, bs)
    assert np.array_equal(tris, bt)


def test_boundary_counts_of_benchmark_beam():
    p, t = O.box6(10, 3, 3, 0.1)
    s2t, tris, _ = O.boundary(len(p), t)
    assert (len(p), len(t), len(s2t), len(tris)) == (176, 540, 140, 276)


# ------------------------------------------------------------------ coloring

def _greedy_property(verts, color, passes):
    """Checks the defining property of greedy coloring in priority order (pass `passes`): each
    constraint's color is the smallest color not taken by a higher-priority neighbour."""
    n = len(verts)
    prev = np.zeros(n, dtype=np.uint64)
    # recompute the previous pass with the oracle itself, then check the last pass by brute force
    if passes > 0:
        prev, _ = O.greedy_color(verts, passes - 1)
    h = np.array([O.hash32(i) for i in range(n)], dtype=np.uint64)
    key = (prev.astype(np.uint64) << np.uint64(32)) | h if passes > 0 else h
    by_vertex = {}
    for i, vs in enumerate(verts):
        for v in vs:
            by_vertex.setdefault(int(v), []).append(i)
    for i in range(n):
        nbrs = {j for v in verts[i] for j in by_vertex[int(v)] if j != i}
        taken = {int(color[j]) for j in nbrs if key[j] > key[i]}
        c = 0
        while c in taken:
            c += 1
        assert color[i] == c, (i, color[i], c)


@pytest.mark.parametrize("passes", [0, 1, 2])
def test_greedy_coloring_is_valid_and_greedy(passes):
    p, t = O.box6(5, 3, 3, 0.1)
    verts = t.astype(np.uint64)
    color, nc = O.greedy_color(verts, passes)
    assert nc == color.max() + 1
    assert tets_conflict_free(t, color)
    _greedy_property(verts, color, passes)


def test_recoloring_never_adds_colors():
    p, t = O.box6(12, 4, 4, 0.1)
    counts = [O.greedy_color(t.astype(np.uint64), k)[1] for k in range(4)]
    assert all(b <= a for a, b in zip(counts, counts[1:]))
    assert counts[0] >= 24  # 24 tets meet at an interior lattice vertex


# ------------------------------------------------------------------ green constraint

def _tet_x0():
    return np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0], [0.0, 0.0, 0.1]])


def _deform(x0, F):
    return x0 @ np.asarray(F).T


def _stvk(F, mu, lam):
    E = 0.5 * (F.T @ F - np.eye(3))
    return mu * np.sum(E * E) + 0.5 * lam * np.trace(E) ** 2, F @ (2 * mu * E + lam * np.trace(E) * np.eye(3))


def _lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def _dminv_v0(x0):
    Dm = (x0[:3] - x0[3]).T
    return np.linalg.inv(Dm), np.linalg.det(Dm) / 6.0


def test_green_rest_state_has_zero_value_and_force():
    x0 = _tet_x0()
    C, f = O.green_eval(x0, x0, 1e6, 0.3)
    assert C == 0.0 and np.all(f == 0.0)
    # rigid rotation + translation: F = R, singular values 1 -> still zero (clamp inactive)
    R = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
    R *= np.sign(np.linalg.det(R))
    C, f = O.green_eval(x0, _deform(x0, R) + 0.3, 1e6, 0.3)
    assert abs(C) < 1e-18 and np.abs(f).max() < 1e-9


@pytest.mark.parametrize("stretch", [(1.1, 0.95, 1.02), (1.3, 1.2, 0.7), (0.9, 0.9, 0.9)])
def test_green_pure_stretch_is_stvk_energy_and_gradient(stretch):
    # With F symmetric positive definite U = V, so the reference's tr(U E^ V^T) equals tr(E) and
    # C = |V0| psi_StVK; the forces must then be minus its gradient (central differences).
    mu, lam = _lame(1e6, 0.3)
    x0 = _tet_x0()
    Q = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
    F = Q @ np.diag(stretch) @ Q.T
    p = _deform(x0, F)
    C, f = O.green_eval(x0, p, 1e6, 0.3)
    DmInv, V0 = _dminv_v0(x0)
    psi, _ = _stvk(F, mu, lam)
    assert np.isclose(C, abs(V0) * psi, rtol=1e-10)
    h = 1e-7
    for k in range(4):
        for d in range(3):
            pp, pm = p.copy(), p.copy()
            pp[k, d] += h
            pm[k, d] -= h
            g = (O.green_value(x0, pp, 1e6, 0.3) - O.green_value(x0, pm, 1e6, 0.3)) / (2 * h)
            assert np.isclose(-f[k, d], g, rtol=1e-5, atol=1e-6 * abs(C) / 0.1)
    assert np.allclose(f.sum(0), 0.0, atol=1e-9)


def test_green_force_is_stvk_first_piola_for_rotated_stretch():
    # For an unclamped, non-inverted F the mapped-back stress U P^ V^T is the StVK first Piola
    # stress F(2 mu E + lam tr(E) I) whatever the rotation; forces are -|V0| P DmInv^T.
    mu, lam = _lame(5e5, 0.45)
    x0 = _tet_x0() + np.array([0.3, -0.1, 0.2])
    for _ in range(20):
        F = np.eye(3) + 0.15 * RNG.normal(size=(3, 3))
        if np.linalg.det(F) <= 0 or np.linalg.svd(F, compute_uv=False).min() <= 0.6:
            continue
        p = _deform(x0, F)
        C, f = O.green_eval(x0, p, 5e5, 0.45)
        DmInv, V0 = _dminv_v0(x0)
        _, P = _stvk(F, mu, lam)
        H = -abs(V0) * P @ DmInv.T
        assert np.allclose(f[:3], H.T, rtol=1e-9, atol=1e-9 * np.abs(H).max())
        # the reference's psi uses tr(U E^ V^T) = sum e_i (u_i . v_i) instead of tr(E)
        U, s, Vt = np.linalg.svd(F)
        e = 0.5 * (s * s - 1)
        tq = np.sum(e * np.einsum("ij,ij->j", U, Vt.T))
        assert np.isclose(C, abs(V0) * (mu * np.sum(e * e) + 0.5 * lam * tq * tq), rtol=1e-9)


def test_green_clamped_and_inverted_closed_forms():
    mu, lam = _lame(1e6, 0.3)
    x0 = _tet_x0()
    DmInv, V0 = _dminv_v0(x0)
    # clamped: sigma_3 = 0.4 < 0.577 is raised to 0.577
    for diag, fhat in [((1.1, 0.9, 0.4), (1.1, 0.9, 0.577)),
                       # inverted: det F < 0 -> smallest singular value negated, then clamped
                       ((1.2, 1.0, -0.8), (1.2, 1.0, 0.577)),
                       ((-1.05, 1.0, 1.1), (1.1, 1.05, 0.577))]:
        d = np.array(diag)
        F = np.diag(d)
        C, f = O.green_eval(x0, _deform(x0, F), 1e6, 0.3)
        # For diagonal F the singular vectors are +-e_a along axis a = order[i]:
        # u_i . v_i = sign(d_a), and the inversion flip of u_3 negates the last one.
        order = np.argsort(-np.abs(d))
        inverted = np.prod(d) < 0
        sgn = np.sign(d[order])
        if inverted:
            sgn[2] = -sgn[2]
        sv = np.abs(d[order]) * np.array([1.0, 1.0, -1.0 if inverted else 1.0])
        fh = np.maximum(sv, 0.577)
        assert np.allclose(fh, fhat)
        e = 0.5 * (fh * fh - 1)
        tq = np.sum(e * sgn)  # the reference's tr(U E^ V^T)
        assert np.isclose(C, abs(V0) * (mu * np.sum(e * e) + 0.5 * lam * tq ** 2), rtol=1e-12)
        ph = fh * (2 * mu * e + lam * np.sum(e))
        P = np.zeros((3, 3))
        for i, ax in enumerate(order):
            P[ax, ax] = ph[i] * sgn[i]
        H = -abs(V0) * P @ DmInv.T
        assert np.allclose(f[:3], H.T, rtol=1e-10, atol=1e-10)


def _one_tet_scene(x0, p, kind, w=(1.0, 1.0, 1.0, 1.0), alpha=0.0, beta=0.0, young=1e6, poisson=0.3):
    sim = O.Simulation()
    sim.set_params(gravity=(0.0, 0.0, 0.0), gravity_is_force=True)
    mass = np.array([1.0 / x if x > 0 else 0.0 for x in w])
    b = sim.add_body(x0, mass)
    sim.set_body_state(b, x=p)
    t = np.array([[0, 1, 2, 3]])
    if kind == "green":
        sim.add_green(b, t, alpha, beta, young, poisson)
    else:
        sim.add_neohookean(b, t, beta, young, poisson)
    sim.set_order(True, 2)
    return sim


@pytest.mark.parametrize("alpha", [0.0, 1e-4])
def test_green_projection_follows_the_xpbd_update(alpha):
    # XPBD (Macklin, Mueller, Chentanez 2016, eq. 18) with the multiplier reset at the start of the
    # substep and accumulated over iterations: dl = -(C + a~ l) / (sum w |grad C|^2 + a~),
    # x += w grad C dl, grad C = -f. Forces come from green_eval, pinned above.
    x0 = _tet_x0()
    F = np.eye(3) + np.array([[0.02, 0.01, 0.0], [-0.01, -0.015, 0.005], [0.0, 0.02, 0.01]])
    p = _deform(x0, F)
    w = np.array([1.0, 0.5, 2.0, 1.0])
    dt = 0.01
    sim = _one_tet_scene(x0, p, "green", w=w, alpha=alpha)
    sim.step(dt, 3, 1)
    x, v = sim.read_body(0)
    q, lam, at = p.copy(), 0.0, alpha / dt ** 2
    Cs = []
    for _ in range(3):
        C, f = O.green_eval(x0, q, 1e6, 0.3)
        Cs.append(C)
        dl = -(C + at * lam) / (np.sum(w[:, None] * f * f) + at)
        lam += dl
        q = q - w[:, None] * f * dl
    assert np.allclose(x, q, rtol=0, atol=1e-14)
    assert Cs[1] < Cs[0] and Cs[2] < Cs[1]
    assert np.allclose(v, (x - p) / dt)


def test_green_rest_state_takes_the_early_out():
    x0 = _tet_x0()
    sim = _one_tet_scene(x0, x0, "green", alpha=1e-4)
    sim.step(0.01, 3, 1)
    x, _ = sim.read_body(0)
    assert np.array_equal(x, x0)


# ------------------------------------------------------------------ neo-Hookean constraint

def _nh_replay(x0, p, w, young, poisson, dt):
    """One projection of the two neo-Hookean constraints, with gradients taken by central
    differences of the closed forms C_D = ||F||, C_H = det F - (1 + mu/lam)."""
    mu, lam = _lame(young, poisson)
    DmInv, V0 = _dminv_v0(x0)
    w = np.asarray(w)

    def F_of(q):
        return (q[:3] - q[3]).T @ DmInv

    cs = [lambda q: np.linalg.norm(F_of(q)), lambda q: np.linalg.det(F_of(q)) - (1 + mu / lam)]
    comp = [1 / (mu * abs(V0)), 1 / (lam * abs(V0))]
    q = p.copy()
    for C, a in zip(cs, comp):
        g = np.zeros((4, 3))
        for k in range(4):
            for d in range(3):
                qp, qm = q.copy(), q.copy()
                qp[k, d] += 1e-7
                qm[k, d] -= 1e-7
                g[k, d] = (C(qp) - C(qm)) / 2e-7
        at = a / dt ** 2
        dl = -C(q) / (np.sum(w[:, None] * g * g) + at)
        q = q + w[:, None] * g * dl
    return q


def test_neohookean_projection_matches_finite_difference_replay():
    x0 = _tet_x0()
    for _ in range(5):
        F = np.eye(3) + 0.1 * RNG.normal(size=(3, 3))
        p = _deform(x0, F) + 0.2
        w = RNG.uniform(0.5, 2.0, size=4)
        sim = _one_tet_scene(x0, p, "neohookean", w=w, young=5e5, poisson=0.45)
        sim.step(0.01, 1, 1)
        x, _ = sim.read_body(0)
        ref = _nh_replay(x0, p, w, 5e5, 0.45, 0.01)
        assert np.allclose(x, ref, atol=1e-7)


def test_neohookean_rest_state_is_stationary_for_the_combined_energy():
    # psi = mu/2 (|F|^2 - 3) + lam/2 (det F - 1 - mu/lam)^2 has zero gradient at F = I
    mu, lam = _lame(5e5, 0.45)
    F = np.eye(3)
    grad = mu * F + lam * (np.linalg.det(F) - 1 - mu / lam) * np.eye(3)
    assert np.allclose(grad, 0.0)


# ------------------------------------------------------------------ distance / contacts / integration

def test_distance_constraint_restores_rest_length():
    x0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    sim = O.Simulation()
    sim.set_params(gravity=(0.0, 0.0, 0.0))
    b = sim.add_body(x0)
    sim.set_body_state(b, x=np.array([[0.0, 0.0, 0.0], [1.5, 0.5, 0.0]]))
    sim.add_distance(np.array([[0, 0, 0, 1]]), 0.0, 0.0)
    sim.set_order(False)
    sim.step(0.01, 1, 1)
    x, _ = sim.read_body(0)
    assert np.isclose(np.linalg.norm(x[1] - x[0]), 1.0, rtol=1e-12)
    assert np.allclose(x.mean(0), [0.75, 0.25, 0.0])


def test_free_fall_closed_form():
    # no constraints, unit mass, gravity as a 9.81 N force: v_n = -9.81 n dt,
    # y_n = y0 - 9.81 dt^2 n (n + 1) / 2 (timestep.cpp:35-57)
    sim = O.Simulation()
    sim.set_params(gravity=(0.0, -9.81, 0.0), gravity_is_force=True)
    sim.add_body(np.array([[0.0, 2.0, 0.0], [1.0, 3.0, 0.0]]))
    sim.set_order(False)
    dt, n = 0.01, 7
    for _ in range(n):
        sim.step(dt, 3, 1)
    x, v = sim.read_body(0)
    assert np.allclose(v[:, 1], -9.81 * n * dt, rtol=1e-12)
    assert np.allclose(x[:, 1], np.array([2.0, 3.0]) - 9.81 * dt * dt * n * (n + 1) / 2, rtol=1e-12)


def test_plane_contact_projects_onto_plane():
    # a particle below the plane with a tiny compliance ends on it (orthogonal projection)
    sim = O.Simulation()
    sim.set_params(gravity=(0.0, 0.0, 0.0), collision_compliance=0.0)
    b = sim.add_body(np.array([[0.3, -0.2, 0.1]]))
    sim.set_surface(b, np.array([0]))
    n = np.array([0.0, 1.0, 1.0]) / np.sqrt(2)
    sim.add_plane(n, (0.0, 0.0, 0.0))
    sim.set_order(False)
    ids, qn = sim.detect_contacts()
    assert len(ids) == 1
    p = np.array([0.3, -0.2, 0.1])
    sd = p @ n
    assert np.allclose(qn[0, :3], p - sd * n) and np.allclose(qn[0, 3:], n)
    sim.step(0.01, 1, 1)
    x, _ = sim.read_body(0)
    assert np.isclose(x[0] @ n, 0.0, atol=1e-15)


def test_contact_set_matches_brute_force_scan():
    p, t = O.box6(6, 3, 3, 0.1, (0.0, -0.15, 0.0))
    s2t, _, _ = O.boundary(len(p), t)
    sim = O.Simulation()
    b = sim.add_body(p)
    sim.set_surface(b, s2t)
    planes = [((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (-1.0, -5.0, -1.0), (0.45, 5.0, 1.0))]
    spheres = [((0.3, 0.1, 0.15), 0.12), ((0.55, -0.1, 0.0), 0.08)]
    for pl in planes:
        sim.add_plane(*pl)
    for c, r in spheres:
        sim.add_sphere(c, r)
    ids, qn = sim.detect_contacts()
    expect = []
    for k, sph in enumerate([None] + spheres):
        for s, vi in enumerate(s2t):
            x = p[vi]
            if k == 0:
                n, pt, lo, hi = planes[0]
                sd = np.dot(n, x - pt)
                inside = np.all(x >= lo) and np.all(x <= hi)
            else:
                sd = np.linalg.norm(x - np.array(sph[0])) - sph[1]
                inside = True
            if sd < 0 and inside:
                expect.append((0, s, vi, k))
    assert [tuple(r) for r in ids] == expect


def test_colored_and_insertion_order_agree_when_constraints_are_disjoint():
    # tets sharing no vertex: any visiting order gives the same result bit for bit
    x0 = np.concatenate([_tet_x0() + [0.5 * i, 0, 0] for i in range(6)])
    t = np.arange(24).reshape(6, 4)
    out = []
    for colored in (False, True):
        sim = O.Simulation()
        sim.add_body(x0)
        sim.add_green(0, t, 1e-4, 0.0, 1e6, 0.3)
        sim.set_order(colored)
        for _ in range(3):
            sim.step(0.01, 4, 2)
        out.append(sim.read_body(0)[0])
    assert np.array_equal(out[0], out[1])


# ------------------------------------------------------------------ clusters

def test_index_order_coloring_of_hex_clusters_is_the_lattice_parity():
    # sequential first-fit in index order over the cells of a box colors a cell by the parity of
    # its lattice coordinates: 8 colors, cells sharing a vertex never equal
    nx, ny, nz = 5, 4, 3
    p, t = O.box6(nx, ny, nz, 0.1)
    sim = O.Simulation()
    sim.add_body(p)
    g = sim.add_green(0, t, 1e-4, 0.0, 1e6, 0.3)
    sim.set_clusters(g, np.arange(0, len(t) + 1, 6))
    sim.set_coloring("index")
    sim.set_order(True)
    color, nc = sim.group_colors(g, len(t))
    assert nc == 8
    cell = np.array([(i, j, k) for i in range(nx) for j in range(ny) for k in range(nz)])
    cc = color[::6]
    assert np.all(color.reshape(-1, 6) == cc[:, None])
    # cells with equal parity vector have equal colors
    par = (cell % 2) @ np.array([4, 2, 1])
    for q in range(8):
        assert len(set(cc[par == q])) == 1


def test_cluster_hash_coloring_is_greedy_over_clusters():
    p, t = O.box6(4, 3, 3, 0.1)
    sim = O.Simulation()
    sim.add_body(p)
    g = sim.add_green(0, t, 1e-4, 0.0, 1e6, 0.3)
    sim.set_clusters(g, np.arange(0, len(t) + 1, 6))
    sim.set_order(True, 1)
    color, nc = sim.group_colors(g, len(t))
    cells = [np.unique(t[6 * i:6 * i + 6]).astype(np.uint64) for i in range(len(t) // 6)]
    width = max(len(c) for c in cells)
    verts = np.array([np.concatenate([c, c[:1].repeat(width - len(c))]) for c in cells])
    _greedy_property(verts, color[::6], 1)


def test_singleton_clusters_equal_the_unclustered_order():
    p, t = O.box6(3, 2, 2, 0.1)
    out = []
    for clustered in (False, True):
        sim = O.Simulation()
        sim.set_params(gravity=(0, -9.81, 0), gravity_is_force=False)
        m = np.ones(len(p))
        m[p[:, 0] == 0] = 0
        sim.add_body(p, m)
        g = sim.add_neohookean(0, t, 0.0, 5e5, 0.45)
        if clustered:
            sim.set_clusters(g, np.arange(len(t) + 1))
        sim.set_order(True, 2)
        for _ in range(3):
            sim.step(1 / 60, 4, 2)
        out.append(sim.read_body(0)[0])
    assert np.array_equal(out[0], out[1])


def test_batched_group_equals_one_group_per_body():
    # bodies share no particle, so visiting one batched group equals visiting per-body groups
    meshes = [O.box6(3, 2, 2, 0.1, (0.0, 0.5, 0.0)), O.box6(2, 2, 3, 0.1, (1.0, 0.7, 0.0))]
    E, nu = [3e5, 8e5], [0.35, 0.45]
    out = []
    for batched in (False, True):
        sim = O.Simulation()
        sim.set_params(gravity=(0, -9.81, 0), gravity_is_force=False)
        for p, t in meshes:
            sim.add_body(p, np.full(len(p), 0.05))
        if batched:
            bodies = np.concatenate([np.full(len(t), b) for b, (p, t) in enumerate(meshes)])
            tets = np.concatenate([t for p, t in meshes])
            sim.add_neohookean_batched(bodies, tets, 0.0, np.repeat(E, [len(t) for p, t in meshes]),
                                       np.repeat(nu, [len(t) for p, t in meshes]))
        else:
            for b, (p, t) in enumerate(meshes):
                sim.add_neohookean(b, t, 0.0, E[b], nu[b])
        sim.set_order(False)
        for _ in range(4):
            sim.step(1 / 60, 5, 2)
        out.append([sim.read_body(b)[0] for b in range(2)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("warm", [0, 1])
def test_threaded_colored_order_is_bit_identical(warm):
    # the elements of one color share no particle, so running them on several threads must give the
    # sequential colored order's bits; the warm-start pass sums its displacements in constraint order
    from sbs_amd import scenes
    from helpers import build_oracle

    sc = scenes.cantilever(12, 4, 4, 0.1, "neohookean", 5e5, 0.45, 1000.0, num_spheres=4, iterations=4, substeps=2,
                           cluster=True, coloring="index")
    a, b = build_oracle(sc), build_oracle(sc)
    b.set_threads(4)
    for s in (a, b):
        if warm:
            s.set_warm_start("previous", 0.5)
    for _ in range(3):
        a.step(sc.dt, sc.iterations, sc.substeps)
        b.step(sc.dt, sc.iterations, sc.substeps)
    xa, va = a.read_body(0)
    xb, vb = b.read_body(0)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert not np.array_equal(xa, sc.bodies[0].x0)
