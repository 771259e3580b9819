"""The oracle's contacts between bodies (oracle/xpbd.cpp, simulation_t::detect_body_contacts) on
the CPU: the grid's nearest triangles against brute force over every triangle, the inside/outside
decision and the normals on placements whose answer is known from the geometry, and the surface
of a resting convex body never holding its own vertices."""
import numpy as np
import pytest

from helpers import build_oracle
from sbs_amd import _lib, scenes
from sbs_amd.scenes import Body, Group, Scene, hex_clusters


def _contacts(sc, brute_force=False, frames=0):
    o = build_oracle(sc)
    o.set_body_contacts(True, *(sc.body_contacts or (0.0, 0.0, 0.0)), brute_force=brute_force)
    for _ in range(frames):
        o.step(sc.dt, sc.iterations, sc.substeps)
    ids, qn = o.detect_contacts()
    return o, ids, qn


def _two_boxes(offset, small=(2, 2, 2), big=(4, 2, 4), edge=0.05):
    """A big box on the ground and a small box whose lowest corner sits at `offset` from the big box's
    top face's lowest corner (no gravity needed: contacts only)."""
    sc = Scene("two_boxes", gravity_is_force=False, iterations=4, substeps=1, coloring="index")
    a, at = _lib.box6(*big, edge)
    b, bt = _lib.box6(*small, edge)
    b = b + np.array([0.0, a[:, 1].max(), 0.0]) + np.asarray(offset, dtype=np.float64)
    conn = []
    for k, (x0, tets) in enumerate(((a, at), (b, bt))):
        s2t, tris = _lib.boundary(len(x0), tets)
        sc.bodies.append(Body(x0=x0, tets=tets, mass=_lib.lumped_mass(x0, tets, 1000.0), surface=s2t,
                              triangles=tris))
        conn.append((k, tets))
    allt = np.concatenate([t for _, t in conn])
    sc.groups.append(Group("neohookean", np.concatenate([np.full(len(t), k, np.uint32) for k, t in conn]), allt,
                           young=np.full(len(allt), 1e5), poisson=np.full(len(allt), 0.4),
                           clusters=hex_clusters(len(allt))))
    sc.body_contacts = (0.0, 0.0, 0.0)
    return sc


def _nsdf(sc):
    return len(sc.planes) + len(sc.spheres)


@pytest.mark.parametrize("name", ["tower", "tower_big_block", "channel"])
def test_grid_and_brute_force_pick_the_same_triangles(name):
    sc = {"tower": scenes.tower, "tower_big_block": lambda: scenes.tower(big_block=True),
          "channel": lambda: scenes.channel(young=1e5, gap=1)}[name]()
    for frames in (0, 12):
        _, ids_g, qn_g = _contacts(sc, False, frames)
        _, ids_b, qn_b = _contacts(sc, True, frames)
        assert np.array_equal(ids_g, ids_b) and np.array_equal(qn_g, qn_b)
        assert np.sum(ids_g[:, 3] >= _nsdf(sc)) > 20


def test_a_vertex_inside_another_body_is_pushed_out_through_the_nearest_face():
    # the small box sunk 1 cm into the big box's top face, 5 cm in from its edges: the small box's
    # bottom vertices lie inside the big box (normal +y, out through the top face) and the big box's
    # top vertices under the small box lie inside the small box (normal -y)
    sc = _two_boxes((0.05, -0.01, 0.05))
    _, ids, qn = _contacts(sc)
    small_bottom = sc.bodies[1].x0[:, 1] < sc.bodies[1].x0[:, 1].min() + 1e-9
    body = ids[:, 3] >= _nsdf(sc)
    hits = [(i, n) for i, n in zip(ids[body], qn[body, 3:]) if i[0] == 1 and small_bottom[i[2]]]
    assert len(hits) == 9  # the 3 x 3 bottom vertices
    for i, n in hits:
        assert np.allclose(n, [0.0, 1.0, 0.0], atol=1e-12), (i, n)
    big = sc.bodies[0].x0
    under = (np.abs(big[:, 1] - big[:, 1].max()) < 1e-9) & (big[:, 0] > 0.05 + 1e-9) & (big[:, 0] < 0.15 - 1e-9) \
        & (big[:, 2] > 0.05 + 1e-9) & (big[:, 2] < 0.15 - 1e-9)
    inner = [(i, n) for i, n in zip(ids[body], qn[body, 3:]) if i[0] == 0 and under[i[2]]]
    assert len(inner) == int(under.sum()) == 1  # the big box's top vertex strictly inside the footprint
    for i, n in inner:
        assert np.allclose(n, [0.0, -1.0, 0.0], atol=1e-12), (i, n)


def test_a_vertex_past_a_convex_edge_is_pushed_away_from_the_edge():
    # the small box beside the big box's +x top edge, 2 mm above and 2 mm out: its lowest inner
    # vertices are nearest to that edge, outside; the normal runs from the edge point to the vertex
    edge_x = 0.2
    sc = _two_boxes((edge_x + 0.002, 0.002, 0.05))
    _, ids, qn = _contacts(sc)
    x1 = sc.bodies[1].x0
    inner = (np.abs(x1[:, 0] - x1[:, 0].min()) < 1e-9) & (np.abs(x1[:, 1] - x1[:, 1].min()) < 1e-9)
    body = ids[:, 3] >= _nsdf(sc)
    got = {int(i[2]): n for i, n in zip(ids[body], qn[body, 3:]) if i[0] == 1 and inner[i[2]]}
    assert len(got) == 3
    for v, n in got.items():
        # the closest point of the big box lies on its edge (x = 0.2, y = top); direction (1, 1, 0)/sqrt 2
        assert np.allclose(n, np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0), atol=1e-9), (v, n)
        assert x1[v, 2] >= 0.05 - 1e-12


def test_a_resting_convex_body_holds_none_of_its_own_vertices():
    sc = _two_boxes((0.05, 0.3, 0.05))  # far apart: no pair within the radius
    _, ids, _ = _contacts(sc)
    assert np.sum(ids[:, 3] >= _nsdf(sc)) == 0


def test_a_fold_of_one_body_rests_on_its_own_surface():
    # the channel's upper arm sags onto its lower arm: after 20 frames its free end's lower vertices
    # hold contacts with the body's own triangles of the lower arm's top face, normal close to +y
    sc = scenes.channel(young=1e5, gap=1)
    o, ids, qn = _contacts(sc, frames=20)
    x = o.read_body(0)[0]
    x0 = sc.bodies[0].x0
    tip = (np.abs(x0[:, 0] - 0.3) < 1e-9) & (np.abs(x0[:, 1] - 0.101) < 1e-9)  # the upper arm's free end, below
    body = ids[:, 3] >= _nsdf(sc)
    own = [(i, n) for i, n in zip(ids[body], qn[body, 3:]) if tip[i[2]]]
    assert own, "the free end holds no contact with its own surface"
    for i, n in own:
        assert n[1] > 0.9, (i, n)
    # and stays above the lower arm's top face
    assert x[tip, 1].min() > 0.05


def _vertex_under_a_face(pinned_face):
    """Two single-tet bodies, no gravity, no elastic constraints: body 1's bottom face (y = 0,
    outward normal -y) above body 0's apex p = (0.25, -0.01, 0.25); thickness 0.05."""
    sc = Scene("vertex_under_face", gravity=(0.0, 0.0, 0.0), gravity_is_force=False, iterations=1, substeps=1,
               coloring="index", body_contacts=(2.0, 0.05, 0.5))
    a = np.array([[0.25, -0.01, 0.25], [0.0, -1.0, 0.0], [1.0, -1.0, 0.0], [0.0, -1.0, 1.0]])
    b = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.25, 1.0, 0.25]])
    tet = np.array([[0, 1, 2, 3]], dtype=np.uint32)
    for x0, mass in ((a, np.array([1.0, 0.0, 0.0, 0.0])), (b, np.zeros(4) if pinned_face else np.ones(4))):
        s2t, tris = _lib.boundary(4, tet)
        sc.bodies.append(Body(x0=x0, tets=tet, mass=mass, surface=s2t, triangles=tris))
    return sc


@pytest.mark.parametrize("pinned_face", [True, False])
def test_one_contact_projection_against_a_face(pinned_face):
    # p lies 0.01 below the face, inside the thickness 0.05: its contact plane is 0.05 below the
    # face (normal -y). Against a pinned face one projection puts p on that plane (the SDF contact's
    # orthogonal projection); against a free face with unit masses p takes its share 1 / (1 + sum
    # b_k^2) of the 0.04 gap, b = (0.5, 0.25, 0.25) the closest point's barycentric weights
    sc = _vertex_under_a_face(pinned_face)
    o = build_oracle(sc)
    ids, qn = o.detect_contacts()
    mine = [(i, q) for i, q in zip(ids, qn) if i[0] == 0 and i[2] == 0]
    assert len(mine) == 1
    assert np.allclose(mine[0][1][3:], [0.0, -1.0, 0.0], atol=1e-15)
    assert np.allclose(sorted(mine[0][1][:3]), [0.25, 0.25, 0.5], atol=1e-15)
    o.step(sc.dt, 1, 1)
    y = o.read_body(0)[0][0, 1]
    share = 1.0 if pinned_face else 1.0 / (1.0 + 0.5 ** 2 + 2 * 0.25 ** 2)
    alpha_tilde = sc.collision_compliance / sc.dt ** 2
    assert abs((-0.01 - y) - 0.04 * share) <= 0.04 * 2 * alpha_tilde
    assert np.allclose(o.read_body(0)[0][0, [0, 2]], [0.25, 0.25], atol=1e-15)
