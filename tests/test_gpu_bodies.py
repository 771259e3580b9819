"""GPU checks of the contacts between bodies and of a body with itself: the broad phase's candidate
pairs against a NumPy restatement of the triangle hash and against brute force over every
(vertex, triangle) pair, the contacts against the FP64 oracle's from the same positions, and a
few-body tower and a self-folding channel stepped side by side with the oracle."""
import os

import numpy as np
import pytest

from helpers import build_oracle
from sbs_amd import scenes
from test_gpu_hash import hash_cell, hash_slot

pytestmark = pytest.mark.gpu

TOL = 1e-4  # max vertex error / bounding-box diagonal


def say(msg):
    print(msg, flush=True)


def _surface_positions(sc, xs):
    return np.concatenate([x[np.asarray(b.surface)] for b, x in zip(sc.bodies, xs)])


def _restated_candidates(P, tris, cell, bits):
    """(s, t) for every surface vertex s and triangle t not having s as a corner that share a slot of
    the 27 cells around s's cell (the triangle's box cells hashed as the device does), or whose box
    covers more than 8 cells (tested by every vertex)."""
    mask = (1 << bits) - 1
    inv = 1.0 / cell
    slots, large = {}, []
    for t, (a, b, c, _) in enumerate(tris):
        corners = P[[a, b, c]]
        lo, hi = hash_cell(corners.min(0), inv), hash_cell(corners.max(0), inv)
        if np.prod(hi - lo + 1) > 8:
            large.append(t)
            continue
        for k in range(lo[2], hi[2] + 1):
            for j in range(lo[1], hi[1] + 1):
                for i in range(lo[0], hi[0] + 1):
                    slots.setdefault(int(hash_slot(i, j, k, mask)), set()).add(t)
    out = set()
    for s, p in enumerate(P):
        ci = hash_cell(p, inv)
        for dk in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    for t in slots.get(int(hash_slot(ci[0] + di, ci[1] + dj, ci[2] + dk, mask)), ()):
                        if s not in tris[t][:3]:
                            out.add((s, t))
        for t in large:
            if s not in tris[t][:3]:
                out.add((s, t))
    return np.array(sorted(out), dtype=np.uint32).reshape(-1, 2)


def _scene(name):
    if name == "tower":
        return scenes.tower()
    if name == "tower_spheres":  # 4 spheres pressing 1.5 cm into the slab's sides and the top box: 5 SDFs, 7 slots
        sc = scenes.tower(name="tower_spheres")
        sc.spheres += [((0.255, 0.05, 0.0), 0.07), ((-0.255, 0.05, 0.05), 0.07), ((0.0, 0.05, -0.255), 0.07),
                       ((-0.08, 0.43, -0.08), 0.075)]
        return sc
    if name == "tower_far":  # 1 km from the origin: the scan's cell-relative FP32 test at large coordinates
        sc = scenes.tower(name="tower_far")
        for b in sc.bodies:
            b.x0 = b.x0 + np.array([1000.25, 0.0, -700.5])
        return sc
    if name == "tower_big_block":  # a body whose triangles cover more than 8 cells of the hash
        return scenes.tower(big_block=True, name="tower_big_block")
    return scenes.channel(young=1e5, gap=1)


@pytest.mark.parametrize("name", ["tower", "tower_big_block", "channel"])
def test_candidate_pairs_are_bit_exact_and_cover_brute_force(name):
    sc = _scene(name)
    say(f"candidate pairs: {sc.name}")
    dev = scenes.build_device(sc)
    cell, h, r, bits, tris = dev.body_contact_setup()
    assert cell > 0 and 0 < h < r < cell
    got = dev.body_contact_candidates()
    P = _surface_positions(sc, [b.x0 for b in sc.bodies])
    want = _restated_candidates(P, tris, cell, bits)
    say(f"{len(got)} candidate pairs")
    assert np.array_equal(got, want)
    # brute force: every non-incident pair whose vertex lies within r of the triangle's box is a
    # candidate (a triangle closer than r to a vertex is one of them)
    lo = np.stack([P[tris[:, k]] for k in range(3)], 1).min(1) - r
    hi = np.stack([P[tris[:, k]] for k in range(3)], 1).max(1) + r
    near = np.all((P[:, None, :] >= lo[None]) & (P[:, None, :] <= hi[None]), axis=2)
    s_idx = np.arange(len(P))[:, None]
    near &= (tris[None, :, 0] != s_idx) & (tris[None, :, 1] != s_idx) & (tris[None, :, 2] != s_idx)
    need = set(map(tuple, np.argwhere(near)))
    have = set(map(tuple, got.astype(np.int64)))
    assert len(need) > 0 and need <= have


def _sorted(ids, qn):
    o = np.lexsort((ids[:, 3], ids[:, 1], ids[:, 0]))
    return ids[o], qn[o]


@pytest.mark.parametrize("name,frames", [("tower", 0), ("tower", 20), ("tower_far", 20), ("tower_spheres", 0),
                                         ("tower_spheres", 20), ("tower_big_block", 10), ("channel", 20)])
def test_contacts_match_the_oracle_from_the_same_positions(name, frames):
    sc = _scene(name)
    say(f"contacts vs oracle: {sc.name} after {frames} frames")
    dev = scenes.build_device(sc)
    for _ in range(frames):
        dev.step(sc.dt, sc.iterations, sc.substeps)
    orc = build_oracle(sc)
    for b in range(len(sc.bodies)):
        x, v = dev.read_body(b)
        dev.write_body(b, x, v)
        orc.set_body_state(b, x, v)
    ids_d, qn_d = _sorted(*dev.read_contacts())
    ids_o, qn_o = _sorted(*orc.detect_contacts())
    nsdf = len(sc.planes) + len(sc.spheres)
    say(f"{len(ids_d)} contacts, {int(np.sum(ids_d[:, 3] >= nsdf))} with triangles, "
        f"{int(np.sum((ids_d[:, 3] >= len(sc.planes)) & (ids_d[:, 3] < nsdf)))} with spheres")
    assert np.sum(ids_d[:, 3] >= nsdf) > 20
    assert dev.contact_overflow() == 0
    assert np.array_equal(ids_d, ids_o)
    assert np.array_equal(qn_d[:, :3], qn_o[:, :3])  # SDF contact points, triangle barycentrics
    assert np.allclose(qn_d[:, 3:], qn_o[:, 3:], rtol=0, atol=1e-7)  # normals are stored in FP32


def _side_by_side(sc, frames, check):
    dev = scenes.build_device(sc)
    orcs = []
    for rounded in (False, True):
        o = build_oracle(sc)
        o.set_threads(max(1, min(16, os.cpu_count() or 1)))
        if rounded:  # the oracle's own sensitivity to rounding-sized perturbations of the start
            for bi, body in enumerate(sc.bodies):
                o.set_body_state(bi, body.x0.astype(np.float32).astype(np.float64), np.zeros_like(body.x0))
        orcs.append(o)
    diag = sc.bbox_diagonal()
    nb = len(sc.bodies)
    for f in range(1, frames + 1):
        dev.step(sc.dt, sc.iterations, sc.substeps)
        for o in orcs:
            o.step(sc.dt, sc.iterations, sc.substeps)
        if f in check:
            xd = [dev.read_body(b)[0] for b in range(nb)]
            assert all(np.all(np.isfinite(x)) for x in xd), f"non-finite positions at frame {f}"
            err = max(float(np.abs(xd[b] - orcs[0].read_body(b)[0]).max()) for b in range(nb)) / diag
            spread = max(float(np.abs(orcs[1].read_body(b)[0] - orcs[0].read_body(b)[0]).max())
                         for b in range(nb)) / diag
            say(f"frame {f} error {err:.3e} (oracle spread {spread:.3e})")
            assert err <= max(TOL, 3.0 * spread), (f, err, spread)
    assert dev.contact_overflow() == 0
    return dev


def test_tower_of_bodies_matches_the_oracle_over_60_frames():
    sc = scenes.tower()
    say(f"{sc.name}: {len(sc.bodies)} bodies, {sc.num_tets} tets")
    dev = _side_by_side(sc, 60, (1, 10, 20, 30, 40, 50, 60))
    # every body rests on its support: none below the ground by more than the thickness, the boxes
    # stacked in order
    ymin = [float(dev.read_body(b)[0][:, 1].min()) for b in range(len(sc.bodies))]
    ymax = [float(dev.read_body(b)[0][:, 1].max()) for b in range(len(sc.bodies))]
    say(f"lowest points {np.round(ymin, 4)}")
    assert min(ymin) > -0.01
    assert ymin[1] > ymax[0] - 0.01 and ymin[2] > ymax[1] - 0.01 and ymin[3] > ymax[0] - 0.01


def test_self_folding_channel_matches_the_oracle_over_30_frames():
    sc = scenes.channel(young=1e5, gap=1)
    say(f"{sc.name}: {sc.num_tets} tets")
    dev = _side_by_side(sc, 30, (1, 10, 20, 30))
    ids, _ = dev.read_contacts()
    tris = dev.body_contact_setup()[4]
    own = ids[ids[:, 3] >= len(sc.planes)]
    assert len(own) > 0 and np.all(tris[own[:, 3] - len(sc.planes), 3] == 0)



def test_contacts_match_the_oracle_on_folded_bodies():
    """32 stiff bodies of 2k-20k tets (Young's modulus up to 1e6) after 15 frames of 8 iterations
    from zero multipliers: their surfaces have folded onto themselves, the densest case of the
    own-surface search (about a fifth of the surface vertices have more than 512 hashed entries in
    their cells, up to ~2,500, and take the near-first order with its held-back queue and second
    scan); the contacts from the same positions equal the oracle's."""
    sc = scenes.batch(32, tet_range=(2e3, 2e4), name="batch_32_folded")
    sc.body_contacts = (0.0, 0.0, 0.0)
    say(f"{sc.name}: {len(sc.bodies)} bodies, {sc.num_tets} tets")
    dev = scenes.build_device(sc)
    for _ in range(15):
        dev.step(sc.dt, sc.iterations, sc.substeps)
    orc = build_oracle(sc)
    for b in range(len(sc.bodies)):
        x, v = dev.read_body(b)
        assert np.all(np.isfinite(x))
        dev.write_body(b, x, v)
        orc.set_body_state(b, x, v)
    ids_d, qn_d = _sorted(*dev.read_contacts())
    ids_o, qn_o = _sorted(*orc.detect_contacts())
    nsdf = len(sc.planes)
    say(f"{len(ids_d)} contacts, {int(np.sum(ids_d[:, 3] >= nsdf))} with triangles")
    assert np.sum(ids_d[:, 3] >= nsdf) > 1000
    assert dev.contact_overflow() == 0
    assert np.array_equal(ids_d, ids_o)
    assert np.array_equal(qn_d[:, :3], qn_o[:, :3])
    assert np.allclose(qn_d[:, 3:], qn_o[:, 3:], rtol=0, atol=1e-7)
