"""GPU checks of the device boundary extraction: the same surface vertices (in first-appearance
order) and triangles (in the mesh's triangle order, first-face vertex order) as the host
extraction and the FP64 oracle on the reference's PLY fixtures, lattice meshes, an ellipsoid and a
random tet soup with faces of every incidence count; on the 6M-tet beam the closed-form counts."""
import os
import time

import numpy as np
import pytest

import oracle as O
import sbs_amd as S
from helpers import GOLDEN, read_ply_tets

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def check_same(nv, tets):
    t0 = time.perf_counter()
    ds, dt = S.extract_boundary(nv, tets)
    t1 = time.perf_counter()
    hs, ht = S.boundary(nv, tets)
    t2 = time.perf_counter()
    say(f"  device {1e3 * (t1 - t0):.1f} ms, host {1e3 * (t2 - t1):.1f} ms, {len(ds)} vertices, {len(dt)} triangles")
    assert np.array_equal(ds, hs) and np.array_equal(dt, ht)
    return ds, dt


@pytest.mark.parametrize("fixture", ["tetrahedron.ply", "2tets.ply", "3tets.ply", "cube_tet.ply", "bar_tet.ply",
                                     "tet_bar_5x2x2.ply"])
def test_fixtures_match_host_and_oracle(fixture):
    say(f"boundary {fixture}")
    pos, tets = read_ply_tets(os.path.join(GOLDEN, fixture))
    ds, dt = check_same(len(pos), tets)
    os_, ot, _ = O.boundary(len(pos), tets)
    assert np.array_equal(ds, os_) and np.array_equal(dt, ot)


@pytest.mark.parametrize("dims", [(10, 3, 3), (220, 28, 27)])
def test_lattice_beams_match_oracle(dims):
    say(f"boundary box6 {dims}")
    pos, tets = S.box6(*dims, 0.1)
    ds, dt = check_same(len(pos), tets)
    os_, ot, _ = O.boundary(len(pos), tets)
    assert np.array_equal(ds, os_) and np.array_equal(dt, ot)


def test_ellipsoid_and_shuffled_bar():
    say("boundary ellipsoid")
    pos, tets = S.ellipsoid6(0.6, 0.4, 0.5, 0.05)
    check_same(len(pos), tets)
    pos, tets = S.bar5(12, 4, 4)
    rng = np.random.default_rng(7)
    check_same(len(pos), tets[rng.permutation(len(tets))])


def test_tet_soup_with_shared_and_repeated_faces():
    # a small vertex set makes faces shared by 1, 2, 3 and more tets; repeated tets too
    say("boundary tet soup")
    rng = np.random.default_rng(11)
    nv = 24
    tets = np.array([rng.choice(nv, 4, replace=False) for _ in range(600)], dtype=np.uint32)
    tets = np.concatenate([tets, tets[rng.integers(0, len(tets), 50)]])
    _, counts = np.unique(np.sort(np.concatenate([tets[:, [0, 1, 3]], tets[:, [1, 2, 3]], tets[:, [2, 0, 3]],
                                                  tets[:, [0, 2, 1]]]), axis=1), axis=0, return_counts=True)
    assert {1, 2, 3}.issubset(set(counts.tolist()))
    ds, dt = check_same(nv, tets)
    os_, ot, _ = O.boundary(nv, tets)
    assert np.array_equal(ds, os_) and np.array_equal(dt, ot)


def test_six_million_tet_beam_counts():
    nx, ny, nz = 400, 50, 50
    say("boundary box6 400x50x50")
    pos, tets = S.box6(nx, ny, nz, 0.1)
    t0 = time.perf_counter()
    ds, dt = S.extract_boundary(len(pos), tets)
    say(f"  device {1e3 * (time.perf_counter() - t0):.1f} ms")
    # a box lattice: every vertex on the box's faces and 2 triangles per boundary quad
    assert len(ds) == (nx + 1) * (ny + 1) * (nz + 1) - (nx - 1) * (ny - 1) * (nz - 1) == 85_002
    assert len(dt) == 4 * (nx * ny + ny * nz + nx * nz) == 170_000
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    p = pos[ds]
    assert np.all(np.any(np.isclose(p, lo) | np.isclose(p, hi), axis=1))
    assert len(np.unique(ds)) == len(ds)
    # every surface vertex appears in a triangle no later than any vertex numbered after it
    first = np.full(len(ds), len(dt) * 3)
    np.minimum.at(first, dt.reshape(-1), np.arange(dt.size))
    assert np.all(np.diff(first) > 0)


def test_empty_mesh():
    s2t, tris = S.extract_boundary(4, np.zeros((0, 4), dtype=np.uint32))
    assert len(s2t) == 0 and len(tris) == 0
