"""GPU checks of grid SDF obstacles: the device interpolation of the cubic Lagrange grid and the
device mesh-distance node values equal the oracle's bit for bit; the environment-body entry point
builds the same field as the oracle; and a soft body falling on a grid obstacle produces the
oracle's contacts and trajectory."""
import os

import numpy as np
import pytest

import oracle as O
import sbs_amd as S
from helpers import GOLDEN, build_oracle, run_pair
from sbs_amd import ply, scenes
from test_oracle_grid import GRIDS, cubic, icosphere

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def bunny():
    g = ply.read_ply(os.path.join(GOLDEN, "data", "meshes", "bunny_simplified.ply"), raise_errors=True)
    return g.positions.astype(np.float64).reshape(-1, 3), g.indices.astype(np.uint32).reshape(-1, 3)


def grid_pair(lo, hi, res, nodes):
    dev = S.Simulation(0)
    dev.add_grid_sdf(lo, hi, res, nodes)
    orc = O.Simulation()
    orc.add_grid_sdf(lo, hi, res, nodes)
    return dev, orc


@pytest.mark.parametrize("g", range(len(GRIDS)))
def test_device_interpolation_matches_oracle(g):
    lo, hi, res = GRIDS[g]
    say(f"grid interpolation {res}")
    rng = np.random.default_rng(g)
    n = S.grid_node_count(res)
    P = rng.uniform(np.asarray(lo) - 0.1, np.asarray(hi) + 0.1, (5000, 3))
    for nodes in (rng.normal(size=n), cubic(S.grid_node_positions(lo, hi, res))[0]):
        dev, orc = grid_pair(lo, hi, res, nodes)
        sd, gr = dev.evaluate_sdf(0, P)
        so, go = orc.evaluate_sdf(0, P)
        assert np.array_equal(sd, so) and np.array_equal(gr, go)


@pytest.mark.parametrize("mesh", ["icosphere", "bunny"])
def test_device_mesh_grid_matches_oracle(mesh):
    V, F = icosphere(0.5, 3) if mesh == "icosphere" else bunny()
    lo, hi = V.min(0) - 0.1, V.max(0) + 0.1
    res = [10, 12, 9]
    say(f"mesh grid {mesh}: {len(F)} triangles, {S.grid_node_count(res)} nodes")
    nd = S.grid_from_mesh(V, F, lo, hi, res)
    no = O.grid_from_mesh(V, F, lo, hi, res)
    assert np.array_equal(nd, no)


def test_mesh_obstacle_entry_point_matches_oracle():
    g = ply.read_ply(os.path.join(GOLDEN, "data", "meshes", "xz_plane.ply"), raise_errors=True)
    V, F = g.positions.astype(np.float64).reshape(-1, 3) * 2.0, g.indices.astype(np.uint32).reshape(-1, 3)
    lo, hi, res = [-2.0, -0.5, -2.0], [2.0, 0.5, 2.0], [10, 10, 10]
    say("mesh obstacle entry point")
    dev = S.Simulation(0)
    dev.add_mesh_obstacle(V, F, lo, hi, res)
    orc = O.Simulation()
    orc.add_mesh_obstacle(V, F, lo, hi, res)
    P = np.random.default_rng(4).uniform([-2.5, -1.0, -2.5], [2.5, 1.0, 2.5], (4000, 3))
    sd, gr = dev.evaluate_sdf(0, P)
    so, go = orc.evaluate_sdf(0, P)
    assert np.array_equal(sd, so) and np.array_equal(gr, go)
    inside = sd < 1e300
    # a horizontal quad: the distance field is the height above it inside its footprint, a linear
    # function the cubic cells reproduce wherever all their nodes lie over the quad (the cells are
    # about 0.4 wide and the quad's edges at +-2)
    foot = inside & (np.abs(P[:, 0]) < 1.5) & (np.abs(P[:, 2]) < 1.5)
    assert foot.sum() > 500
    assert np.allclose(sd[foot], P[foot, 1], rtol=0, atol=1e-12)


def bunny_drop_scene():
    V, F = bunny()
    lo, hi, res = V.min(0) - 0.1, V.max(0) + 0.1, [16, 16, 16]
    nodes = S.grid_from_mesh(V, F, lo, hi, res)
    sc = scenes.cantilever(8, 3, 3, 0.08, "neohookean", 2e5, 0.4, 1000.0, ground_y=-0.6, iterations=8,
                           substeps=2, cluster=True, coloring="index", name="bunny_drop")
    body = sc.bodies[0]
    body.x0 = body.x0 + np.array([-0.3, 0.62, -0.1])
    body.mass = np.full(len(body.x0), body.mass.max())  # nothing pinned: the body falls
    sc.grids.append((lo, hi, res, nodes, lo, hi))
    return sc


def test_body_falling_on_grid_obstacle_matches_oracle():
    sc = bunny_drop_scene()
    say("bunny drop")
    errs, dev, _ = run_pair(sc, 20, check_frames=(10, 15, 20))  # the body lands on the bunny at frame 13-20
    say(f"  max error / bbox diagonal by frame {errs}")
    assert max(errs.values()) <= 1e-4
    # contacts on one state: the device's positions given to the oracle
    xd, vd = dev.read_body(0)
    orc = build_oracle(sc)
    orc.set_body_state(0, xd, vd)
    ids_d, qn_d = dev.read_contacts()
    ids_o, qn_o = orc.detect_contacts()
    o = np.lexsort((ids_o[:, 3], ids_o[:, 1], ids_o[:, 0]))
    ids_o, qn_o = ids_o[o], qn_o[o]
    grid_contacts = int(np.sum(ids_d[:, 3] == 1))
    say(f"  {len(ids_d)} contacts, {grid_contacts} on the grid obstacle")
    assert grid_contacts > 5
    assert np.array_equal(ids_d, ids_o)
    assert np.array_equal(qn_d[:, :3], qn_o[:, :3])
    assert np.allclose(qn_d[:, 3:], qn_o[:, 3:], rtol=0, atol=1e-7)  # normals are stored in FP32
