"""GPU check of the per-frame surface refresh (update_visual_model + compute_normals on the device)
against a NumPy restatement of the reference's loops on the device's own positions."""
import numpy as np
import pytest

from sbs_amd import ply, scenes

pytestmark = pytest.mark.gpu


def reference_normals(prev, x, s2t, tris):
    """tetrahedral_mesh_boundary.cpp:122-145: the normals keep their previous value, every triangle
    (in order) adds (p2 - p1) x (p3 - p1) to its three vertices, then each is normalized."""
    n = prev.copy()
    p = x[s2t]
    for a, b, c in tris:
        e1, e2 = p[b] - p[a], p[c] - p[a]
        f = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]])
        n[a] += f
        n[b] += f
        n[c] += f
    for i in range(len(n)):
        z = n[i, 0] * n[i, 0] + n[i, 1] * n[i, 1] + n[i, 2] * n[i, 2]
        if z > 0:
            n[i] /= np.sqrt(z)
    return n


@pytest.mark.parametrize("graphs", [False, True])
def test_surface_refresh_matches_reference_loops(graphs):
    sc = scenes.reference_demo()
    print(f"surface refresh graphs={graphs}", flush=True)
    dev = scenes.build_device(sc)
    dev.use_graphs(graphs)
    dev.set_surface_refresh(True)
    b = sc.bodies[0]
    n = np.zeros((len(b.surface), 3))
    for _ in range(6):
        dev.step(sc.dt, sc.iterations, sc.substeps)
        x = dev.read_body(0)[0]
        pos, nrm = dev.read_surface(0)
        assert np.array_equal(pos, x[b.surface])
        n = reference_normals(n, x, b.surface, b.triangles)
        assert np.array_equal(nrm, n)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0)


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_surface_frame_dump(tmp_path, fmt):
    sc = scenes.reference_demo()
    print(f"surface dump {fmt}", flush=True)
    dev = scenes.build_device(sc)
    dev.set_surface_refresh(True)
    for _ in range(3):
        dev.step(sc.dt, sc.iterations, sc.substeps)
    path = tmp_path / "surface.ply"
    dev.write_surface_ply(0, path, fmt)
    pos, nrm = dev.read_surface(0)
    g = ply.read_ply(path, raise_errors=True)
    assert g.is_triangle_mesh()
    assert np.array_equal(g.indices.reshape(-1, 3), sc.bodies[0].triangles.astype(np.int32))
    if fmt == "ascii":  # "%f" keeps 6 decimals
        assert np.allclose(g.positions.reshape(-1, 3), pos, rtol=0, atol=5e-7 + 1e-6 * np.abs(pos).max())
        assert np.allclose(g.normals.reshape(-1, 3), nrm, rtol=0, atol=1e-6)
        expect = ply.write_ply_bytes(ply.triangle_geometry(pos, sc.bodies[0].triangles, nrm), "ascii")
        assert open(path, "rb").read() == expect
    else:
        assert np.array_equal(g.positions.reshape(-1, 3), pos.astype(np.float32))
        assert np.array_equal(g.normals.reshape(-1, 3), nrm.astype(np.float32))
