"""Shared test helpers: build the same scene on the device and in the FP64 oracle, compare."""
import os

import numpy as np

import oracle as O
from sbs_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_ply_tets(path):
    """ASCII PLY with `vertex` (x y z) and `tet` (list uchar int) elements, as the reference's
    fixtures in data/meshes are written."""
    lines = open(path).read().split("\n")
    end = lines.index("end_header")
    nv = int(next(l for l in lines if l.startswith("element vertex")).split()[-1])
    nt = int(next(l for l in lines if l.startswith("element tet")).split()[-1])
    pos = np.array([[float(x) for x in l.split()] for l in lines[end + 1:end + 1 + nv]])
    tets = np.array([[int(x) for x in l.split()][1:] for l in lines[end + 1 + nv:end + 1 + nv + nt]], dtype=np.uint32)
    return pos, tets


def build_oracle(scene, colored=True, recolor_passes=2):
    sim = O.Simulation()
    scene.apply(sim)
    sim.set_order(colored, recolor_passes)
    scene.set_initial_state(sim)
    return sim


def run_pair(scene, frames, check_frames=(), recolor_passes=2, device_sim=None):
    """Steps the device and the colored-order oracle side by side; returns {frame: max bbox-relative
    position error over all bodies} for the requested frames (1-based)."""
    dev = device_sim if device_sim is not None else scenes.build_device(scene, recolor_passes=recolor_passes)
    orc = build_oracle(scene, True, recolor_passes)
    diag = scene.bbox_diagonal()
    errs = {}
    for f in range(1, frames + 1):
        dev.step(scene.dt, scene.iterations, scene.substeps)
        orc.step(scene.dt, scene.iterations, scene.substeps)
        if f in check_frames:
            e = 0.0
            for b in range(len(scene.bodies)):
                xd, _ = dev.read_body(b)
                xo, _ = orc.read_body(b)
                assert np.all(np.isfinite(xd)), f"non-finite device positions at frame {f}"
                e = max(e, float(np.abs(xd - xo).max()) / diag)
            errs[f] = e
    return errs, dev, orc


def tets_conflict_free(tets, colors):
    """True when no two constraints of the same color share a vertex."""
    tets = np.asarray(tets)
    seen = {}
    for t, c in enumerate(colors):
        for v in tets[t]:
            key = (int(c), int(v))
            if key in seen:
                return False
            seen[key] = t
    return True
