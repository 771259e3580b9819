"""GPU check of a scene loaded from the reference's scene file (data/scenes/scene1.json): the
yellow bar falls on the floor mesh (a grid SDF built on the device) and follows the oracle; the
contacts on one shared state are the oracle's."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, build_oracle, run_pair
from sbs_amd import scene_file

pytestmark = pytest.mark.gpu

TOL = 1e-4  # max vertex error / bounding-box diagonal


def test_scene1_matches_oracle():
    sc, _ = scene_file.to_scene(os.path.join(GOLDEN, "data", "scenes", "scene1.json"))
    print("scene1.json", flush=True)
    errs, dev, _ = run_pair(sc, 90, check_frames=(30, 60, 90))
    print(f"  max error / bbox diagonal by frame {errs}", flush=True)
    assert max(errs.values()) <= TOL
    xd, vd = dev.read_body(0)
    assert xd[:, 1].min() > -0.05  # resting on the floor
    orc = build_oracle(sc)
    orc.write_body(0, xd, vd)
    ids_d, qn_d = dev.read_contacts()
    ids_o, qn_o = orc.detect_contacts()
    o = np.lexsort((ids_o[:, 3], ids_o[:, 1], ids_o[:, 0]))
    assert len(ids_d) > 20
    assert np.array_equal(ids_d, ids_o[o]) and np.array_equal(qn_d[:, :3], qn_o[o, :3])
