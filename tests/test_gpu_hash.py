"""GPU checks of the contact broad phase: the device prefix sum and radix sort against NumPy, the
obstacle hash table against a NumPy restatement of its cells and slots, and the hashed contacts
against the brute-force device path and the FP64 oracle (bit for bit)."""
import numpy as np
import pytest

import sbs_amd as S
from helpers import build_oracle
from sbs_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("n", [0, 1, 5, 1023, 1024, 1025, 100_000, 3_000_000])
def test_exclusive_scan_matches_numpy(n):
    say(f"scan n={n}")
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1000, n).astype(np.uint32)
    out = S.exclusive_scan(a)
    ref = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)]).astype(np.uint32)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("n,bits", [(1, 8), (100, 8), (1025, 13), (300_000, 24), (300_000, 32), (2_000_000, 21)])
def test_radix_sort_is_stable_and_matches_numpy(n, bits):
    say(f"sort n={n} bits={bits}")
    rng = np.random.default_rng(bits * 7 + n)
    keys = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    if bits < 32:
        keys &= np.uint32((1 << bits) - 1)
        keys[::3] = keys[0]  # many equal keys: stability matters
    vals = np.arange(n, dtype=np.uint32)
    ok, ov = S.sort_pairs(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ok, keys[order]) and np.array_equal(ov, vals[order])


def hash_slot(i, j, k, mask):
    i, j, k = (np.asarray(v, dtype=np.int64) & 0xFFFFFFFF for v in (i, j, k))
    return ((i * 73856093) & 0xFFFFFFFF ^ (j * 19349663) & 0xFFFFFFFF ^ (k * 83492791) & 0xFFFFFFFF) & mask


def hash_cell(x, inv_h):
    return np.clip(np.floor(np.asarray(x, dtype=np.float64) * inv_h), -2.0 ** 30, 2.0 ** 30).astype(np.int64)


def test_obstacle_hash_table_matches_restatement():
    sc = scenes.obstacle_field(3000)
    say("obstacle hash table")
    dev = scenes.build_device(sc)
    h, bits, start, end, entries, glob = dev.read_contact_hash()
    assert h > 0 and bits >= 10
    mask = (1 << bits) - 1
    inv_h = 1.0 / h
    # sphere k is SDF 1 + k (the plane comes first), its volume the box c -+ r in FP64
    slots = {}
    for k, (c, r) in enumerate(sc.spheres):
        c = np.asarray(c)
        lo, hi = c - r, c + r
        a, b = hash_cell(lo, inv_h), hash_cell(hi, inv_h)
        for i in range(a[0], b[0] + 1):
            for j in range(a[1], b[1] + 1):
                for m in range(a[2], b[2] + 1):
                    slots.setdefault(int(hash_slot(i, j, m, mask)), set()).add(1 + k)
    assert list(glob) == [0]  # the unbounded ground plane is tested by every vertex
    assert len(entries) == sum(len(v) for v in slots.values())
    for sl in range(1 << bits):
        got = list(entries[start[sl]:end[sl]])
        assert got == sorted(slots.get(sl, ())), sl


def test_hashed_contacts_match_brute_force_and_oracle():
    sc = scenes.obstacle_field(3000)
    say("hashed contacts")
    hashed = scenes.build_device(sc)
    sc_bf = scenes.obstacle_field(3000, contact_hash=("off", 0.0))
    brute = scenes.build_device(sc_bf)
    assert hashed.read_contact_hash()[0] > 0 and brute.read_contact_hash()[0] == 0
    ids_h, qn_h = hashed.read_contacts()
    ids_b, qn_b = brute.read_contacts()
    orc = build_oracle(sc)
    ids_o, qn_o = orc.detect_contacts()
    # the oracle generates contacts SDF by SDF (the reference's order); the device reports them by
    # (body, surface vertex, SDF), the order they are projected in per vertex
    o = np.lexsort((ids_o[:, 3], ids_o[:, 1], ids_o[:, 0]))
    ids_o, qn_o = ids_o[o], qn_o[o]
    say(f"{len(ids_h)} contacts")
    assert len(ids_h) > 300
    assert hashed.contact_overflow() == 0 and brute.contact_overflow() == 0
    assert np.array_equal(ids_h, ids_b) and np.array_equal(qn_h, qn_b)
    assert np.array_equal(ids_h, ids_o) and np.array_equal(qn_h[:, :3], qn_o[:, :3])
    assert np.allclose(qn_h[:, 3:], qn_o[:, 3:], rtol=0, atol=1e-7)  # normals are stored in FP32


def test_hashed_and_brute_force_steps_are_bit_identical():
    sc = scenes.obstacle_field(3000)
    say("hashed vs brute-force steps")
    a = scenes.build_device(sc)
    b = scenes.build_device(scenes.obstacle_field(3000, contact_hash=("off", 0.0)))
    for _ in range(5):
        a.step(sc.dt, sc.iterations, sc.substeps)
        b.step(sc.dt, sc.iterations, sc.substeps)
    assert np.array_equal(a.read_body(0)[0], b.read_body(0)[0])


def test_contact_overflow_is_an_error():
    # a beam corner inside 10 concentric spheres: 10 contacts at one surface vertex, 2 more than the
    # slots (the beam rests on the ground plane, sd = 0, no contact)
    sc = scenes.cantilever(4, 2, 2, 0.1, "neohookean", 5e5, 0.45, 1000.0, ground_y=0.0, iterations=2, substeps=1,
                           name="overflow_beam")
    for k in range(10):
        sc.spheres.append(((0.4, 0.2, 0.2), 0.05 + 0.01 * k))
    say("contact overflow")
    dev = scenes.build_device(sc)
    dev.step(sc.dt, sc.iterations, sc.substeps)
    with pytest.raises(S.SbsError) as e:
        dev.synchronize()
    assert e.value.code == 9  # SBS_ERR_CONTACT_OVERFLOW
    assert dev.contact_overflow() > 0
    with pytest.raises(S.SbsError):  # sticky: the next frame is refused
        dev.step(sc.dt, sc.iterations, sc.substeps)
