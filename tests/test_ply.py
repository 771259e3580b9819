"""PLY reading and writing (src/io/ply.cpp) against the reference's own reader and writer, compiled
from its sources into oracle/_ref (tests that need it skip where it is not built), and against the
reference's fixture files: ASCII files byte for byte, binary triangle meshes byte for byte, the
arrays every reader returns, and the number parsing of std::stof / std::stoi."""
import glob
import os
import time

import numpy as np
import pytest

import ref_ply
import sbs_amd as S
from helpers import GOLDEN, read_ply_tets
from sbs_amd import ply

needs_ref = pytest.mark.skipif(not ref_ply.available(), reason="oracle/_ref/libref_ply.so is not built")

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "*.ply")))
ARRAYS = ("positions", "indices", "normals", "uvs", "colors")


def same_arrays(g, r, check_type=True):
    for name in ARRAYS:
        a, b = getattr(g, name), r[name]
        assert a.shape == b.shape, name
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name  # bit for bit, NaNs included
    if check_type:
        assert g.geometry_type == r["geometry_type"]


def sample_geometry(rng, nv, ne, kind, extras=True):
    pos = rng.normal(0, 3, nv * 3).astype(np.float32)
    # values that stress "%f": rounding ties at the 7th decimal, large and tiny magnitudes, signed zero
    special = np.array([0.0078125, -0.0078125, 1 / 3, 2.5e-7, 5e-7, 1e30, -3.4e38, 1e-30, -0.0, 0.5, 123456.789],
                       np.float32)
    pos[:len(special)] = special[:min(len(special), len(pos))]
    g = ply.Geometry(positions=pos, indices=rng.integers(0, max(nv, 1), ne * (3 if kind == ply.TRIANGLE else 4))
                     .astype(np.int32), geometry_type=kind)
    if extras:
        g.normals = rng.normal(0, 1, nv * 3).astype(np.float32)
        g.colors = rng.integers(0, 256, nv * 3).astype(np.uint8)
        g.uvs = rng.random(nv * 2).astype(np.float32)
    return g


def ref_write(g, fmt):
    return ref_ply.write(g.positions, g.indices, g.normals, g.uvs, g.colors, g.geometry_type, fmt)


@needs_ref
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_reference_fixtures_read_like_the_reference(path):
    data = open(path, "rb").read()
    g = ply.read_ply(path)
    st, r = ref_ply.read(data)
    assert st == 0 and g is not None
    same_arrays(g, r)
    pos, tets = read_ply_tets(path)
    assert g.is_tetrahedral_mesh()
    assert np.array_equal(g.positions, pos.astype(np.float32).reshape(-1))
    assert np.array_equal(g.indices, tets.astype(np.int32).reshape(-1))


@needs_ref
@pytest.mark.parametrize("kind", [ply.TRIANGLE, ply.TETRAHEDRON])
@pytest.mark.parametrize("extras", [False, True])
def test_ascii_writes_are_byte_identical(kind, extras):
    g = sample_geometry(np.random.default_rng(kind * 2 + extras), 2000, 3000, kind, extras)
    ours = ply.write_ply_bytes(g, "ascii")
    assert ours == ref_write(g, 0)
    # and both readers read the file back alike
    st, r = ref_ply.read(ours)
    assert st == 0
    same_arrays(ply.read_ply_bytes(ours), r)


@needs_ref
@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian"])
def test_binary_triangle_writes_are_byte_identical_and_round_trip(fmt):
    g = sample_geometry(np.random.default_rng(3), 1500, 2500, ply.TRIANGLE)
    ours = ply.write_ply_bytes(g, fmt)
    assert ours == ref_write(g, ply.FORMATS[fmt])
    back = ply.read_ply_bytes(ours)
    st, r = ref_ply.read(ours)
    assert st == 0
    same_arrays(back, r)
    for name in ARRAYS:
        assert np.array_equal(getattr(back, name).view(np.uint8), getattr(g, name).view(np.uint8)), name


@needs_ref
@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian"])
def test_binary_tets_keep_all_four_indices(fmt):
    # the reference writes a list size of 4 followed by 3 indices (ply.cpp:245) and reads tets at
    # 3 per tet (ply.cpp:974); this library writes and reads all 4
    g = sample_geometry(np.random.default_rng(5), 300, 700, ply.TETRAHEDRON)
    ours = ply.write_ply_bytes(g, fmt)
    ref = ref_write(g, ply.FORMATS[fmt])
    assert len(ours) == len(ref) + 4 * 700
    head = ours[:ours.index(b"end_header\n") + 11]
    assert ref.startswith(head)
    back = ply.read_ply_bytes(ours)
    for name in ARRAYS:
        assert np.array_equal(getattr(back, name).view(np.uint8), getattr(g, name).view(np.uint8)), name
    assert back.is_tetrahedral_mesh()


def header(fmt, *lines):
    return ("ply\nformat %s 1.0\n" % fmt + "".join(l + "\n" for l in lines) + "end_header\n").encode()


CASES = {
    "comments_and_unknown_lines": header("ascii", "comment made by hand", "obj_info x", "element vertex 3",
                                         "property float x", "property float y", "property float z",
                                         "element face 1", "property list uchar int indices")
    + b"0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n",
    "extra_vertex_property_and_element": header("ascii", "element vertex 3", "property float x", "property float y",
                                                "property float confidence", "property float z", "element edge 2",
                                                "property int a", "property int b", "element face 2",
                                                "property uchar flags", "property list uchar int indices")
    + b"1 2 9 3\n4 5 9 6\n7 8 9 9\n0 1\n1 2\n3 0 1 2\n4 0 1 2 0\n",
    "crlf_and_signs": header("ascii", "element vertex 2", "property float x", "property float y", "property float z",
                             "property uchar r", "property uchar g", "property uchar b")
    .replace(b"\n", b"\r\n") + b"+1.5 -2e3 .25 300 -1 +7\r\n1e-3 5. -0 0 255 12\r\n",
    "unknown_format_is_ascii": header("binary_middle_endian", "element vertex 1", "property float x",
                                      "property float y", "property float z") + b"1 2 3\n",
    "tet_lists_of_other_sizes": header("ascii", "element vertex 4", "property double x", "property double y",
                                       "property double z", "element tet 3", "property list uchar int indices")
    + b"0 0 0\n1 0 0\n0 1 0\n0 0 1\n4 0 1 2 3\n3 0 1 2\n4 3 2 1 0\n",
    "prefix_numbers": header("ascii", "element vertex 2", "property float x", "property float y", "property float z",
                             "element face 1", "property list uchar int indices")
    + b"1.5abc 0x10 inf\n-nan 2e 3.e1\n3.0 0x 1z 2q\n",
    "many_digits": header("ascii", "element vertex 2", "property float x", "property float y", "property float z")
    + b"0.1000000000000000055511151231257827 123456789012345678901234567890 1.00000005960464477539062500001\n"
      b"0.000000000000000000000000000000000000011754943508222875 3.4028234663852886e38 -16777217\n",
    "uv_and_normals": header("ascii", "element vertex 2", "property float x", "property float y", "property float z",
                             "property float nx", "property float ny", "property float nz", "property float u",
                             "property float v") + b"1 2 3 0 0 1 0.5 0.25\n4 5 6 1 0 0 0.75 1\n",
    "no_mesh_element": header("ascii", "element vertex 1", "property float x", "property float y",
                              "property float z") + b"1 2 3\n",
}

BAD = {
    "not_ply": b"plx\nformat ascii 1.0\nend_header\n",
    "bad_property_type": header("ascii", "element vertex 1", "property float32 x"),
    "element_without_count": header("ascii", "element vertex"),
    "property_before_element": header("ascii", "property float x"),
    "bad_float": header("ascii", "element vertex 1", "property float x", "property float y", "property float z")
    + b"1 abc 3\n",
    "float_overflow": header("ascii", "element vertex 1", "property float x", "property float y", "property float z")
    + b"1 1e39 3\n",
    "float_underflow": header("ascii", "element vertex 1", "property float x", "property float y",
                              "property float z") + b"1 1e-46 3\n",
    "int_overflow": header("ascii", "element vertex 3", "property float x", "property float y", "property float z",
                           "element face 1", "property list uchar int indices")
    + b"0 0 0\n1 0 0\n0 1 0\n3 0 1 4294967296\n",
}


def binary_cases():
    out = {}
    for fmt, end in (("binary_little_endian", "<"), ("binary_big_endian", ">")):
        h = header(fmt, "element vertex 3", "property double x", "property double y", "property double z",
                   "property uchar r", "property uchar g", "property uchar b", "element face 2",
                   "property list uchar uint indices")
        body = b""
        for i in range(3):
            body += np.array([i + 0.1, -i, 2.5 * i], end + "f8").tobytes() + bytes([i, 100 + i, 250])
        body += bytes([3]) + np.array([0, 1, 2], end + "u4").tobytes()
        body += bytes([3]) + np.array([2, 1, 0], end + "u4").tobytes()
        out["double_vertices_" + fmt] = h + body
        h = header(fmt, "element vertex 2", "property float x", "property float y", "property float z",
                   "element face 1", "property list uchar short indices")
        body = np.array([1, 2, 3, 4, 5, 6], end + "f4").tobytes() + bytes([3]) + np.array([0, 1, 1], end + "i2").tobytes()
        out["short_indices_" + fmt] = h + body
    return out


@needs_ref
@pytest.mark.parametrize("name", sorted(CASES) + sorted(binary_cases()))
def test_reader_matches_reference(name):
    data = CASES.get(name) or binary_cases()[name]
    st, r = ref_ply.read(data)
    assert st == 0
    g = ply.read_ply_bytes(data, raise_errors=True)
    # with neither a face nor a tet element the reference leaves geometry_type uninitialized
    same_arrays(g, r, check_type=name != "no_mesh_element" and name != "unknown_format_is_ascii")


@needs_ref
@pytest.mark.parametrize("name", sorted(BAD))
def test_malformed_files_fail_like_the_reference(name):
    st, _ = ref_ply.read(BAD[name])
    assert st != 0  # no value (1) or an exception (2)
    assert ply.read_ply_bytes(BAD[name]) is None
    with pytest.raises(S.SbsError):
        ply.read_ply_bytes(BAD[name], raise_errors=True)


@needs_ref
def test_float_tokens_parse_like_stof():
    rng = np.random.default_rng(17)
    toks = []
    vals = rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000)
    for i, v in enumerate(vals):
        k = i % 6
        toks.append("%.*f" % (int(rng.integers(0, 12)), v) if k == 0 else "%.*e" % (int(rng.integers(0, 20)), v)
                    if k == 1 else repr(float(v)) if k == 2 else "%.9g" % v if k == 3
                    else "".join(rng.choice(list("0123456789"), int(rng.integers(1, 30)))) + "." +
                    "".join(rng.choice(list("0123456789"), int(rng.integers(0, 30)))) if k == 4
                    else "%.7f" % (int(rng.integers(-10**6, 10**6)) / 128.0))
    toks = [t for t in toks if "e+3" not in t and "e-3" not in t and "e+4" not in t and "e-4" not in t]
    toks = toks[:len(toks) // 3 * 3]
    nv = len(toks) // 3
    data = header("ascii", f"element vertex {nv}", "property float x", "property float y", "property float z") + \
        "".join(" ".join(toks[3 * i:3 * i + 3]) + "\n" for i in range(nv)).encode()
    st, r = ref_ply.read(data)
    assert st == 0
    same_arrays(ply.read_ply_bytes(data, raise_errors=True), r, check_type=False)


def test_round_trips_and_path_rules(tmp_path):
    g = sample_geometry(np.random.default_rng(9), 500, 800, ply.TETRAHEDRON)
    for fmt in ply.FORMATS:
        p = tmp_path / f"mesh_{fmt}.ply"
        ply.write_ply(p, g, fmt)
        back = ply.read_ply(p)
        assert back is not None and back.is_tetrahedral_mesh()
        assert np.array_equal(back.indices, g.indices)
        if fmt != "ascii":
            assert np.array_equal(back.positions.view(np.uint32), g.positions.view(np.uint32))
    assert ply.read_ply(tmp_path / "missing.ply") is None
    assert ply.read_ply(tmp_path / "mesh_ascii.txt") is None
    with pytest.raises(S.SbsError):
        ply.write_ply(tmp_path / "mesh.obj", g)
    with pytest.raises(S.SbsError):
        ply.read_ply(tmp_path / "missing.ply", raise_errors=True)


def test_truncated_files_are_errors():
    g = sample_geometry(np.random.default_rng(2), 50, 40, ply.TRIANGLE)
    for fmt in ply.FORMATS:
        data = ply.write_ply_bytes(g, fmt)
        assert ply.read_ply_bytes(data) is not None
        assert ply.read_ply_bytes(data[:-7]) is None


def test_library_mesh_round_trip_through_files(tmp_path):
    pos, tets = S.box6(6, 3, 2, 0.1)
    p = tmp_path / "beam.ply"
    ply.write_ply(p, ply.tet_geometry(pos, tets), "binary_little_endian")
    g = ply.read_ply(p)
    assert np.array_equal(g.indices.reshape(-1, 4), tets.astype(np.int32))
    assert np.array_equal(g.positions.reshape(-1, 3), pos.astype(np.float32))
    s2t, tris = S.boundary(len(pos), g.indices.reshape(-1, 4))
    assert (len(s2t), len(tris)) == (7 * 4 * 3 - 5 * 2 * 1, 4 * (6 * 3 + 3 * 2 + 6 * 2))


@needs_ref
def test_million_tet_beam_file_matches_reference():
    pos, tets = S.box6(220, 28, 27, 0.1)
    g = ply.tet_geometry(pos, tets)
    t0 = time.perf_counter()
    ours = ply.write_ply_bytes(g, "ascii")
    t1 = time.perf_counter()
    ref = ref_write(g, 0)
    t2 = time.perf_counter()
    assert ours == ref
    back = ply.read_ply_bytes(ours)
    t3 = time.perf_counter()
    st, r = ref_ply.read(ours)
    t4 = time.perf_counter()
    assert st == 0
    same_arrays(back, r)
    print(f"\n{len(ours) / 1e6:.1f} MB: write {t1 - t0:.3f}s (reference {t2 - t1:.3f}s), "
          f"read {t3 - t2:.3f}s (reference {t4 - t3:.3f}s)")


def test_second_vertex_element_larger_than_the_arrays_is_an_error():
    # the arrays are sized from the last vertex element naming x, y, z (one item); a second vertex
    # element of 64 items naming only x would write past them
    for fmt in ("ascii", "binary_little_endian", "binary_big_endian"):
        h = header(fmt, "element vertex 1", "property float x", "property float y", "property float z",
                   "element vertex 64", "property float x")
        if fmt == "ascii":
            body = b"1 2 3\n" + b"".join(b"%d\n" % i for i in range(64))
        else:
            end = "<" if fmt.endswith("little_endian") else ">"
            body = np.array([1, 2, 3], end + "f4").tobytes() + np.arange(64, dtype=end + "f4").tobytes()
        data = h + body
        assert ply.read_ply_bytes(data) is None
        with pytest.raises(S.SbsError) as e:
            ply.read_ply_bytes(data, raise_errors=True)
        assert e.value.code == 8  # SBS_ERR_FORMAT
    # the same elements in the other order fit: the arrays hold 64 positions
    h = header("ascii", "element vertex 64", "property float x", "property float y", "property float z",
               "element vertex 1", "property float x")
    g = ply.read_ply_bytes(h + b"".join(b"%d 0 0\n" % i for i in range(64)) + b"7\n", raise_errors=True)
    assert g.positions.reshape(-1, 3)[0, 0] == 7.0 and g.positions.reshape(-1, 3)[63, 0] == 63.0
