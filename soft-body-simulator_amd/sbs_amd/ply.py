"""PLY files of the geometry record (src/io/ply.cpp, include/sbs/io/ply.h): read_ply and write_ply
with the reference's names, formats and array layout, over the library's C entry points."""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import GeometryRecord, SbsError, check, lib

ASCII, BINARY_LITTLE_ENDIAN, BINARY_BIG_ENDIAN = 0, 1, 2  # io::ply_format_t
FORMATS = {"ascii": ASCII, "binary_little_endian": BINARY_LITTLE_ENDIAN, "binary_big_endian": BINARY_BIG_ENDIAN}
UNSET, TRIANGLE, TETRAHEDRON = -1, 0, 1  # common::geometry_t::geometry_type_t (UNSET: neither element)


@dataclass
class Geometry:
    """common::geometry_t (include/sbs/common/geometry.h:10-30): flat arrays."""
    positions: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    indices: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    normals: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    uvs: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    colors: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint8))
    geometry_type: int = UNSET

    def is_triangle_mesh(self):
        return self.geometry_type == TRIANGLE

    def is_tetrahedral_mesh(self):
        return self.geometry_type == TETRAHEDRON

    def record(self):
        """A GeometryRecord viewing this geometry's arrays (keep the geometry alive while in use)."""
        self.positions = np.ascontiguousarray(self.positions, np.float32).reshape(-1)
        self.indices = np.ascontiguousarray(self.indices, np.int32).reshape(-1)
        self.normals = np.ascontiguousarray(self.normals, np.float32).reshape(-1)
        self.uvs = np.ascontiguousarray(self.uvs, np.float32).reshape(-1)
        self.colors = np.ascontiguousarray(self.colors, np.uint8).reshape(-1)
        r = GeometryRecord()
        for name in ("positions", "indices", "normals", "uvs", "colors"):
            a = getattr(self, name)
            setattr(r, name, a.ctypes.data_as(type(getattr(r, name))) if a.size else None)
            setattr(r, "num_" + name, a.size)
        r.geometry_type = self.geometry_type
        return r


def _take(rec):
    def arr(ptr, n, dt):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)
    try:
        return Geometry(arr(rec.positions, rec.num_positions, np.float32), arr(rec.indices, rec.num_indices, np.int32),
                        arr(rec.normals, rec.num_normals, np.float32), arr(rec.uvs, rec.num_uvs, np.float32),
                        arr(rec.colors, rec.num_colors, np.uint8), int(rec.geometry_type))
    finally:
        lib().sbs_geometry_release(ctypes.byref(rec))


def _format(fmt):
    return FORMATS[fmt] if isinstance(fmt, str) else int(fmt)


def read_ply(path, raise_errors=False):
    """io::read_ply(path): the geometry, or None where the reference returns no value (a missing
    file, a path not ending in .ply, a malformed file); raise_errors=True raises SbsError instead."""
    rec = GeometryRecord()
    code = lib().sbs_ply_read(os.fsencode(path), ctypes.byref(rec))
    if code != 0:
        if raise_errors:
            check(code)
        return None
    return _take(rec)


def read_ply_bytes(data, raise_errors=False):
    """io::read_ply(std::istream&) over a file image."""
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    rec = GeometryRecord()
    code = lib().sbs_ply_read_memory(ctypes.cast(buf, ctypes.c_void_p), len(data), ctypes.byref(rec))
    if code != 0:
        if raise_errors:
            check(code)
        return None
    return _take(rec)


def write_ply(path, geometry, format="ascii"):
    """io::write_ply(path, geometry, format); raises SbsError when the file cannot be written."""
    rec = geometry.record()
    check(lib().sbs_ply_write(os.fsencode(path), ctypes.byref(rec), _format(format)))


def write_ply_bytes(geometry, format="ascii"):
    """io::write_ply(std::ostream&, geometry, format): the file image as bytes."""
    rec = geometry.record()
    n = ctypes.c_uint64()
    check(lib().sbs_ply_write_memory(ctypes.byref(rec), _format(format), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(max(n.value, 1))
    check(lib().sbs_ply_write_memory(ctypes.byref(rec), _format(format), ctypes.cast(buf, ctypes.c_void_p), n.value,
                                     ctypes.byref(n)))
    return buf.raw[:n.value]


def tet_geometry(positions, tets):
    """A tetrahedral geometry from (V, 3) positions and (T, 4) tets."""
    return Geometry(positions=np.asarray(positions, np.float32).reshape(-1),
                    indices=np.asarray(tets, np.int64).astype(np.int32).reshape(-1), geometry_type=TETRAHEDRON)


def triangle_geometry(positions, triangles, normals=None):
    """A triangle geometry from (V, 3) positions, (F, 3) triangles and optional (V, 3) normals."""
    g = Geometry(positions=np.asarray(positions, np.float32).reshape(-1),
                 indices=np.asarray(triangles, np.int64).astype(np.int32).reshape(-1), geometry_type=TRIANGLE)
    if normals is not None:
        g.normals = np.asarray(normals, np.float32).reshape(-1)
    return g


def read_tet_mesh(path):
    """(positions (V, 3) float64, tets (T, 4) uint32) of a PLY tet mesh, as tetrahedral_body_t takes
    them from a geometry_t (src/physics/tetrahedral_body.cpp:30-62: float positions widened)."""
    g = read_ply(path, raise_errors=True)
    if not g.is_tetrahedral_mesh():
        raise ValueError(f"{path} holds no tet element")
    if g.indices.size and (g.indices.min() < 0 or g.indices.max() >= g.positions.size // 3):
        raise ValueError(f"{path}: tet index out of range")
    return g.positions.astype(np.float64).reshape(-1, 3), g.indices.astype(np.uint32).reshape(-1, 4)


__all__ = ["ASCII", "BINARY_BIG_ENDIAN", "BINARY_LITTLE_ENDIAN", "FORMATS", "Geometry", "SbsError", "TETRAHEDRON",
           "TRIANGLE", "UNSET", "read_ply", "read_ply_bytes", "read_tet_mesh", "tet_geometry", "triangle_geometry", "write_ply",
           "write_ply_bytes"]
