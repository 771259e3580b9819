"""Test infrastructure only: ctypes access to the reference's own PLY reader and writer, built from
the reference's sources into oracle/_ref/libref_ply.so (`make -C oracle ref`). available() is
False where that build is absent; tests that need it skip then."""
import ctypes
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libref_ply.so")


class RefGeometry(ctypes.Structure):
    _fields_ = [("positions", ctypes.POINTER(ctypes.c_float)), ("num_positions", ctypes.c_uint64),
                ("indices", ctypes.POINTER(ctypes.c_int32)), ("num_indices", ctypes.c_uint64),
                ("normals", ctypes.POINTER(ctypes.c_float)), ("num_normals", ctypes.c_uint64),
                ("uvs", ctypes.POINTER(ctypes.c_float)), ("num_uvs", ctypes.c_uint64),
                ("colors", ctypes.POINTER(ctypes.c_uint8)), ("num_colors", ctypes.c_uint64),
                ("geometry_type", ctypes.c_int32)]


_L = None


def available():
    return os.path.exists(PATH)


def _lib():
    global _L
    if _L is None:
        L = ctypes.CDLL(PATH)
        L.ref_ply_read_memory.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(RefGeometry)]
        L.ref_ply_read_memory.restype = ctypes.c_int
        L.ref_ply_write_memory.argtypes = [ctypes.POINTER(RefGeometry), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64)]
        L.ref_ply_write_memory.restype = ctypes.c_int
        L.ref_free.argtypes = [ctypes.c_void_p]
        L.ref_free.restype = None
        _L = L
    return _L


def read(data):
    """(status, dict of arrays): status 0 a geometry, 1 no value, 2 an exception."""
    g = RefGeometry()
    st = _lib().ref_ply_read_memory(bytes(data), len(data), ctypes.byref(g))
    if st != 0:
        return st, None
    out = {}
    for name, dt in (("positions", np.float32), ("indices", np.int32), ("normals", np.float32), ("uvs", np.float32),
                     ("colors", np.uint8)):
        n = getattr(g, "num_" + name)
        p = getattr(g, name)
        out[name] = np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)
        if n:
            _lib().ref_free(ctypes.cast(p, ctypes.c_void_p))
    out["geometry_type"] = int(g.geometry_type)
    return 0, out


def write(positions=(), indices=(), normals=(), uvs=(), colors=(), geometry_type=0, fmt=0):
    arrs = {"positions": np.ascontiguousarray(positions, np.float32).reshape(-1),
            "indices": np.ascontiguousarray(indices, np.int32).reshape(-1),
            "normals": np.ascontiguousarray(normals, np.float32).reshape(-1),
            "uvs": np.ascontiguousarray(uvs, np.float32).reshape(-1),
            "colors": np.ascontiguousarray(colors, np.uint8).reshape(-1)}
    g = RefGeometry()
    for name, a in arrs.items():
        setattr(g, name, a.ctypes.data_as(type(getattr(g, name))) if a.size else None)
        setattr(g, "num_" + name, a.size)
    g.geometry_type = geometry_type
    out = ctypes.c_void_p()
    n = ctypes.c_uint64()
    _lib().ref_ply_write_memory(ctypes.byref(g), fmt, ctypes.byref(out), ctypes.byref(n))
    data = ctypes.string_at(out, n.value)
    _lib().ref_free(out)
    return data
