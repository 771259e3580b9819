"""ctypes bindings of the C ABI in include/sbs_gpu.h (libsbs_gpu.so).

This module is plumbing for tests and benchmarks; the product is the native library. Loading
fails loudly when the library has not been built -- there is no CPU fallback.
"""
import ctypes
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "_build", "libsbs_gpu.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "sbs_gpu.h")

SBS_OK = 0
(KERNEL_PREDICT, KERNEL_CONTACT, KERNEL_GREEN, KERNEL_NEOHOOKEAN, KERNEL_DISTANCE, KERNEL_COMMIT, KERNEL_DETECT,
 KERNEL_WARM_START) = range(8)
KERNEL_NAMES = ["predict", "contact", "green", "neohookean", "distance", "commit", "detect", "warm_start"]
WARM_START = {"none": 0, "previous": 1, "learned": 2, "explicit": 3}

_d = ctypes.c_double
_i = ctypes.c_int
_u32 = ctypes.c_uint32
_p = ctypes.c_void_p
_pd = ctypes.POINTER(ctypes.c_double)
_pu = ctypes.POINTER(ctypes.c_uint32)
_pu64 = ctypes.POINTER(ctypes.c_uint64)
_u64 = ctypes.c_uint64
_f = ctypes.c_float
_pf = ctypes.POINTER(ctypes.c_float)


class XpbdParams(ctypes.Structure):
    _fields_ = [
        ("compliance", ctypes.c_double),
        ("damping", ctypes.c_double),
        ("young_modulus", ctypes.c_double),
        ("poisson_ratio", ctypes.c_double),
        ("hooke_coefficient", ctypes.c_double),
        ("collision_compliance", ctypes.c_double),
        ("collision_damping", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("gravity_is_force", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class GeometryRecord(ctypes.Structure):
    """struct sbs_geometry of include/sbs_gpu.h."""
    _fields_ = [("positions", ctypes.POINTER(ctypes.c_float)), ("num_positions", ctypes.c_uint64),
                ("indices", ctypes.POINTER(ctypes.c_int32)), ("num_indices", ctypes.c_uint64),
                ("normals", ctypes.POINTER(ctypes.c_float)), ("num_normals", ctypes.c_uint64),
                ("uvs", ctypes.POINTER(ctypes.c_float)), ("num_uvs", ctypes.c_uint64),
                ("colors", ctypes.POINTER(ctypes.c_uint8)), ("num_colors", ctypes.c_uint64),
                ("geometry_type", ctypes.c_int32)]


_pgeo = ctypes.POINTER(GeometryRecord)

_SIGS = {
    "sbs_ply_read": [ctypes.c_char_p, _pgeo],
    "sbs_ply_read_memory": [_p, _u64, _pgeo],
    "sbs_geometry_release": [_pgeo],
    "sbs_ply_write": [ctypes.c_char_p, _pgeo, _i],
    "sbs_ply_write_memory": [_pgeo, _i, _p, _u64, _pu64],
    "sbs_gpu_default_params": [ctypes.POINTER(XpbdParams)],
    "sbs_gpu_create": [_i, ctypes.POINTER(_p)],
    "sbs_gpu_destroy": [_p],
    "sbs_gpu_set_params": [_p, ctypes.POINTER(XpbdParams)],
    "sbs_gpu_add_body": [_p, _pd, _pd, _u32, _pu],
    "sbs_gpu_set_collision_surface": [_p, _u32, _pu, _u32],
    "sbs_gpu_add_green_constraints": [_p, _u32, _pu, _u32, _d, _d, _d, _d, _pu],
    "sbs_gpu_add_neohookean_constraints": [_p, _u32, _pu, _u32, _d, _d, _d, _pu],
    "sbs_gpu_add_distance_constraints": [_p, _pu, _u32, _d, _d, _pu],
    "sbs_gpu_add_green_constraints_batched": [_p, _pu, _pu, _u32, _d, _d, _pd, _pd, _pu],
    "sbs_gpu_add_neohookean_constraints_batched": [_p, _pu, _pu, _u32, _d, _pd, _pd, _pu],
    "sbs_mesh_ellipsoid6": [_d, _d, _d, _d, _pd, _pu, _pu, _pu],
    "sbs_gpu_add_plane": [_p, _pd, _pd, _pd, _pd],
    "sbs_gpu_add_sphere": [_p, _pd, _d],
    "sbs_gpu_add_grid_sdf": [_p, _pd, _pd, _pu, _pd, _pd, _pd],
    "sbs_gpu_add_mesh_obstacle": [_p, _pd, _u32, _pu, _u32, _pd, _pd, _pu],
    "sbs_gpu_evaluate_sdf": [_p, _u32, _pd, _u32, _pd, _pd],
    "sbs_sdf_grid_node_count": [_pu, _pu64],
    "sbs_sdf_grid_node_positions": [_pd, _pd, _pu, _pd],
    "sbs_sdf_environment_domain": [_pd, _u32, _pd, _pd, _pd, _pd],
    "sbs_sdf_grid_from_mesh": [_i, _pd, _u32, _pu, _u32, _pd, _pd, _pu, _pd],
    "sbs_gpu_set_group_clusters": [_p, _u32, _pu, _u32],
    "sbs_gpu_set_coloring": [_p, _i],
    "sbs_gpu_set_launch_config": [_p, _i, _i],
    "sbs_gpu_set_contact_fusion": [_p, _i],
    "sbs_gpu_commit": [_p, _u32],
    "sbs_gpu_step": [_p, _d, _u32, _u32],
    "sbs_gpu_synchronize": [_p],
    "sbs_gpu_use_graphs": [_p, _i],
    "sbs_gpu_read_body": [_p, _u32, _pd, _pd],
    "sbs_gpu_write_body": [_p, _u32, _pd, _pd],
    "sbs_gpu_group_colors": [_p, _u32, _pu, _pu],
    "sbs_gpu_read_contacts": [_p, _pu, _pd, _u32, _pu],
    "sbs_gpu_residual": [_p, _u32, _d, _pd, _u32],
    "sbs_gpu_energy": [_p, _pd],
    "sbs_gpu_set_timing": [_p, _i],
    "sbs_gpu_kernel_stats": [_p, _i, _pu64, _pd, _pu64],
    "sbs_gpu_reset_stats": [_p],
    "sbs_mlp_create": [_i, _u32, _u32, _u64, ctypes.POINTER(_p)],
    "sbs_mlp_destroy": [_p],
    "sbs_mlp_num_params": [_pu],
    "sbs_mlp_get_params": [_p, _pf],
    "sbs_mlp_set_params": [_p, _pf],
    "sbs_mlp_set_normalization": [_p, _pf, _pf, _pf],
    "sbs_mlp_get_normalization": [_p, _pf, _pf, _pf],
    "sbs_mlp_forward": [_p, _pf, _u32, _pf],
    "sbs_mlp_loss_grad": [_p, _pf, _pf, _u32, _pd, _pf],
    "sbs_mlp_dataset_reserve": [_p, _u64],
    "sbs_mlp_dataset_append": [_p, _pf, _pf, _u64],
    "sbs_mlp_dataset_size": [_p, _pu64],
    "sbs_mlp_dataset_clear": [_p],
    "sbs_mlp_dataset_read": [_p, _u64, _u64, _pf, _pf],
    "sbs_mlp_fit_normalization": [_p],
    "sbs_mlp_train": [_p, _u32, _u32, _f, _u64, _pd],
    "sbs_gpu_set_warm_start": [_p, _i, _p],
    "sbs_gpu_set_contact_hash": [_p, _i, _d],
    "sbs_gpu_set_surface_triangles": [_p, _u32, _pu, _u32],
    "sbs_gpu_set_surface_refresh": [_p, _i],
    "sbs_gpu_read_surface": [_p, _u32, _pd, _pd],
    "sbs_gpu_write_surface_ply": [_p, _u32, ctypes.c_char_p, _i],
    "sbs_gpu_read_contact_hash": [_p, _pd, _pu, _pu, _pu, _pu, _u32, _pu, _pu, _pu],
    "sbs_gpu_contact_overflow": [_p, _pu64],
    "sbs_gpu_set_body_contacts": [_p, _i, _d, _d, _d],
    "sbs_gpu_read_body_contact_setup": [_p, _pd, _pd, _pd, _pu, _pu],
    "sbs_gpu_read_body_contact_candidates": [_p, _pu, _u64, _pu64],
    "sbs_util_exclusive_scan": [_i, _pu, _u64, _pu],
    "sbs_util_sort_pairs": [_i, _pu, _pu, _u64, _i, _pu, _pu],
    "sbs_util_sort_pairs_noalloc": [_i, _p, _p, _u64, _i, _p],
    "sbs_gpu_extract_boundary": [_i, _u32, _pu, _u32, _pu, _pu, _pu, _pu],
    "sbs_gpu_capture": [_p, _p, _u32],
    "sbs_gpu_set_warm_start_scale": [_p, _d],
    "sbs_gpu_set_warm_start_scales": [_p, _d, _d],
    "sbs_gpu_snapshot": [_p],
    "sbs_gpu_restore": [_p],
    "sbs_gpu_snapshot_slot": [_p, _i],
    "sbs_gpu_restore_slot": [_p, _i],
    "sbs_gpu_set_fallback": [_p, _d, _u32],
    "sbs_gpu_fallback_count": [_p, _pu64],
    "sbs_mesh_bar5": [_u32, _u32, _u32, _pd, _pu],
    "sbs_mesh_box6": [_u32, _u32, _u32, _d, _pd, _pd, _pu],
    "sbs_mesh_boundary": [_u32, _pu, _u32, _pu, _pu, _pu, _pu],
    "sbs_mesh_lumped_mass": [_u32, _pd, _pu, _u32, _d, _pd],
}

_lib = None


class SbsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sbs error {code}: {msg}")
        self.code = code


def header_functions():
    """Names of the functions include/sbs_gpu.h declares."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sbs_\w+)\s*\(", text, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.restype = _i
            f.argtypes = args
        L.sbs_gpu_last_error.restype = ctypes.c_char_p
        L.sbs_gpu_last_error.argtypes = []
        _lib = L
    return _lib


def check(code):
    if code != SBS_OK:
        raise SbsError(code, lib().sbs_gpu_last_error().decode())


def _dp(a):
    return None if a is None else a.ctypes.data_as(_pd)


def _up(a):
    return None if a is None else a.ctypes.data_as(_pu)


def _fp(a):
    return None if a is None else a.ctypes.data_as(_pf)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ------------------------------------------------------------------ host mesh utilities

def bar5(w, h, d):
    nt = max(w - 1, 0) * max(h - 1, 0) * max(d - 1, 0) * 5
    pos = np.zeros(w * h * d * 3)
    tets = np.zeros(nt * 4, dtype=np.uint32)
    check(lib().sbs_mesh_bar5(w, h, d, _dp(pos), _up(tets)))
    return pos.reshape(-1, 3), tets.reshape(-1, 4)


def box6(nx, ny, nz, edge, origin=(0.0, 0.0, 0.0)):
    nv = (nx + 1) * (ny + 1) * (nz + 1)
    pos = np.zeros(nv * 3)
    tets = np.zeros(nx * ny * nz * 24, dtype=np.uint32)
    check(lib().sbs_mesh_box6(nx, ny, nz, edge, _dp(f64(origin)), _dp(pos), _up(tets)))
    return pos.reshape(-1, 3), tets.reshape(-1, 4)


def boundary(nv, tets):
    tets = u32(tets).reshape(-1)
    ns, ntr = _u32(), _u32()
    check(lib().sbs_mesh_boundary(nv, _up(tets), len(tets) // 4, None, None, ctypes.byref(ns), ctypes.byref(ntr)))
    s2t = np.zeros(ns.value, dtype=np.uint32)
    tris = np.zeros(ntr.value * 3, dtype=np.uint32)
    check(lib().sbs_mesh_boundary(nv, _up(tets), len(tets) // 4, _up(s2t), _up(tris), ctypes.byref(ns),
                                  ctypes.byref(ntr)))
    return s2t, tris.reshape(-1, 3)


def ellipsoid6(rx, ry, rz, edge):
    """Hex cells (edge `edge`) whose centers lie inside the ellipsoid, 6 tets each."""
    nv, nt = _u32(), _u32()
    check(lib().sbs_mesh_ellipsoid6(rx, ry, rz, edge, None, None, ctypes.byref(nv), ctypes.byref(nt)))
    pos = np.zeros(nv.value * 3)
    tets = np.zeros(nt.value * 4, dtype=np.uint32)
    check(lib().sbs_mesh_ellipsoid6(rx, ry, rz, edge, _dp(pos), _up(tets), ctypes.byref(nv), ctypes.byref(nt)))
    return pos.reshape(-1, 3), tets.reshape(-1, 4)


def lumped_mass(pos, tets, density):
    pos = f64(pos).reshape(-1)
    tets = u32(tets).reshape(-1)
    m = np.zeros(len(pos) // 3)
    check(lib().sbs_mesh_lumped_mass(len(m), _dp(pos), _up(tets), len(tets) // 4, density, _dp(m)))
    return m


# ------------------------------------------------------------------ device primitives

def exclusive_scan(values, device=0):
    """Exclusive prefix sum on the device: n + 1 words (the last is the total)."""
    a = u32(values)
    out = np.zeros(len(a) + 1, dtype=np.uint32)
    check(lib().sbs_util_exclusive_scan(device, _up(a), len(a), _up(out)))
    return out


def extract_boundary(nv, tets, device=0):
    """Boundary surface of a tet mesh computed on the device; the same result as boundary()."""
    tets = u32(tets).reshape(-1)
    nt = len(tets) // 4
    ns, ntr = _u32(), _u32()
    s2t = np.zeros(max(nv, 1), dtype=np.uint32)
    tris = np.zeros(max(12 * nt, 1), dtype=np.uint32)
    check(lib().sbs_gpu_extract_boundary(device, nv, _up(tets), nt, _up(s2t), _up(tris), ctypes.byref(ns),
                                         ctypes.byref(ntr)))
    return s2t[:ns.value].copy(), tris[:3 * ntr.value].reshape(-1, 3).copy()


def grid_node_count(res):
    """Node count of a cubic Lagrange grid of res[0] x res[1] x res[2] cells."""
    n = ctypes.c_uint64()
    check(lib().sbs_sdf_grid_node_count(_up(u32(res)), ctypes.byref(n)))
    return n.value


def grid_node_positions(lo, hi, res):
    """(N, 3) node positions of a cubic Lagrange grid over the box [lo, hi]."""
    out = np.zeros(3 * grid_node_count(res))
    check(lib().sbs_sdf_grid_node_positions(_dp(f64(lo)), _dp(f64(hi)), _up(u32(res)), _dp(out)))
    return out.reshape(-1, 3)


def environment_domain(positions, lo, hi):
    """environment_body_t's extended grid domain for a mesh and a requested domain."""
    p = f64(positions).reshape(-1)
    a, b = np.zeros(3), np.zeros(3)
    check(lib().sbs_sdf_environment_domain(_dp(p), len(p) // 3, _dp(f64(lo)), _dp(f64(hi)), _dp(a), _dp(b)))
    return a, b


def grid_from_mesh(positions, triangles, lo, hi, res, device=0):
    """Node values of a triangle mesh's signed distance on a cubic Lagrange grid, on the device."""
    p, t = f64(positions).reshape(-1), u32(triangles).reshape(-1)
    out = np.zeros(grid_node_count(res))
    check(lib().sbs_sdf_grid_from_mesh(device, _dp(p), len(p) // 3, _up(t), len(t) // 3, _dp(f64(lo)), _dp(f64(hi)),
                                       _up(u32(res)), _dp(out)))
    return out


def sort_pairs(keys, vals, bits=32, device=0):
    """Stable device radix sort of (keys, vals) by the low `bits` bits of the keys."""
    k, v = u32(keys), u32(vals)
    ok, ov = np.zeros_like(k), np.zeros_like(v)
    check(lib().sbs_util_sort_pairs(device, _up(k), _up(v), len(k), bits, _up(ok), _up(ov)))
    return ok, ov


# ------------------------------------------------------------------ warm-start network

class Mlp:
    """The warm-start network behind the C ABI (one sbs_mlp handle)."""

    def __init__(self, num_inputs=39, num_outputs=12, seed=0, device=0):
        self._L = lib()
        h = _p()
        check(self._L.sbs_mlp_create(device, num_inputs, num_outputs, seed, ctypes.byref(h)))
        self.h = h
        self.nin, self.nout = num_inputs, num_outputs
        n = _u32()
        check(self._L.sbs_mlp_num_params(ctypes.byref(n)))
        self.num_params = n.value

    def close(self):
        if getattr(self, "h", None):
            check(self._L.sbs_mlp_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        check(self._L.sbs_mlp_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = f32(p)
        assert len(p) == self.num_params
        check(self._L.sbs_mlp_set_params(self.h, _fp(p)))

    def set_normalization(self, mean, inv_std, out_scale):
        check(self._L.sbs_mlp_set_normalization(self.h, _fp(f32(mean)), _fp(f32(inv_std)), _fp(f32(out_scale))))

    def get_normalization(self):
        m, s, o = np.zeros(self.nin, np.float32), np.zeros(self.nin, np.float32), np.zeros(self.nout, np.float32)
        check(self._L.sbs_mlp_get_normalization(self.h, _fp(m), _fp(s), _fp(o)))
        return m, s, o

    def forward(self, x):
        x = f32(x).reshape(-1, self.nin)
        y = np.zeros((len(x), self.nout), dtype=np.float32)
        check(self._L.sbs_mlp_forward(self.h, _fp(x), len(x), _fp(y)))
        return y

    def loss_grad(self, x, t):
        x = f32(x).reshape(-1, self.nin)
        t = f32(t).reshape(-1, self.nout)
        loss = ctypes.c_double()
        g = np.zeros(self.num_params, dtype=np.float32)
        check(self._L.sbs_mlp_loss_grad(self.h, _fp(x), _fp(t), len(x), ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def reserve(self, rows):
        check(self._L.sbs_mlp_dataset_reserve(self.h, rows))

    def append(self, x, t):
        x = f32(x).reshape(-1, self.nin)
        t = f32(t).reshape(-1, self.nout)
        check(self._L.sbs_mlp_dataset_append(self.h, _fp(x), _fp(t), len(x)))

    def dataset_size(self):
        n = ctypes.c_uint64()
        check(self._L.sbs_mlp_dataset_size(self.h, ctypes.byref(n)))
        return n.value

    def clear(self):
        check(self._L.sbs_mlp_dataset_clear(self.h))

    def dataset(self, first=0, rows=None):
        """Rows of the device-resident training set: features [rows][num_inputs], targets [rows][num_outputs]."""
        rows = self.dataset_size() - first if rows is None else rows
        x = np.zeros((rows, self.nin), dtype=np.float32)
        t = np.zeros((rows, self.nout), dtype=np.float32)
        check(self._L.sbs_mlp_dataset_read(self.h, first, rows, _fp(x), _fp(t)))
        return x, t

    def fit_normalization(self):
        check(self._L.sbs_mlp_fit_normalization(self.h))

    def train(self, steps, batch, lr, seed=0):
        loss = ctypes.c_double()
        check(self._L.sbs_mlp_train(self.h, steps, batch, lr, seed, ctypes.byref(loss)))
        return loss.value


# ------------------------------------------------------------------ simulation handle

class Simulation:
    """A device simulation behind the C ABI (one sbs_gpu handle)."""

    def __init__(self, device=0):
        self._L = lib()
        h = _p()
        check(self._L.sbs_gpu_create(device, ctypes.byref(h)))
        self.h = h
        self.bodies = []
        self.groups = []  # constraint counts
        self.params = XpbdParams()
        check(self._L.sbs_gpu_default_params(ctypes.byref(self.params)))

    def close(self):
        if getattr(self, "h", None):
            check(self._L.sbs_gpu_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, gravity=(0.0, -9.81, 0.0), gravity_is_force=True, collision_compliance=1e-8,
                   collision_damping=0.0):
        p = self.params
        p.gravity[0], p.gravity[1], p.gravity[2] = gravity
        p.gravity_is_force = int(gravity_is_force)
        p.collision_compliance = collision_compliance
        p.collision_damping = collision_damping
        check(self._L.sbs_gpu_set_params(self.h, ctypes.byref(p)))

    def add_body(self, x0, mass=None):
        x0 = f64(x0).reshape(-1)
        m = None if mass is None else f64(mass)
        b = _u32()
        check(self._L.sbs_gpu_add_body(self.h, _dp(x0), _dp(m), len(x0) // 3, ctypes.byref(b)))
        self.bodies.append(len(x0) // 3)
        return b.value

    def set_surface(self, body, s2t):
        s2t = u32(s2t)
        check(self._L.sbs_gpu_set_collision_surface(self.h, body, _up(s2t), len(s2t)))
        self._surface_sizes = getattr(self, "_surface_sizes", {})
        self._surface_sizes[body] = len(s2t)

    num_triangles = 0

    def set_surface_triangles(self, body, triangles):
        t = u32(triangles).reshape(-1)
        check(self._L.sbs_gpu_set_surface_triangles(self.h, body, _up(t), len(t) // 3))
        self.num_triangles += len(t) // 3

    def set_surface_refresh(self, enable=True):
        check(self._L.sbs_gpu_set_surface_refresh(self.h, int(enable)))

    def read_surface(self, body):
        """(positions, normals) of the body's surface vertices, (n, 3) each."""
        n = self._surface_sizes.get(body, 0)
        x, nv = np.zeros(3 * n), np.zeros(3 * n)
        check(self._L.sbs_gpu_read_surface(self.h, body, _dp(x), _dp(nv)))
        return x.reshape(-1, 3), nv.reshape(-1, 3)

    def write_surface_ply(self, body, path, format="ascii"):
        """Frame dump of the body's surface mesh as a PLY triangle mesh (with the refreshed normals
        when the surface refresh is on)."""
        fmt = {"ascii": 0, "binary_little_endian": 1, "binary_big_endian": 2}[format] if isinstance(format, str) \
            else int(format)
        check(self._L.sbs_gpu_write_surface_ply(self.h, body, os.fsencode(path), fmt))

    def _group(self, code, out, n):
        check(code)
        self.groups.append(n)
        return out.value

    def add_green(self, body, tets, alpha, beta, young, poisson):
        t = u32(tets).reshape(-1)
        g = _u32()
        return self._group(self._L.sbs_gpu_add_green_constraints(self.h, body, _up(t), len(t) // 4, alpha, beta, young,
                                                                 poisson, ctypes.byref(g)), g, len(t) // 4)

    def add_neohookean(self, body, tets, beta, young, poisson):
        t = u32(tets).reshape(-1)
        g = _u32()
        return self._group(self._L.sbs_gpu_add_neohookean_constraints(self.h, body, _up(t), len(t) // 4, beta, young,
                                                                      poisson, ctypes.byref(g)), g, len(t) // 4)

    def add_green_batched(self, bodies, tets, alpha, beta, young, poisson):
        b, t = u32(bodies), u32(tets).reshape(-1)
        g = _u32()
        return self._group(self._L.sbs_gpu_add_green_constraints_batched(
            self.h, _up(b), _up(t), len(b), alpha, beta, _dp(f64(young)), _dp(f64(poisson)), ctypes.byref(g)), g, len(b))

    def add_neohookean_batched(self, bodies, tets, beta, young, poisson):
        b, t = u32(bodies), u32(tets).reshape(-1)
        g = _u32()
        return self._group(self._L.sbs_gpu_add_neohookean_constraints_batched(
            self.h, _up(b), _up(t), len(b), beta, _dp(f64(young)), _dp(f64(poisson)), ctypes.byref(g)), g, len(b))

    def add_distance(self, b1v1b2v2, alpha, beta):
        e = u32(b1v1b2v2).reshape(-1)
        g = _u32()
        return self._group(self._L.sbs_gpu_add_distance_constraints(self.h, _up(e), len(e) // 4, alpha, beta,
                                                                    ctypes.byref(g)), g, len(e) // 4)

    def add_plane(self, normal, point, vmin=(-1e30,) * 3, vmax=(1e30,) * 3):
        check(self._L.sbs_gpu_add_plane(self.h, _dp(f64(normal)), _dp(f64(point)), _dp(f64(vmin)), _dp(f64(vmax))))

    def add_sphere(self, center, radius):
        check(self._L.sbs_gpu_add_sphere(self.h, _dp(f64(center)), radius))

    def add_grid_sdf(self, lo, hi, res, nodes, vmin=None, vmax=None):
        """Grid SDF obstacle from node values (cubic Lagrange grid); volume defaults to the domain."""
        nodes = f64(nodes)
        if len(nodes) != grid_node_count(res):
            raise ValueError(f"{len(nodes)} node values for a grid of {grid_node_count(res)} nodes")
        check(self._L.sbs_gpu_add_grid_sdf(self.h, _dp(f64(lo)), _dp(f64(hi)), _up(u32(res)), _dp(nodes),
                                           _dp(f64(lo if vmin is None else vmin)), _dp(f64(hi if vmax is None else vmax))))

    def add_mesh_obstacle(self, positions, triangles, lo, hi, res=(10, 10, 10)):
        """environment_body_t(geometry, domain, resolution): a triangle-mesh obstacle with its grid
        SDF built on the device."""
        p, t = f64(positions).reshape(-1), u32(triangles).reshape(-1)
        check(self._L.sbs_gpu_add_mesh_obstacle(self.h, _dp(p), len(p) // 3, _up(t), len(t) // 3, _dp(f64(lo)),
                                                _dp(f64(hi)), _up(u32(res))))

    def evaluate_sdf(self, k, points):
        """(signed distances (n,), gradients (n, 3)) of obstacle k at points, on the device."""
        p = f64(points).reshape(-1)
        n = len(p) // 3
        sd, g = np.zeros(n), np.zeros(3 * n)
        check(self._L.sbs_gpu_evaluate_sdf(self.h, k, _dp(p), n, _dp(sd), _dp(g)))
        return sd, g.reshape(-1, 3)

    def set_clusters(self, group, offsets):
        off = u32(offsets)
        check(self._L.sbs_gpu_set_group_clusters(self.h, group, _up(off), len(off) - 1))

    def set_coloring(self, order):
        """order: "hash" (device Jones-Plassmann) or "index" (sequential greedy in index order)."""
        check(self._L.sbs_gpu_set_coloring(self.h, {"hash": 0, "index": 1}[order]))

    def set_launch_config(self, block=256, cluster_block=64):
        check(self._L.sbs_gpu_set_launch_config(self.h, block, cluster_block))

    def set_contact_fusion(self, enable=True):
        check(self._L.sbs_gpu_set_contact_fusion(self.h, int(enable)))

    def commit(self, recolor_passes=2):
        check(self._L.sbs_gpu_commit(self.h, recolor_passes))

    def use_graphs(self, enable=True):
        check(self._L.sbs_gpu_use_graphs(self.h, int(enable)))

    def step(self, dt, iterations, substeps):
        check(self._L.sbs_gpu_step(self.h, dt, iterations, substeps))

    def synchronize(self):
        check(self._L.sbs_gpu_synchronize(self.h))

    def read_body(self, body):
        n = self.bodies[body]
        x = np.zeros(n * 3)
        v = np.zeros(n * 3)
        check(self._L.sbs_gpu_read_body(self.h, body, _dp(x), _dp(v)))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def write_body(self, body, x=None, v=None):
        x = None if x is None else f64(x).reshape(-1)
        v = None if v is None else f64(v).reshape(-1)
        check(self._L.sbs_gpu_write_body(self.h, body, _dp(x), _dp(v)))

    def group_colors(self, group):
        nc = _u32()
        c = np.zeros(self.groups[group], dtype=np.uint32)
        check(self._L.sbs_gpu_group_colors(self.h, group, ctypes.byref(nc), _up(c)))
        return c, nc.value

    def read_contacts(self):
        cnt = _u32()
        check(self._L.sbs_gpu_read_contacts(self.h, None, None, 0, ctypes.byref(cnt)))
        n = cnt.value
        ids = np.zeros(max(n, 1) * 4, dtype=np.uint32)
        qn = np.zeros(max(n, 1) * 6)
        check(self._L.sbs_gpu_read_contacts(self.h, _up(ids), _dp(qn), n, ctypes.byref(cnt)))
        return ids[:4 * n].reshape(-1, 4), qn[:6 * n].reshape(-1, 6)

    def residual(self, group, dt):
        _, nc = self.group_colors(group)
        out = np.zeros(max(nc, 1))
        check(self._L.sbs_gpu_residual(self.h, group, dt, _dp(out), len(out)))
        return out[:nc]

    def energy(self):
        out = np.zeros(6)
        check(self._L.sbs_gpu_energy(self.h, _dp(out)))
        return out

    warm_mode = "none"
    warm_net = None

    def set_warm_start(self, mode, net=None):
        """mode: "none", "previous", "explicit" or "learned" (needs net, an Mlp(39, 12))."""
        check(self._L.sbs_gpu_set_warm_start(self.h, WARM_START[mode], None if net is None else net.h))
        self.warm_mode, self.warm_net = mode, net

    def set_warm_start_scale(self, scale, hydrostatic=None):
        """Warm-start factor of the multipliers; with `hydrostatic`, scale applies to the deviatoric
        multipliers and `hydrostatic` to the hydrostatic ones."""
        if hydrostatic is None:
            check(self._L.sbs_gpu_set_warm_start_scale(self.h, scale))
        else:
            check(self._L.sbs_gpu_set_warm_start_scales(self.h, scale, hydrostatic))

    def capture(self, net, stride):
        check(self._L.sbs_gpu_capture(self.h, None if net is None else net.h, stride))

    def snapshot(self, slot=0):
        check(self._L.sbs_gpu_snapshot_slot(self.h, slot))

    def restore(self, slot=0):
        check(self._L.sbs_gpu_restore_slot(self.h, slot))

    def set_fallback(self, limit, iterations=0):
        """Residual limit of a warm-started frame, and the plain solver's iterations per substep
        when it is repeated (0: the step's own count)."""
        check(self._L.sbs_gpu_set_fallback(self.h, limit, iterations))

    def fallback_count(self):
        n = ctypes.c_uint64()
        check(self._L.sbs_gpu_fallback_count(self.h, ctypes.byref(n)))
        return n.value

    def set_contact_hash(self, mode="auto", cell_size=0.0):
        check(self._L.sbs_gpu_set_contact_hash(self.h, {"auto": 0, "on": 1, "off": 2}[mode], cell_size))

    def read_contact_hash(self):
        """(cell size, table bits, start, end, entries, global list) of the obstacle hash."""
        h = ctypes.c_double()
        bits, ne, ng = _u32(), _u32(), _u32()
        check(self._L.sbs_gpu_read_contact_hash(self.h, ctypes.byref(h), ctypes.byref(bits), None, None, None, 0,
                                                ctypes.byref(ne), None, ctypes.byref(ng)))
        T = (1 << bits.value) if h.value > 0 else 0
        start, end = np.zeros(T, np.uint32), np.zeros(T, np.uint32)
        entries, glob = np.zeros(max(ne.value, 1), np.uint32), np.zeros(max(ng.value, 1), np.uint32)
        check(self._L.sbs_gpu_read_contact_hash(self.h, None, None, _up(start) if T else None, _up(end) if T else None,
                                                _up(entries), ne.value, ctypes.byref(ne), _up(glob), ctypes.byref(ng)))
        return h.value, bits.value, start, end, entries[:ne.value], glob[:ng.value]

    def set_body_contacts(self, enable=True, cell_size=0.0, thickness=0.0, radius=0.0):
        """Contacts between the bodies' surfaces (before commit; 0 selects the defaults)."""
        check(self._L.sbs_gpu_set_body_contacts(self.h, int(enable), cell_size, thickness, radius))

    def body_contact_setup(self):
        """(cell size, thickness, radius, table bits, triangles [n][4]: outward surface vertices, body)."""
        cell, h, r, bits = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), _u32()
        check(self._L.sbs_gpu_read_body_contact_setup(self.h, ctypes.byref(cell), ctypes.byref(h), ctypes.byref(r),
                                                      ctypes.byref(bits), None))
        tris = np.zeros((self.num_triangles if cell.value > 0 else 0, 4), dtype=np.uint32)
        if len(tris):
            check(self._L.sbs_gpu_read_body_contact_setup(self.h, None, None, None, None, _up(tris)))
        return cell.value, h.value, r.value, bits.value, tris

    def body_contact_candidates(self):
        """Broad-phase (surface vertex, triangle) pairs at the committed positions."""
        n = ctypes.c_uint64()
        check(self._L.sbs_gpu_read_body_contact_candidates(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros((n.value, 2), dtype=np.uint32)
        if n.value:
            check(self._L.sbs_gpu_read_body_contact_candidates(self.h, _up(out), n.value, ctypes.byref(n)))
        return out

    def contact_overflow(self):
        n = ctypes.c_uint64()
        check(self._L.sbs_gpu_contact_overflow(self.h, ctypes.byref(n)))
        return n.value

    def set_timing(self, enable):
        check(self._L.sbs_gpu_set_timing(self.h, int(enable)))

    def reset_stats(self):
        check(self._L.sbs_gpu_reset_stats(self.h))

    def kernel_stats(self, kind):
        launches = ctypes.c_uint64()
        ms = ctypes.c_double()
        units = ctypes.c_uint64()
        check(self._L.sbs_gpu_kernel_stats(self.h, kind, ctypes.byref(launches), ctypes.byref(ms), ctypes.byref(units)))
        return launches.value, ms.value, units.value
