"""ctypes bindings to the FP64 oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is a CPU restatement of the reference XPBD step and is used
as the checker of the GPU path (tests/, __graft_entry__.smoke()) and as the timed CPU baseline of
bench.py. Nothing in the product imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SBS_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

_lib = None

_d = ctypes.c_double
_u32 = ctypes.c_uint32
_i = ctypes.c_int
_p = ctypes.c_void_p
_pd = ctypes.POINTER(ctypes.c_double)
_pu = ctypes.POINTER(ctypes.c_uint32)
_pu64 = ctypes.POINTER(ctypes.c_uint64)
_pf = ctypes.POINTER(ctypes.c_float)

_SIGS = {
    "orc_create": (_p, []),
    "orc_destroy": (None, [_p]),
    "orc_set_params": (None, [_p, _d, _d, _d, _i, _d, _d]),
    "orc_add_body": (_u32, [_p, _pd, _pd, _u32]),
    "orc_set_body_state": (None, [_p, _u32, _pd, _pd]),
    "orc_set_surface": (None, [_p, _u32, _pu, _u32]),
    "orc_add_green": (_i, [_p, _u32, _pu, _u32, _d, _d, _d, _d]),
    "orc_add_neohookean": (_i, [_p, _u32, _pu, _u32, _d, _d, _d]),
    "orc_add_distance": (_i, [_p, _pu, _u32, _d, _d]),
    "orc_add_green_batched": (_i, [_p, _pu, _pu, _u32, _d, _d, _pd, _pd]),
    "orc_add_neohookean_batched": (_i, [_p, _pu, _pu, _u32, _d, _pd, _pd]),
    "orc_add_plane": (None, [_p, _pd, _pd, _pd, _pd]),
    "orc_add_sphere": (None, [_p, _pd, _d]),
    "orc_set_order": (None, [_p, _i, _u32]),
    "orc_set_threads": (None, [_p, _i]),
    "orc_set_surface_triangles": (None, [_p, _u32, _pu, _u32]),
    "orc_set_body_contacts": (None, [_p, _i, _d, _d, _d, _i]),
    "orc_set_network": (None, [_p, _pf, _pf, _pf, _pf, _i, _i]),
    "orc_network_forward": (None, [_p, _pd, _u32, _pd]),
    "orc_cluster_features": (_u32, [_p, _u32, _d, _pd]),
    "orc_multipliers": (None, [_p, _u32, _pd]),
    "orc_green_trace": (None, [_p, _i]),
    "orc_green_trace_read": (ctypes.c_uint64, [_p, ctypes.c_void_p, ctypes.c_uint64]),
    "orc_set_clusters": (None, [_p, _u32, _pu, _u32]),
    "orc_set_coloring": (None, [_p, _i]),
    "orc_set_warm_start": (None, [_p, _i, _d, _d]),
    "orc_group_colors": (_u32, [_p, _u32, _pu]),
    "orc_step": (None, [_p, _d, _u32, _u32]),
    "orc_read_body": (None, [_p, _u32, _pd, _pd]),
    "orc_detect_contacts": (_u32, [_p]),
    "orc_read_contacts": (_u32, [_p, _pu, _pd, _u32]),
    "orc_residual": (None, [_p, _u32, _d, _pd, _u32]),
    "orc_energy": (None, [_p, _pd]),
    "orc_svd3": (None, [_pd, _pd, _pd, _pd]),
    "orc_inverse3": (None, [_pd, _pd, _pd]),
    "orc_green_value": (_d, [_pd, _pd, _d, _d]),
    "orc_green_eval": (_d, [_pd, _pd, _d, _d, _pd]),
    "orc_bar5": (_u32, [_u32, _u32, _u32, _pd, _pu]),
    "orc_box6": (_u32, [_u32, _u32, _u32, _d, _pd, _pd, _pu]),
    "orc_boundary": (None, [_u32, _pu, _u32, _pu, _pu, _pu, _pu, _pu]),
    "orc_greedy_color": (_u32, [_u32, _u32, _pu64, _u32, _pu]),
    "orc_hash32": (_u32, [_u32]),
    "orc_add_grid_sdf": (None, [_p, _pd, _pd, _pu, _pd, _pd, _pd]),
    "orc_evaluate_sdf": (None, [_p, _u32, _pd, _u32, _pd, _pd]),
    "orc_grid_node_count": (ctypes.c_uint64, [_pu]),
    "orc_grid_node_positions": (None, [_pd, _pd, _pu, _pd]),
    "orc_grid_from_mesh": (None, [_pd, _u32, _pu, _u32, _pd, _pd, _pu, _pd]),
    "orc_mesh_signed_distance": (None, [_pd, _u32, _pu, _u32, _pd, _u32, _pd]),
    "orc_environment_domain": (None, [_pd, _u32, _pd, _pd, _pd, _pd]),
}


def build():
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_pd)


def _up(a):
    return None if a is None else a.ctypes.data_as(_pu)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- pure helpers

def svd3(F):
    F = f64(F).reshape(9)
    U = np.zeros(9)
    s = np.zeros(3)
    V = np.zeros(9)
    lib().orc_svd3(_dp(F), _dp(U), _dp(s), _dp(V))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def inverse3(A):
    A = f64(A).reshape(9)
    out = np.zeros(9)
    det = ctypes.c_double()
    lib().orc_inverse3(_dp(A), _dp(out), ctypes.byref(det))
    return out.reshape(3, 3), det.value


def green_value(x0, p, young, poisson):
    return lib().orc_green_value(_dp(f64(x0).reshape(12)), _dp(f64(p).reshape(12)), young, poisson)


def green_eval(x0, p, young, poisson):
    """(C, f) of one green constraint: C = |V0| psi, f (4, 3) = -grad C."""
    f = np.zeros(12)
    C = lib().orc_green_eval(_dp(f64(x0).reshape(12)), _dp(f64(p).reshape(12)), young, poisson, _dp(f))
    return C, f.reshape(4, 3)


def bar5(w, h, d):
    nt = max(w - 1, 0) * max(h - 1, 0) * max(d - 1, 0) * 5
    pos = np.zeros(w * h * d * 3)
    tets = np.zeros(nt * 4, dtype=np.uint32)
    lib().orc_bar5(w, h, d, _dp(pos), _up(tets))
    return pos.reshape(-1, 3), tets.reshape(-1, 4)


def box6(nx, ny, nz, edge, origin=(0.0, 0.0, 0.0)):
    nv = (nx + 1) * (ny + 1) * (nz + 1)
    pos = np.zeros(nv * 3)
    tets = np.zeros(nx * ny * nz * 24, dtype=np.uint32)
    lib().orc_box6(nx, ny, nz, edge, _dp(f64(origin)), _dp(pos), _up(tets))
    return pos.reshape(-1, 3), tets.reshape(-1, 4)


def boundary(nv, tets):
    tets = u32(tets).reshape(-1)
    ns, ntr = _u32(), _u32()
    lib().orc_boundary(nv, _up(tets), len(tets) // 4, None, None, None, ctypes.byref(ns), ctypes.byref(ntr))
    s2t = np.zeros(ns.value, dtype=np.uint32)
    tris = np.zeros(ntr.value * 3, dtype=np.uint32)
    t2f = np.zeros(ntr.value, dtype=np.uint32)
    lib().orc_boundary(nv, _up(tets), len(tets) // 4, _up(s2t), _up(tris), _up(t2f), ctypes.byref(ns),
                       ctypes.byref(ntr))
    return s2t, tris.reshape(-1, 3), t2f


def greedy_color(verts, passes):
    """verts: (n, arity) uint64 vertex keys of each constraint."""
    verts = np.ascontiguousarray(verts, dtype=np.uint64)
    n, arity = verts.shape
    color = np.zeros(n, dtype=np.uint32)
    nc = lib().orc_greedy_color(n, arity, verts.ctypes.data_as(_pu64), passes, _up(color))
    return color, nc


def hash32(x):
    return lib().orc_hash32(x)


# ---------------------------------------------------------------- scene

class Simulation:
    """FP64 oracle scene; call sequence mirrors the GPU C-ABI (include/sbs_gpu.h)."""

    def __init__(self):
        self._L = lib()
        self.h = self._L.orc_create()
        self.bodies = []  # vertex counts

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_destroy(self.h)
            self.h = None

    def set_params(self, gravity=(0.0, -9.81, 0.0), gravity_is_force=True, collision_compliance=1e-8,
                   collision_damping=0.0):
        self._L.orc_set_params(self.h, gravity[0], gravity[1], gravity[2], int(gravity_is_force),
                               collision_compliance, collision_damping)

    def add_body(self, x0, mass=None):
        x0 = f64(x0).reshape(-1)
        m = None if mass is None else f64(mass)
        b = self._L.orc_add_body(self.h, _dp(x0), _dp(m), len(x0) // 3)
        self.bodies.append(len(x0) // 3)
        return b

    def set_body_state(self, body, x=None, v=None):
        x = None if x is None else f64(x).reshape(-1)
        v = None if v is None else f64(v).reshape(-1)
        self._L.orc_set_body_state(self.h, body, _dp(x), _dp(v))

    def write_body(self, body, x=None, v=None):
        self.set_body_state(body, x, v)

    def set_surface(self, body, s2t):
        s2t = u32(s2t)
        self._L.orc_set_surface(self.h, body, _up(s2t), len(s2t))

    def add_green(self, body, tets, alpha, beta, young, poisson):
        t = u32(tets).reshape(-1)
        return self._L.orc_add_green(self.h, body, _up(t), len(t) // 4, alpha, beta, young, poisson)

    def add_neohookean(self, body, tets, beta, young, poisson):
        t = u32(tets).reshape(-1)
        return self._L.orc_add_neohookean(self.h, body, _up(t), len(t) // 4, beta, young, poisson)

    def add_green_batched(self, bodies, tets, alpha, beta, young, poisson):
        b, t = u32(bodies), u32(tets).reshape(-1)
        return self._L.orc_add_green_batched(self.h, _up(b), _up(t), len(b), alpha, beta, _dp(f64(young)),
                                             _dp(f64(poisson)))

    def add_neohookean_batched(self, bodies, tets, beta, young, poisson):
        b, t = u32(bodies), u32(tets).reshape(-1)
        return self._L.orc_add_neohookean_batched(self.h, _up(b), _up(t), len(b), beta, _dp(f64(young)),
                                                  _dp(f64(poisson)))

    def add_distance(self, b1v1b2v2, alpha, beta):
        e = u32(b1v1b2v2).reshape(-1)
        return self._L.orc_add_distance(self.h, _up(e), len(e) // 4, alpha, beta)

    def add_plane(self, normal, point, vmin=(-1e30,) * 3, vmax=(1e30,) * 3):
        self._L.orc_add_plane(self.h, _dp(f64(normal)), _dp(f64(point)), _dp(f64(vmin)), _dp(f64(vmax)))

    def add_sphere(self, center, radius):
        self._L.orc_add_sphere(self.h, _dp(f64(center)), radius)

    def add_grid_sdf(self, lo, hi, res, nodes, vmin=None, vmax=None):
        """Grid SDF obstacle (Discregrid cubic Lagrange grid); volume defaults to the domain."""
        nodes = f64(nodes)
        assert len(nodes) == grid_node_count(res)
        self._L.orc_add_grid_sdf(self.h, _dp(f64(lo)), _dp(f64(hi)), _up(u32(res)), _dp(nodes),
                                 _dp(f64(lo if vmin is None else vmin)), _dp(f64(hi if vmax is None else vmax)))

    def add_mesh_obstacle(self, positions, triangles, lo, hi, res=(10, 10, 10)):
        """environment_body_t(geometry, domain, resolution): extended domain, mesh distance at the
        grid nodes (on the CPU), the extended domain as englobing volume."""
        a, b = environment_domain(positions, lo, hi)
        self.add_grid_sdf(a, b, res, grid_from_mesh(positions, triangles, a, b, res), a, b)

    def evaluate_sdf(self, k, points):
        p = f64(points).reshape(-1)
        n = len(p) // 3
        sd, g = np.zeros(n), np.zeros(3 * n)
        self._L.orc_evaluate_sdf(self.h, k, _dp(p), n, _dp(sd), _dp(g))
        return sd, g.reshape(-1, 3)

    def set_clusters(self, group, offsets):
        off = u32(offsets)
        self._L.orc_set_clusters(self.h, group, _up(off), len(off) - 1)

    def set_coloring(self, order):
        """order: "hash" or "index" (sequential greedy in index order)."""
        self._L.orc_set_coloring(self.h, {"hash": 0, "index": 1}[order])

    def set_warm_start(self, mode, scale=1.0, hydrostatic=None):
        """mode: "none", "previous" (neo-Hookean multipliers carried over) or "explicit" (set to
        -C/alpha~ at the predicted positions), times scale (the hydrostatic multipliers times
        `hydrostatic` when given); their displacement is applied first."""
        self._L.orc_set_warm_start(self.h, {"none": 0, "previous": 1, "explicit": 3}[mode], scale,
                                   scale if hydrostatic is None else hydrostatic)

    def set_learned_warm_start(self, net, scale=1.0, hydrostatic=None):
        """The learned warm start with the network `net` (an sbs_amd Mlp of 39 inputs and 12 outputs:
        its FP32 parameters and normalization are read from the device), times the scales."""
        self.set_network(net.get_params(), *net.get_normalization(), nin=net.nin, nout=net.nout)
        self._L.orc_set_warm_start(self.h, 2, scale, scale if hydrostatic is None else hydrostatic)

    def set_network(self, params, mean, inv_std, out_scale, nin=39, nout=12):
        """The network of the learned warm start from its FP32 parameters and normalization."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        m48, s48 = np.zeros(48, np.float32), np.ones(48, np.float32)
        m48[:len(mean)] = mean
        s48[:len(inv_std)] = inv_std
        o = np.ascontiguousarray(out_scale, dtype=np.float32)
        f = lambda a: a.ctypes.data_as(_pf)  # noqa: E731
        self._keep = (p, m48, s48, o)
        self._L.orc_set_network(self.h, f(p), f(m48), f(s48), f(o), int(nin), int(nout))

    def network_forward(self, x, nout=12):
        """The oracle's network on raw feature rows x [n][39] (test hook)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 39)
        y = np.zeros((len(x), nout))
        self._L.orc_network_forward(self.h, _dp(x), len(x), _dp(y))
        return y

    def cluster_features(self, group, dt, num_clusters):
        """Features [clusters][39] of a clustered neo-Hookean group at the current positions (test hook)."""
        out = np.zeros((num_clusters, 39))
        n = self._L.orc_cluster_features(self.h, group, dt, _dp(out))
        return out[:n]

    def multipliers(self, group, count):
        """[count][4]: (deviatoric, hydrostatic) multipliers and their values one substep before, per
        neo-Hookean constraint of the group (test hook)."""
        out = np.zeros((count, 4))
        self._L.orc_multipliers(self.h, group, _dp(out))
        return out

    def green_trace(self, on=True):
        """Start (or stop) recording the branches of every green projection (test hook)."""
        self._L.orc_green_trace(self.h, int(on))

    def read_green_trace(self):
        n = self._L.orc_green_trace_read(self.h, None, 0)
        out = np.zeros(n, dtype=np.uint8)
        self._L.orc_green_trace_read(self.h, out.ctypes.data, n)
        return out

    def set_surface_triangles(self, body, tris):
        t = u32(tris).reshape(-1)
        self._L.orc_set_surface_triangles(self.h, body, _up(t), len(t) // 3)

    def set_body_contacts(self, enable=True, cell_size=0.0, thickness=0.0, radius=0.0, brute_force=False):
        """Contacts between bodies (see simulation_t::detect_body_contacts); brute_force tests every
        triangle instead of the grid's."""
        self._L.orc_set_body_contacts(self.h, int(enable), cell_size, thickness, radius, int(brute_force))

    def set_order(self, colored, recolor_passes=2):
        self._L.orc_set_order(self.h, int(colored), recolor_passes)

    def set_threads(self, threads):
        """Threads for the colored order's color stages (bit-identical to one thread)."""
        self._L.orc_set_threads(self.h, int(threads))

    def group_colors(self, group, n):
        c = np.zeros(n, dtype=np.uint32)
        nc = self._L.orc_group_colors(self.h, group, _up(c))
        return c, nc

    def step(self, dt, iterations, substeps):
        self._L.orc_step(self.h, dt, iterations, substeps)

    def read_body(self, body):
        n = self.bodies[body]
        x = np.zeros(n * 3)
        v = np.zeros(n * 3)
        self._L.orc_read_body(self.h, body, _dp(x), _dp(v))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def detect_contacts(self):
        n = self._L.orc_detect_contacts(self.h)
        ids = np.zeros(n * 4, dtype=np.uint32)
        qn = np.zeros(n * 6)
        self._L.orc_read_contacts(self.h, _up(ids), _dp(qn), n)
        return ids.reshape(-1, 4), qn.reshape(-1, 6)

    def residual(self, group, dt, ncolors):
        out = np.zeros(max(ncolors, 1))
        self._L.orc_residual(self.h, group, dt, _dp(out), len(out))
        return out

    def energy(self):
        out = np.zeros(6)
        self._L.orc_energy(self.h, _dp(out))
        return out


def grid_node_count(res):
    return int(lib().orc_grid_node_count(_up(u32(res))))


def grid_node_positions(lo, hi, res):
    out = np.zeros(3 * grid_node_count(res))
    lib().orc_grid_node_positions(_dp(f64(lo)), _dp(f64(hi)), _up(u32(res)), _dp(out))
    return out.reshape(-1, 3)


def grid_from_mesh(pos, tris, lo, hi, res):
    p, t = f64(pos).reshape(-1), u32(tris).reshape(-1)
    out = np.zeros(grid_node_count(res))
    lib().orc_grid_from_mesh(_dp(p), len(p) // 3, _up(t), len(t) // 3, _dp(f64(lo)), _dp(f64(hi)), _up(u32(res)),
                             _dp(out))
    return out


def mesh_signed_distance(pos, tris, points):
    p, t, q = f64(pos).reshape(-1), u32(tris).reshape(-1), f64(points).reshape(-1)
    out = np.zeros(len(q) // 3)
    lib().orc_mesh_signed_distance(_dp(p), len(p) // 3, _up(t), len(t) // 3, _dp(q), len(q) // 3, _dp(out))
    return out


def environment_domain(pos, lo, hi):
    p = f64(pos).reshape(-1)
    a, b = np.zeros(3), np.zeros(3)
    lib().orc_environment_domain(_dp(p), len(p) // 3, _dp(f64(lo)), _dp(f64(hi)), _dp(a), _dp(b))
    return a, b
