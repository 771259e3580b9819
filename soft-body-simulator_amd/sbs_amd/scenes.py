"""Scene descriptions: the reference demo scene and the benchmark workloads.

A Scene is plain data (numpy arrays and numbers). `Scene.apply(sim)` issues the scene through any
object with the simulation call set of include/sbs_gpu.h (add_body, set_surface, add_green, ...),
so the same description drives the device library and, in tests, the CPU oracle.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Body:
    x0: np.ndarray          # (n, 3) rest positions
    tets: np.ndarray        # (t, 4) body-local tets
    mass: np.ndarray = None  # (n,) or None for unit mass
    surface: np.ndarray = None  # surface vertex -> body vertex
    triangles: np.ndarray = None  # (f, 3) boundary triangles in surface vertex indices
    v0: np.ndarray = None  # (n, 3) initial velocities, or None for rest


@dataclass
class Group:
    kind: str               # "green" | "neohookean" | "distance"
    body: object            # body index, or per-constraint body array (batched group)
    conn: np.ndarray        # tets (t, 4) or (b1, v1, b2, v2) rows
    alpha: float = 0.0
    beta: float = 0.0
    young: object = 1e6     # scalar, or per-constraint array (batched group)
    poisson: object = 0.3
    clusters: np.ndarray = None  # cluster offsets over conn, or None


@dataclass
class Scene:
    name: str
    bodies: list = field(default_factory=list)
    groups: list = field(default_factory=list)
    planes: list = field(default_factory=list)   # (normal, point, vmin, vmax)
    spheres: list = field(default_factory=list)  # (center, radius)
    grids: list = field(default_factory=list)    # grid SDFs: (lo, hi, res, node values, vmin, vmax)
    meshes: list = field(default_factory=list)   # mesh obstacles: (positions, triangles, lo, hi, res)
    gravity: tuple = (0.0, -9.81, 0.0)
    gravity_is_force: bool = True
    collision_compliance: float = 1e-8
    dt: float = 1.0 / 60.0
    iterations: int = 10
    substeps: int = 1
    coloring: str = "hash"  # "hash" (device Jones-Plassmann) or "index" (sequential greedy)
    contact_hash: tuple = None  # (mode "auto" | "on" | "off", cell size) of the contact broad phase
    body_contacts: tuple = None  # (cell size, thickness, radius) of the contacts between bodies (0: defaults)

    @property
    def num_tets(self):
        return sum(len(g.conn) for g in self.groups if g.kind != "distance")

    @property
    def num_vertices(self):
        return sum(len(b.x0) for b in self.bodies)

    def apply(self, sim):
        sim.set_params(gravity=self.gravity, gravity_is_force=self.gravity_is_force,
                       collision_compliance=self.collision_compliance)
        sim.set_coloring(self.coloring)
        for b in self.bodies:
            bi = sim.add_body(b.x0, b.mass)
            if b.surface is not None:
                sim.set_surface(bi, b.surface)
                if b.triangles is not None and hasattr(sim, "set_surface_triangles"):
                    sim.set_surface_triangles(bi, b.triangles)
        for gi, g in enumerate(self.groups):
            if np.ndim(g.body) > 0:
                if g.kind == "green":
                    sim.add_green_batched(g.body, g.conn, g.alpha, g.beta, g.young, g.poisson)
                else:
                    sim.add_neohookean_batched(g.body, g.conn, g.beta, g.young, g.poisson)
            elif g.kind == "green":
                sim.add_green(g.body, g.conn, g.alpha, g.beta, g.young, g.poisson)
            elif g.kind == "neohookean":
                sim.add_neohookean(g.body, g.conn, g.beta, g.young, g.poisson)
            elif g.kind == "distance":
                sim.add_distance(g.conn, g.alpha, g.beta)
            else:
                raise ValueError(g.kind)
            if g.clusters is not None:
                sim.set_clusters(gi, g.clusters)
        for n, p, lo, hi in self.planes:
            sim.add_plane(n, p, lo, hi)
        for c, r in self.spheres:
            sim.add_sphere(c, r)
        for lo, hi, res, nodes, vmin, vmax in self.grids:
            sim.add_grid_sdf(lo, hi, res, nodes, vmin, vmax)
        for pos, tris, lo, hi, res in self.meshes:
            sim.add_mesh_obstacle(pos, tris, lo, hi, res)
        if self.contact_hash is not None and hasattr(sim, "set_contact_hash"):
            sim.set_contact_hash(*self.contact_hash)
        if self.body_contacts is not None:
            sim.set_body_contacts(True, *self.body_contacts)

    def set_initial_state(self, sim):
        """Initial velocities of the bodies that have them (after commit on the device)."""
        for bi, b in enumerate(self.bodies):
            if b.v0 is not None:
                sim.write_body(bi, None, b.v0)

    def bbox_diagonal(self):
        x = np.concatenate([b.x0 for b in self.bodies])
        return float(np.linalg.norm(x.max(0) - x.min(0)))


def build_device(scene, device=0, recolor_passes=2):
    sim = _lib.Simulation(device)
    scene.apply(sim)
    sim.commit(recolor_passes)
    scene.set_initial_state(sim)
    return sim


def _rotation(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def reference_demo():
    """main.cpp:22-124: get_simple_bar_model(4, 4, 12), moved by T(-10, 5, -1) R(pi/2 about
    (0, 1, 0.2)) S(1, 0.8, 2), one green constraint per tet with simulation_parameters_t defaults,
    a floor plane through the origin with normal +y and volume [-20,-5,-20]-[20,5,20], unit masses,
    gravity as a 9.81 N force, dt 0.016, 5 iterations, 1 substep."""
    pos, tets = _lib.bar5(4, 4, 12)
    R = _rotation((0.0, 1.0, 0.2), 3.14159 / 2.0)
    x0 = (R @ (pos * np.array([1.0, 0.8, 2.0])).T).T + np.array([-10.0, 5.0, -1.0])
    s2t, tris = _lib.boundary(len(pos), tets)
    sc = Scene("reference_demo", dt=0.016, iterations=5, substeps=1, gravity_is_force=True)
    sc.bodies.append(Body(x0=x0, tets=tets, surface=s2t, triangles=tris))
    sc.groups.append(Group("green", 0, tets, alpha=1e-4, beta=0.0, young=1e6, poisson=0.30))
    sc.planes.append(((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (-20.0, -5.0, -20.0), (20.0, 5.0, 20.0)))
    return sc


def obstacle_field(num_spheres=4000, nx=24, ny=6, nz=6, edge=0.1, seed=7, contact_hash=("auto", 0.0)):
    """A neo-Hookean beam (hex clusters) among many small static spheres: sphere centres uniform in
    the beam's box grown by 0.1 m, radii U[0.02, 0.06] m, so that many surface vertices start inside
    one or more spheres. The broad phase of the contact detection is what the scene exercises."""
    sc = cantilever(nx, ny, nz, edge, "neohookean", 5e5, 0.45, 1000.0, num_spheres=0, iterations=8, substeps=2,
                    cluster=True, coloring="index", name=f"obstacle_field_{num_spheres}")
    rng = np.random.default_rng(seed)
    x = sc.bodies[0].x0
    lo, hi = x.min(0) - 0.1, x.max(0) + 0.1
    for _ in range(num_spheres):
        sc.spheres.append((tuple(rng.uniform(lo, hi)), float(rng.uniform(0.02, 0.06))))
    sc.contact_hash = contact_hash
    return sc


def hex_clusters(num_tets):
    """Cluster offsets grouping the 6 tets of each hex cell of a box6 mesh."""
    return np.arange(0, num_tets + 1, 6, dtype=np.uint32)


def cantilever(nx, ny, nz, edge=0.1, model="green", young=1e6, poisson=0.45, density=1000.0, alpha=1e-4,
               beta=0.0, ground_y=-1.0, num_spheres=0, seed=0x5B5, dt=1.0 / 60.0, iterations=10, substeps=1,
               name=None, cluster=False, coloring="hash"):
    """Box of nx*ny*nz hex cells of edge `edge` from the origin, 6 tets per cell, lumped masses of
    the given density, the x = 0 face pinned (mass 0), gravity as an acceleration, a ground plane
    at y = ground_y and `num_spheres` static spheres of radius U[0.05, 0.2] m under the beam.
    cluster=True runs the 6 tets of a hex cell as one cluster."""
    pos, tets = _lib.box6(nx, ny, nz, edge)
    mass = _lib.lumped_mass(pos, tets, density)
    mass[pos[:, 0] <= pos[:, 0].min()] = 0.0
    s2t, tris = _lib.boundary(len(pos), tets)
    sc = Scene(name or f"cantilever_{nx}x{ny}x{nz}_{model}", dt=dt, iterations=iterations, substeps=substeps,
               gravity_is_force=False, coloring=coloring)
    sc.bodies.append(Body(x0=pos, tets=tets, mass=mass, surface=s2t, triangles=tris))
    sc.groups.append(Group(model, 0, tets, alpha=alpha, beta=beta, young=young, poisson=poisson,
                           clusters=hex_clusters(len(tets)) if cluster else None))
    sc.planes.append(((0.0, 1.0, 0.0), (0.0, ground_y, 0.0), (-1e30,) * 3, (1e30,) * 3))
    if num_spheres:
        rng = np.random.default_rng(seed)
        L, H, W = nx * edge, ny * edge, nz * edge
        for _ in range(num_spheres):
            r = rng.uniform(0.05, 0.2)
            c = (rng.uniform(0.2 * L, L), rng.uniform(ground_y + r, -0.05), rng.uniform(0.0, W))
            sc.spheres.append((c, r))
    return sc


def beam_540():
    """Plain XPBD cantilever: 10x3x3 cells of 0.1 m, 6 tets each (540 tets, 176 vertices), left
    face pinned, green constraints E=1e6, nu=0.45, density 1000, ground at y=-1, dt 1/60,
    1 substep x 10 iterations."""
    return cantilever(10, 3, 3, 0.1, "green", 1e6, 0.45, 1000.0, iterations=10, substeps=1, name="green_beam_540")


def beam_1m(cluster=True, coloring="index"):
    """Large beam, 1M-tet variant: 220x28x27 cells (997,920 tets, 179,452 vertices), stable
    neo-Hookean E=5e5, nu=0.45, 4 substeps x 8 iterations, ground plane and 16 static spheres.
    By default the 6 tets of a cell form one cluster and clusters are colored in index order
    (8 colors on the lattice)."""
    return cantilever(220, 28, 27, 0.1, "neohookean", 5e5, 0.45, 1000.0, num_spheres=16, iterations=8, substeps=4,
                      name="beam_1M_neohookean", cluster=cluster, coloring=coloring)


def beam_6m(cluster=True, coloring="index"):
    """Large beam: 400x50x50 cells (6.0M tets, 1,043,001 vertices), same material and solver."""
    return cantilever(400, 50, 50, 0.1, "neohookean", 5e5, 0.45, 1000.0, num_spheres=16, iterations=8, substeps=4,
                      name="beam_6M_neohookean", cluster=cluster, coloring=coloring)


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def batch(num_bodies=512, tet_range=(2e3, 2e5), edge=0.05, seed=0xB0D1E5, density=1000.0, iterations=8,
          substeps=4, name=None, large=None, young_range=(1e4, 1e6), poisson_range=(0.3, 0.48), drop_range=(0.5, 3.0)):
    """Batched bodies: `num_bodies` boxes and ellipsoids (alternating) of 6-tet hex cells with tet
    counts log-uniform in tet_range, random rotations, dropped from heights U[drop_range] (U[0.5, 3] m) onto a
    ground plane at y = 0, Young's modulus log-uniform in young_range ([1e4, 1e6]), Poisson ratio
    uniform in poisson_range ([0.3, 0.48]),
    stable neo-Hookean constraints in one batched group with hex-cell clusters (index-order
    coloring), 4 substeps x 8 iterations. large = (count, tets): that many more bodies of about
    `tets` tets each (same shapes and materials), in a row beside the grid of the others, in the
    same group."""
    rng = np.random.default_rng(seed)
    sc = Scene(name or f"batch_{num_bodies}_bodies", gravity_is_force=False, iterations=iterations,
               substeps=substeps, coloring="index")
    lo, hi = np.log(tet_range[0]), np.log(tet_range[1])
    meshes = []
    for b in range(num_bodies):
        target = float(np.exp(rng.uniform(lo, hi)))
        cells = max(target / 6.0, 1.0)
        aspect = np.exp(rng.uniform(np.log(0.5), np.log(2.0), size=3))
        aspect /= np.cbrt(np.prod(aspect))
        if b % 2 == 0:
            n = np.maximum(np.round(aspect * np.cbrt(cells)).astype(int), 1)
            pos, tets = _lib.box6(int(n[0]), int(n[1]), int(n[2]), edge)
            pos = pos - pos.mean(0)
        else:
            r = aspect * np.cbrt(cells * edge ** 3 * 3.0 / (4.0 * np.pi))
            r = np.maximum(r, 1.5 * edge)
            pos, tets = _lib.ellipsoid6(r[0], r[1], r[2], edge)
        meshes.append((pos @ _random_rotation(rng).T, tets))
    extent = max(float(np.ptp(p, axis=0).max()) for p, _ in meshes) + 0.2
    side = int(np.ceil(np.sqrt(num_bodies)))
    if large:
        for k in range(int(large[0])):
            cells = max(float(large[1]) / 6.0, 1.0)
            if k % 2 == 0:
                n = max(int(round(np.cbrt(cells))), 1)
                pos, tets = _lib.box6(n, n, max(int(round(cells / (n * n))), 1), edge)
                pos = pos - pos.mean(0)
            else:
                r = np.full(3, np.cbrt(cells * edge ** 3 * 3.0 / (4.0 * np.pi)))
                pos, tets = _lib.ellipsoid6(r[0], r[1], r[2], edge)
            meshes.append((pos @ _random_rotation(rng).T, tets))
    big = max((float(np.ptp(p, axis=0).max()) for p, _ in meshes[num_bodies:]), default=0.0) + 0.2
    bodies, bconn, young, poisson, clusters = [], [], [], [], [0]
    for b, (pos, tets) in enumerate(meshes):
        gx, gz = b % side, b // side
        drop = rng.uniform(*drop_range)
        corner = np.array([gx * extent, 0.0, gz * extent])
        if b >= num_bodies:  # the large bodies: one row beyond the grid
            corner = np.array([(b - num_bodies) * big, 0.0, side * extent + 0.5 * big])
        x0 = pos + corner + np.array([0.0, drop - pos[:, 1].min(), 0.0])
        mass = _lib.lumped_mass(x0, tets, density)
        s2t, tris = _lib.boundary(len(x0), tets)
        sc.bodies.append(Body(x0=x0, tets=tets, mass=mass, surface=s2t, triangles=tris))
        bodies.append(np.full(len(tets), b, dtype=np.uint32))
        bconn.append(tets)
        E = float(np.exp(rng.uniform(np.log(young_range[0]), np.log(young_range[1]))))
        nu = float(rng.uniform(*poisson_range))
        young.append(np.full(len(tets), E))
        poisson.append(np.full(len(tets), nu))
        clusters.append(clusters[-1] + len(tets))
    conn = np.concatenate(bconn)
    sc.groups.append(Group("neohookean", np.concatenate(bodies), conn, young=np.concatenate(young),
                           poisson=np.concatenate(poisson), clusters=hex_clusters(len(conn))))
    sc.planes.append(((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (-1e30,) * 3, (1e30,) * 3))
    return sc


def stack(edge=0.05, young=2e5, poisson=0.4, iterations=8, substeps=4, gap=0.02, name="stack_4_bodies",
          body_contacts=(0.0, 0.0, 0.0), boxes=3, ellipsoid=True):
    """A few bodies falling onto each other: three boxes of 4 x 4 x 4 hex cells stacked with gaps of
    `gap` m, each shifted sideways by a quarter of its width and turned about y, and an ellipsoid
    dropped onto the top box's corner; stable neo-Hookean constraints in one batched group with hex
    clusters, a ground plane at y = 0, contacts between the bodies (cell size, thickness and radius,
    0 for the defaults)."""
    sc = Scene(name, gravity_is_force=False, iterations=iterations, substeps=substeps, coloring="index",
               body_contacts=body_contacts)
    meshes, y = [], 0.02
    for k in range(boxes):
        pos, tets = _lib.box6(4, 4, 4, edge)
        pos = pos - pos.mean(0)
        pos = pos @ _rotation((0.0, 1.0, 0.0), 0.3 * k).T
        pos = pos + np.array([0.05 * k, y - pos[:, 1].min(), 0.03 * k])
        y = pos[:, 1].max() + gap
        meshes.append((pos, tets))
    if ellipsoid:
        pos, tets = _lib.ellipsoid6(0.12, 0.08, 0.1, edge)
        pos = pos - pos.mean(0) + np.array([0.15, y + 0.05 - (pos[:, 1].min() - pos[:, 1].mean()), 0.1])
        meshes.append((pos, tets))
    bodies, bconn, clusters = [], [], [0]
    for b, (x0, tets) in enumerate(meshes):
        mass = _lib.lumped_mass(x0, tets, 1000.0)
        s2t, tris = _lib.boundary(len(x0), tets)
        sc.bodies.append(Body(x0=x0, tets=tets, mass=mass, surface=s2t, triangles=tris))
        bodies.append(np.full(len(tets), b, dtype=np.uint32))
        bconn.append(tets)
    conn = np.concatenate(bconn)
    sc.groups.append(Group("neohookean", np.concatenate(bodies), conn, young=np.full(len(conn), young),
                           poisson=np.full(len(conn), poisson), clusters=hex_clusters(len(conn))))
    sc.planes.append(((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (-1e30,) * 3, (1e30,) * 3))
    return sc


def cells6(mask, edge):
    """Hex cells of edge `edge` where mask[i, j, k] is set (cell (i, j, k) spans [i, i + 1] x [j, j + 1]
    x [k, k + 1] edges), each split into the 6 tets of sbs_mesh_box6 (the same monotone lattice paths
    and vertex order, so neighbouring cells share their faces' diagonals); the used lattice corners
    numbered in lattice order."""
    mask = np.asarray(mask, dtype=bool)
    nx, ny, nz = mask.shape
    vy, vz = ny + 1, nz + 1
    order = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
    tets = []
    for i, j, k in zip(*np.nonzero(mask)):
        for p, axes in enumerate(order):
            c = [int(i), int(j), int(k)]
            v = [(c[0] * vy + c[1]) * vz + c[2]]
            for a in axes:
                c[a] += 1
                v.append((c[0] * vy + c[1]) * vz + c[2])
            if p in (1, 2, 5):
                v[1], v[2] = v[2], v[1]
            tets.append(v)
    tets = np.asarray(tets, dtype=np.int64)
    used = np.unique(tets)
    remap = np.full((nx + 1) * vy * vz, -1, dtype=np.int64)
    remap[used] = np.arange(len(used))
    li, lj, lk = np.unravel_index(used, (nx + 1, vy, vz))
    pos = np.stack([li, lj, lk], 1).astype(np.float64) * edge
    return pos, remap[tets].astype(np.uint32)


def channel(edge=0.05, young=5e4, poisson=0.4, iterations=16, substeps=8, length=6, gap=2, beta=1e-3,
            name="channel_self_contact", body_contacts=(0.0, 0.0, 0.0)):
    """One body folding onto itself: a channel lying on its side on the ground plane, its upper arm
    `length` x 1 x 3 hex cells and its lower arm 3 cells longer, joined by a wall of `gap` cells at
    x = 0, the arms `gap` cells apart; the upper arm sags under gravity until its free end rests on
    the lower arm (a contact of the body's surface with its own non-incident triangles). Stable
    neo-Hookean constraints with damping `beta` and hex clusters."""
    sc = Scene(name, gravity_is_force=False, iterations=iterations, substeps=substeps, coloring="index",
               body_contacts=body_contacts)
    mask = np.zeros((length + 3, gap + 2, 3), dtype=bool)
    mask[:, 0, :] = True
    mask[:length, gap + 1, :] = True
    mask[0, :, :] = True
    x0, tets = cells6(mask, edge)
    x0 = x0 + np.array([0.0, 0.001, 0.0])
    mass = _lib.lumped_mass(x0, tets, 1000.0)
    s2t, tris = _lib.boundary(len(x0), tets)
    sc.bodies.append(Body(x0=x0, tets=tets, mass=mass, surface=s2t, triangles=tris))
    sc.groups.append(Group("neohookean", 0, tets, beta=beta, young=young, poisson=poisson,
                           clusters=hex_clusters(len(tets))))
    sc.planes.append(((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (-1e30,) * 3, (1e30,) * 3))
    return sc


def tower(edge=0.05, young=2e5, poisson=0.4, iterations=16, substeps=8, gap=0.01, beta=1e-3, ellipsoid=True,
          name="tower_4_bodies", body_contacts=(0.0, 0.0, 0.0), big_block=False):
    """A few bodies resting on each other: a slab of 8 x 2 x 8 hex cells on the ground plane, a box of
    4 x 3 x 4 cells turned 0.3 rad about y on the slab, a box of 2 x 2 x 2 cells turned 0.6 rad on
    that box, and an ellipsoid of radii (0.1, 0.06, 0.08) m on the slab beside them, each starting
    `gap` m above its support; every body's footprint lies inside its support's top face (without
    friction an overhanging soft body slides off). Stable neo-Hookean constraints with damping `beta`
    in one batched group with hex clusters, contacts between the bodies. big_block: one more body, a
    single hex cell of edge 0.3 m on the ground beside the slab, whose triangles are larger than
    the contact hash's cells."""
    sc = Scene(name, gravity_is_force=False, iterations=iterations, substeps=substeps, coloring="index",
               body_contacts=body_contacts)
    pos, tets = _lib.box6(8, 2, 8, edge)
    meshes = [(pos - np.array([4 * edge, -0.001, 4 * edge]), tets)]
    top = meshes[0][0][:, 1].max()
    for (nx, ny, nz), angle in (((4, 3, 4), 0.3), ((2, 2, 2), 0.6)):
        pos, tets = _lib.box6(nx, ny, nz, edge)
        pos = (pos - pos.mean(0)) @ _rotation((0.0, 1.0, 0.0), angle).T
        pos = pos + np.array([-0.08, top + gap - pos[:, 1].min(), -0.08])
        meshes.append((pos, tets))
        top = pos[:, 1].max()
    if ellipsoid:
        pos, tets = _lib.ellipsoid6(0.1, 0.06, 0.08, edge)
        pos = pos - pos.mean(0)
        meshes.append((pos + np.array([0.1, meshes[0][0][:, 1].max() + gap - pos[:, 1].min(), 0.1]), tets))
    if big_block:
        pos, tets = _lib.box6(1, 1, 1, 0.3)
        meshes.append((pos + np.array([0.25, 0.001, -0.15]), tets))
    bodies, bconn = [], []
    for b, (x0, tets) in enumerate(meshes):
        mass = _lib.lumped_mass(x0, tets, 1000.0)
        s2t, tris = _lib.boundary(len(x0), tets)
        sc.bodies.append(Body(x0=x0, tets=tets, mass=mass, surface=s2t, triangles=tris))
        bodies.append(np.full(len(tets), b, dtype=np.uint32))
        bconn.append(tets)
    conn = np.concatenate(bconn)
    sc.groups.append(Group("neohookean", np.concatenate(bodies), conn, beta=beta, young=np.full(len(conn), young),
                           poisson=np.full(len(conn), poisson), clusters=hex_clusters(len(conn))))
    sc.planes.append(((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (-1e30,) * 3, (1e30,) * 3))
    return sc


def box_on_box(edge=0.05, young=2e5, poisson=0.4, iterations=8, substeps=4, gap=0.01, drop=0.0, name="box_on_box",
               body_contacts=(0.0, 0.0, 0.0)):
    """A box of 3 x 3 x 3 hex cells resting `gap` m above the middle of a wider box of 6 x 2 x 6
    cells that lies on the ground plane, moving down at the speed of a fall from `drop` m; stable
    neo-Hookean constraints in one batched group with hex clusters, contacts between the bodies."""
    sc = Scene(name, gravity_is_force=False, iterations=iterations, substeps=substeps, coloring="index",
               body_contacts=body_contacts)
    lo, lt = _lib.box6(6, 2, 6, edge)
    lo = lo - np.array([3 * edge, 0.0, 3 * edge]) + np.array([0.0, 0.001, 0.0])
    up, ut = _lib.box6(3, 3, 3, edge)
    up = up - np.array([1.5 * edge, 0.0, 1.5 * edge]) + np.array([0.0, lo[:, 1].max() + gap, 0.0])
    bconn, bodies = [], []
    for b, (x0, tets) in enumerate(((lo, lt), (up, ut))):
        mass = _lib.lumped_mass(x0, tets, 1000.0)
        s2t, tris = _lib.boundary(len(x0), tets)
        v0 = None
        if b == 1 and drop > 0.0:
            v0 = np.tile([0.0, -np.sqrt(2 * 9.81 * drop), 0.0], (len(x0), 1))
        sc.bodies.append(Body(x0=x0, tets=tets, mass=mass, surface=s2t, triangles=tris, v0=v0))
        bodies.append(np.full(len(tets), b, dtype=np.uint32))
        bconn.append(tets)
    conn = np.concatenate(bconn)
    sc.groups.append(Group("neohookean", np.concatenate(bodies), conn, young=np.full(len(conn), young),
                           poisson=np.full(len(conn), poisson), clusters=hex_clusters(len(conn))))
    sc.planes.append(((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (-1e30,) * 3, (1e30,) * 3))
    return sc


def batch_512(num_bodies=512):
    """512 batched bodies (about 22M tets)."""
    return batch(num_bodies, name=f"batch_{num_bodies}_bodies")


def batch_512_contacts(num_bodies=512, **kw):
    """The 512-body batch with contacts between bodies and of each body with its own surface (cell
    2 mean edges, thickness 0.1 and radius 1 mean edge). With the batch's stiffer bodies (Young's
    modulus up to 1e6) 8 iterations that start every substep from zero multipliers do not hold the
    rest shape: their surfaces fold onto themselves in free fall, the own-surface search's densest
    case. kw: batch()'s other arguments."""
    sc = batch(num_bodies, name=kw.pop("name", f"batch_{num_bodies}_bodies_with_body_contacts"), **kw)
    sc.body_contacts = (0.0, 0.0, 0.0)
    return sc


def batch_512_large_contacts(num_bodies=512, large=(8, 5e5)):
    """The combined configuration with self-collision: batch_512_large with contacts between bodies
    and of each body with its own surface."""
    sc = batch_512_large(num_bodies, large)
    sc.name += "_with_body_contacts"
    sc.body_contacts = (0.0, 0.0, 0.0)
    return sc


def batch_512_large(num_bodies=512, large=(8, 5e5)):
    """The 512-body batch plus 8 bodies of about 500k tets in the same batched group
    (about 26M tets); the learned warm start serves all of them with one network."""
    return batch(num_bodies, large=large, name=f"batch_{num_bodies}_plus_{int(large[0])}_large")
