"""JSON scene files (src/io/load_scene.cpp, include/sbs/io/load_scene.h, include/sbs/common/scene.h).

load_scene(path, environment_body_factory, physics_body_factory) reads the scene format of the
reference: lights, "environment" bodies (triangle meshes) and "objects" (tet meshes with a physics
block), each geometry read from a PLY file relative to the scene file, coloured when the file has no
colours, rescaled into its "box" and translated by its "translation" with the reference's float and
double arithmetic (load_scene.cpp:11-80), then handed to the caller's factories; the nodes the
factories return are tagged environment or physics and collideable or not (load_scene.cpp:243-257,
334-347). to_scene() turns a loaded file into a simulation Scene with default factories: tet bodies
with lumped masses from their mass density, their initial velocity and one Green constraint per tet,
and collideable environment meshes as grid SDF obstacles (environment_body_t(geometry, domain,
resolution), environment_body.cpp:12-78).
"""
import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib, ply
from .scenes import Body, Group, Scene

FLOAT_MAX = float(np.finfo(np.float32).max)


@dataclass
class SceneBodyInfo:
    """scene::scene_body_info (load_scene.h:15-19)."""
    id: str
    geometry: ply.Geometry


@dataclass
class PhysicsBodyInfo(SceneBodyInfo):
    """scene::physics_body_info (load_scene.h:21-29)."""
    mass_density: float = 0.0
    velocity: tuple = (0.0, 0.0, 0.0)


@dataclass
class SceneNode:
    """What load_scene records on a factory's node: id, environment or physics body, collideable."""
    id: str
    body: object
    environment: bool
    collideable: bool


@dataclass
class LoadedScene:
    """common::scene_t (scene.h:11-53): nodes and the two lights (float32 fields)."""
    nodes: list = field(default_factory=list)
    directional_light: dict = field(default_factory=dict)
    point_light: dict = field(default_factory=dict)


def rescale(geometry, bmin, bmax):
    """load_scene.cpp:11-67: affine map of the float positions' bounding box onto [bmin, bmax] per
    axis in double arithmetic, axes thinner than 1e-8 left as they are, rounded back to float."""
    p = geometry.positions.reshape(-1, 3)
    lo = np.full(3, FLOAT_MAX)
    hi = np.full(3, -FLOAT_MAX)
    if len(p):  # comparisons with NaN are false in the reference: NaNs never set a bound
        lo = np.fmin(lo, np.fmin.reduce(p, axis=0).astype(np.float64))
        hi = np.fmax(hi, np.fmax.reduce(p, axis=0).astype(np.float64))
    d = hi - lo
    out = p.astype(np.float64)
    for a in range(3):
        if not abs(d[a]) < 1e-8:
            out[:, a] = bmin[a] + (bmax[a] - bmin[a]) * (out[:, a] - lo[a]) / d[a]
    geometry.positions = out.astype(np.float32).reshape(-1)


def translate(geometry, t):
    """load_scene.cpp:69-81: float positions plus the float-rounded translation, in float."""
    p = geometry.positions.reshape(-1, 3)
    geometry.positions = (p + np.asarray(t, np.float64).astype(np.float32)).astype(np.float32).reshape(-1)


def _color(geometry, spec):
    if geometry.colors.size == 0:
        c = spec["color"]
        rgb = np.array([int(c["r"]), int(c["g"]), int(c["b"])], dtype=np.int64).astype(np.uint8)
        geometry.colors = np.tile(rgb, len(geometry.positions) // 3)


def _transform(geometry, spec):
    if "box" in spec:
        b = spec["box"]
        rescale(geometry, [float(b["min"][k]) for k in "xyz"], [float(b["max"][k]) for k in "xyz"])
    if "translation" in spec:
        t = spec["translation"]
        translate(geometry, [float(t[k]) for k in "xyz"])


def _asset(path, spec):
    asset = os.path.join(os.path.dirname(path), spec["geometry"]["path"])
    if not asset.endswith(".ply") or os.path.basename(asset) == ".ply" or not os.path.isfile(asset):
        return None
    return asset


def _light(spec, keys):
    f = np.float32
    out = {}
    for k in keys:
        v = spec[k]
        out[k] = {kk: f(vv) for kk, vv in v.items()} if isinstance(v, dict) else f(v)
    return out


def load_scene(path, environment_body_factory, physics_body_factory):
    """io::load_scene (load_scene.cpp:83-363). Returns an empty LoadedScene when the path is not an
    existing .json file; bodies whose geometry file is missing, not .ply or unreadable are skipped."""
    path = os.fspath(path)
    if not (os.path.isfile(path) and path.endswith(".json") and os.path.basename(path) != ".json"):
        return LoadedScene()
    with open(path) as f:
        spec = json.load(f)
    scene = LoadedScene()
    lights = spec["lights"]
    d, p = lights["directional"], lights["point"]
    scene.directional_light = {"direction": _light(d["direction"], "xyz"),
                               **_light(d, ("ambient", "diffuse", "specular"))}
    scene.point_light = {"position": _light(p["position"], "xyz"),
                         **_light(p, ("ambient", "diffuse", "specular", "attenuation"))}
    for es in spec["environment"]:
        _ = es["geometry"]["type"]
        asset = _asset(path, es)
        if asset is None:
            continue
        geometry = ply.read_ply(asset)
        if geometry is None:
            continue
        _color(geometry, es)
        _transform(geometry, es)
        info = SceneBodyInfo(str(es["id"]), geometry)
        node = environment_body_factory(info)
        scene.nodes.append(SceneNode(info.id, node, True, bool(es.get("collideable", False))))
    for os_ in spec["objects"]:
        _ = os_["geometry"]["type"]
        asset = _asset(path, os_)
        if asset is None:
            continue
        phys = os_["physics"]
        _ = str(phys["type"])
        density = float(phys["mass"])
        vel = tuple(float(phys["velocity"][k]) for k in "xyz")
        geometry = ply.read_ply(asset)
        if geometry is None:
            continue
        info = PhysicsBodyInfo(str(os_["id"]), geometry, density, vel)
        _color(geometry, os_)
        _transform(geometry, os_)
        node = physics_body_factory(info)
        scene.nodes.append(SceneNode(info.id, node, False, bool(os_.get("collideable", False))))
    return scene


def to_scene(path, young=1e6, poisson=0.3, alpha=1e-4, beta=0.0, resolution=(10, 10, 10), dt=1.0 / 60.0,
             iterations=10, substeps=1, name=None):
    """A simulation Scene from a scene file with the default factories (see the module comment).
    Environment meshes take their own bounding box as the grid domain; physics bodies are tet meshes
    whose boundary surface is their collision surface."""
    loaded = load_scene(path, lambda info: info, lambda info: info)
    sc = Scene(name or os.path.splitext(os.path.basename(os.fspath(path)))[0], dt=dt, iterations=iterations,
               substeps=substeps, gravity=(0.0, -9.81, 0.0), gravity_is_force=False)
    for node in loaded.nodes:
        g = node.body.geometry
        x = g.positions.astype(np.float64).reshape(-1, 3)
        if node.environment:
            if node.collideable and g.is_triangle_mesh() and g.indices.size:
                tris = g.indices.reshape(-1, 3).astype(np.uint32)
                sc.meshes.append((x, tris, x.min(axis=0), x.max(axis=0), tuple(resolution)))
            continue
        if not g.is_tetrahedral_mesh():
            continue
        tets = g.indices.reshape(-1, 4).astype(np.uint32)
        mass = _lib.lumped_mass(x, tets, node.body.mass_density)
        s2t, tris = _lib.boundary(len(x), tets)
        b = len(sc.bodies)
        sc.bodies.append(Body(x0=x, tets=tets, mass=mass, surface=s2t if node.collideable else None,
                              triangles=tris if node.collideable else None,
                              v0=np.tile(np.asarray(node.body.velocity, np.float64), (len(x), 1))))
        sc.groups.append(Group("green", b, tets, alpha=alpha, beta=beta, young=young, poisson=poisson))
    return sc, loaded
