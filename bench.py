#!/usr/bin/env python3
"""Benchmark of the MI355X XPBD time step on the 1M-tet cantilever beam.

One "step" is one timestep_t::step() call (one frame = `substeps` substeps of `iterations`
Gauss-Seidel sweeps over the 997,920 stable neo-Hookean tets, contacts against the ground plane and
16 spheres). Timed regions replay the frame as a hipGraph; inputs are resident in HBM.

The metric compares warm-started solvers against the plain solver at equal residual, per state.
After `--warmup` frames the tuner takes the next `--tune-frames` frames as its horizon: a
warm-start candidate (previous multipliers, then a network trained on the device from captured
episodes; iterations and a pair of multiplier scales) qualifies when on every frame its FP64 frame
residual is at most the plain solver's from the same state. The plain solver then carries the state
to the start of the timed horizon (`--steps` frames, disjoint from the tuning frames); every leg is
timed there from one snapshot (`--reps` repeats, median) with the per-frame fallback check on, and
every warm-started leg is compared with the plain solver per state over the timed frames again; a
leg with a frame above the plain residual is timed and compared again with one more iteration (and
the scale the tuner found best for that count on the tuning frames).
`value` is the learned leg when all its timed frames reach the plain residual, else the plain
solver's.

After the timed regions an instrumented region (HIP events around every launch on the library's
stream) measures the average duration of the tet projection kernel for the roofline figure. The
CPU baseline times the FP64 oracle in the reference's insertion order on one pinned core over a
bounded sample (5 frames after 5 warm-up frames of the same scene, or of a 16-body batch scaled by
tet count for the batches); the accuracy legs compare the device with the oracle from rest.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "soft-body-simulator_amd"))

METRIC = "substeps/s at 1M tets to equal constraint residual, learned+batched vs plain"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one tet projection: indices 16 + DmInv 36 + V0 4 + per-tet material 8
# + multiplier read/write 8 + 4 gathers of (position 12 + inverse mass 4) + 4 scatters of 12.
BYTES_PER_TET = 184


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def copy_calibration(nbytes=256 << 20, reps=4):
    """Device-to-device copies of a known size (the runtime's copy kernel), used to scale the
    FETCH_SIZE / WRITE_SIZE counters of a --pmc profile of this command."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {reps} device copies of {nbytes} bytes")


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def cpu_baseline(scene, sample=None, warmup_frames=5, frames=5):
    """The FP64 oracle in the reference's insertion order, one thread pinned to one host core
    (os.sched_setaffinity of the calling thread, restored afterwards): `warmup_frames` untimed
    frames, then `frames` frames timed one by one; the median frame gives substeps/s (of the same
    scene, or of a sample scene scaled to the full scene by tet count)."""
    O = _oracle()
    run = sample or scene
    t0 = time.perf_counter()
    sim = O.Simulation()
    run.apply(sim)
    sim.set_order(False)
    setup = time.perf_counter() - t0
    mask = os.sched_getaffinity(0)
    core = min(mask)
    os.sched_setaffinity(0, {core})
    try:
        log(f"cpu baseline: {run.name} built in {setup:.1f}s; {warmup_frames} warm-up and {frames} timed frames "
            f"of {run.substeps} substeps on core {core}")
        for _ in range(warmup_frames):
            sim.step(run.dt, run.iterations, run.substeps)
        ts = []
        for _ in range(frames):
            t0 = time.perf_counter()
            sim.step(run.dt, run.iterations, run.substeps)
            ts.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, mask)
    med = sorted(ts)[len(ts) // 2]
    scale = scene.num_tets / run.num_tets
    what = (f"median of {frames} frames (after {warmup_frames} warm-up frames) of {run.substeps} substeps x "
            f"{run.iterations} iterations of {run.name} ({run.num_tets} tets)"
            + (f", scaled by tet count to {scene.num_tets} tets" if sample else ""))
    return {"value": run.substeps / (med * scale), "unit": "substeps/s", "cores": 1, "kind": "port",
            "sample": f"{what}; FP64 oracle in reference insertion order, one thread pinned to core {core}",
            "frame_seconds": [round(t, 4) for t in ts], "host_cpu": cpu_model(), "nproc": os.cpu_count(),
            "affinity_cores": len(mask),
            "projections_per_s": run.num_tets * run.iterations * run.substeps / med}


def oracle_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def accuracy_legs(scene, legs, launch, frames=60, check=(1, 60)):
    """Position error of the device against the FP64 oracle, relative to the bounding-box diagonal,
    after the frames in `check`, from rest. legs: a list of {"name", "iterations", "mode", "scale",
    "net"} (mode "none" for the plain solver; no fallback, so the warm start's own trajectory is
    measured). Each device leg is compared with the oracle running the same solver (colored order,
    the same warm start), and each warm-started device leg also with the plain oracle leg, the
    solution the warm starts are measured against ("<name>_vs_plain"). The oracle runs the colored order's stages on
    oracle_threads() host threads (bit-identical to one thread)."""
    import numpy as np

    from sbs_amd import autotune, scenes

    O = _oracle()
    diag = scene.bbox_diagonal()
    dev_x, orc_x = {}, {}

    def positions(sim):
        return np.concatenate([sim.read_body(b)[0] for b in range(len(scene.bodies))])

    for leg in legs:
        name, k, mode = leg["name"], leg["iterations"], leg["mode"]
        dev = scenes.build_device(scene)
        dev.set_launch_config(launch["block"], launch["cluster_block"])
        dev.use_graphs(True)
        if mode != "none":
            dev.set_warm_start(mode, leg.get("net"))
            dev.set_warm_start_scale(*autotune._scales(leg["scale"]))
        xs = {}
        for f in range(1, frames + 1):
            dev.step(scene.dt, k, scene.substeps)
            if f in check:
                xs[f] = positions(dev)
        dev_x[name] = xs
        dev.close()
        if mode == "learned" and not hasattr(O.Simulation, "set_learned_warm_start"):
            continue
        t0 = time.perf_counter()
        orc = O.Simulation()
        scene.apply(orc)
        orc.set_order(True, 2)
        scene.set_initial_state(orc)
        orc.set_threads(oracle_threads())
        if mode == "learned":
            orc.set_learned_warm_start(leg["net"], *autotune._scales(leg["scale"]))
        elif mode != "none":
            orc.set_warm_start(mode, *autotune._scales(leg["scale"]))
        xs = {}
        for f in range(1, frames + 1):
            orc.step(scene.dt, k, scene.substeps)
            if f in check:
                xs[f] = positions(orc)
        orc_x[name] = xs
        del orc
        log(f"accuracy: {name} ({k} iterations) on the device and in the oracle, {frames} frames; oracle "
            f"{time.perf_counter() - t0:.1f}s on {oracle_threads()} threads")

    def err(a, b):
        e = float(np.abs(a - b).max()) / diag
        return e if np.isfinite(e) else None

    out = {}
    for name in dev_x:
        if name in orc_x:
            out[name] = {f"frame_{f}": err(dev_x[name][f], orc_x[name][f]) for f in check}
        if name != "plain" and "plain" in orc_x:
            out[f"{name}_vs_plain"] = {f"frame_{f}": err(dev_x[name][f], orc_x["plain"][f]) for f in check}
    return out


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def row_boundary(args):
    """Boundary surface extraction of the 6M-tet beam (400 x 50 x 50 cells, 6-split) on the device,
    per call including the upload of the tets and the download of the surface; beside it the
    library's host extraction and the oracle (reference algorithm) on the same mesh."""
    import numpy as np

    import sbs_amd as S

    pos, tets = S.box6(400, 50, 50, 0.1)
    s2t, tris = S.extract_boundary(len(pos), tets)
    dev = _median_time(lambda: S.extract_boundary(len(pos), tets), 5)
    t0 = time.perf_counter()
    hs, ht = S.boundary(len(pos), tets)
    host = time.perf_counter() - t0
    assert np.array_equal(s2t, hs) and np.array_equal(tris, ht)
    log(f"boundary: device {1e3 * dev:.1f} ms, host library {1e3 * host:.0f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        O.boundary(len(pos), tets)
        el = time.perf_counter() - t0
        cpu = {"value": round(len(tets) / el, 1), "unit": "tets/s", "cores": 1, "kind": "port",
               "sample": f"the whole 6M-tet mesh, oracle boundary extraction (ordered maps), {el:.1f}s on {cpu_model()}"}
    return {"metric": "boundary surface extraction, 6M-tet beam", "value": round(len(tets) / dev, 1), "unit": "tets/s",
            "higher_is_better": True, "dtype": "u32", "data": "synthetic",
            "config": {"workload": "box6 400x50x50", "tets": int(len(tets)), "surface_vertices": int(len(s2t)),
                       "triangles": int(len(tris))},
            "ms_per_call": round(1e3 * dev, 3), "host_library_ms": round(1e3 * host, 1), "cpu_baseline": cpu}


def icosphere(r, levels):
    import numpy as np

    t = (1 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(levels):
        mid, nf = {}, []

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return r * np.array(v), np.array(f, dtype=np.uint32)


def row_grid_sdf(args):
    """Grid SDF of a triangle mesh: the signed distance of an 81,920-triangle sphere mesh at the
    245,025 nodes of a 32^3-cell cubic Lagrange grid, on the device (per call, with the host's
    pseudonormals and the transfers); the oracle times a 4^3-cell grid of the same mesh."""
    import sbs_amd as S

    V, F = icosphere(0.5, 6)
    lo, hi, res = [-0.6] * 3, [0.6] * 3, [32, 32, 32]
    nodes = S.grid_node_count(res)
    S.grid_from_mesh(V, F, lo, hi, res)
    dev = _median_time(lambda: S.grid_from_mesh(V, F, lo, hi, res), 3)
    pairs = nodes * len(F)
    log(f"grid sdf: {nodes} nodes x {len(F)} triangles in {1e3 * dev:.1f} ms ({pairs / dev:.3g} pairs/s)")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        small = [4, 4, 4]
        t0 = time.perf_counter()
        O.grid_from_mesh(V, F, lo, hi, small)
        el = time.perf_counter() - t0
        n_small = O.grid_node_count(small)
        cpu = {"value": round(n_small * len(F) / el, 1), "unit": "node-triangle pairs/s", "cores": 1, "kind": "port",
               "sample": f"{n_small} nodes (4^3 cells) x {len(F)} triangles, oracle, {el:.1f}s on {cpu_model()}"}
    return {"metric": "grid SDF construction from a triangle mesh", "value": round(pairs / dev, 1),
            "unit": "node-triangle pairs/s", "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "icosphere 81920 triangles, 32^3 cells", "nodes": nodes, "triangles": int(len(F))},
            "ms_per_call": round(1e3 * dev, 3), "cpu_baseline": cpu}


BEAM_SIZES = ((110, 14, 14), (160, 20, 20), (220, 28, 27), (280, 35, 34), (350, 44, 43), (400, 50, 50),
              (500, 63, 62))


def row_beam_sizes(args):
    """The plain solver's projection kernel on beams of the 1M beam's shape from 130k to 11.7M tets
    (same material, contacts, 4 substeps x 8 iterations): its average launch and HBM roofline
    fraction against the clusters per launch, i.e. the waves each SIMD gets per color."""
    import sbs_amd as S
    from sbs_amd import scenes

    rows = []
    for nx, ny, nz in BEAM_SIZES:
        sc = scenes.cantilever(nx, ny, nz, 0.1, "neohookean", 5e5, 0.45, 1000.0, num_spheres=16, iterations=8,
                               substeps=4, name=f"beam_{nx}x{ny}x{nz}_neohookean", cluster=True, coloring="index")
        r = plain_throughput(sc, S._lib.KERNEL_NEOHOOKEAN, args)
        per_launch = sc.num_tets / 8
        r["waves_per_simd_per_launch"] = round(per_launch / 6 / 64 / 1024, 3)
        rows.append(r)
        log(f"{sc.name}: {sc.num_tets} tets, {r['waves_per_simd_per_launch']} waves per SIMD per launch, "
            f"{r['projection_avg_launch_us']} us per launch, {r['roofline_frac']} of HBM")
    return {"metric": "projection kernel roofline fraction against beam size", "unit": "fraction of 8 TB/s",
            "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "beams of 220x28x27 cells' proportions", "sizes": [list(s) for s in BEAM_SIZES]},
            "rows": rows}


def row_ply(args):
    """ASCII PLY writing and reading of the 1M-tet beam (179,452 vertices, 997,920 tets) from and
    to memory; beside it the reference's own ply.cpp (built from its sources as a checker, when
    present)."""
    import numpy as np

    import sbs_amd as S
    from sbs_amd import ply

    pos, tets = S.box6(220, 28, 27, 0.1)
    g = ply.tet_geometry(pos, tets)
    data = ply.write_ply_bytes(g)
    w = _median_time(lambda: ply.write_ply_bytes(g), 3)
    r = _median_time(lambda: ply.read_ply_bytes(data), 3)
    mb = len(data) / 1e6
    log(f"ply: {mb:.1f} MB, write {1e3 * w:.0f} ms, read {1e3 * r:.0f} ms")
    cpu = None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_ply
    if ref_ply.available() and not args.no_cpu_baseline:
        rw = _median_time(lambda: ref_ply.write(g.positions, g.indices, geometry_type=1, fmt=0), 1)
        rr = _median_time(lambda: ref_ply.read(data), 1)
        assert ref_ply.write(g.positions, g.indices, geometry_type=1, fmt=0) == data
        cpu = {"value": round(mb / rr, 1), "unit": "MB/s read", "write_MB_per_s": round(mb / rw, 1), "cores": 1,
               "kind": "reference", "sample": f"the same {mb:.1f} MB file, the reference's ply.cpp, on {cpu_model()}"}
    return {"metric": "ASCII PLY read throughput, 1M-tet beam", "value": round(mb / r, 1), "unit": "MB/s",
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "box6 220x28x27 tet mesh", "bytes": len(data)},
            "write_MB_per_s": round(mb / w, 1), "threads": min(16, os.cpu_count() or 1), "cpu_baseline": cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed frames")
    ap.add_argument("--warmup", type=int, default=5, help="untimed frames")
    ap.add_argument("--profile-steps", type=int, default=3, help="frames of the instrumented region")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="beam_1m",
                    choices=["beam_1m", "beam_6m", "beam_540", "batch_512", "batch_512_large", "batch_512_contacts",
                             "batch_512_large_contacts"])
    ap.add_argument("--no-learned", action="store_true", help="plain solver only")
    ap.add_argument("--tune-frames", type=int, default=16, help="frames after warm-up the warm starts are tuned on")
    ap.add_argument("--reps", type=int, default=5, help="repeats of the timed frames (median)")
    ap.add_argument("--accuracy-frames", type=int, default=60,
                    help="frames of the accuracy legs from rest (0: none); errors at frame 1 and the last")
    ap.add_argument("--train-frames", type=int, default=None, help="captured frames of the network's training data")
    ap.add_argument("--also", nargs="*", default=["beam_6m", "batch_512", "batch_512_contacts"],
                    choices=["beam_1m", "beam_6m", "beam_540", "batch_512", "batch_512_large", "batch_512_contacts",
                             "batch_512_large_contacts"],
                    help="further workloads measured with the plain solver (no tuning)")
    ap.add_argument("--no-graphs", action="store_true", help="stream launches instead of hipGraph replays")
    ap.add_argument("--no-contact-fusion", action="store_true",
                    help="the first iteration's contacts in their own launch, not in the warm-start gather")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run device copies of known size (to scale PMC byte counters)")
    ap.add_argument("--profile-frames", type=int, default=0,
                    help="profiler run: this many plain and learned frames without tuning or timing (see profile_run)")
    ap.add_argument("--warm-config", default="4:0.7,0.65",
                    help="the previous-multiplier start the profiler run trains its network on: iterations:scales")
    ap.add_argument("--row", default="step", choices=["step", "boundary", "grid_sdf", "ply", "beam_sizes"],
                    help="step: the time step (the headline metric); the others measure the surface, grid SDF "
                         "and PLY components on their own workloads")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the time step runs on a single MI355X; sharding the bodies or the "
              "beam across GPUs is not implemented, so --gpus must be 1.", file=sys.stderr)
        sys.exit(2)
    if args.row != "step":
        out = {"boundary": row_boundary, "grid_sdf": row_grid_sdf, "ply": row_ply,
               "beam_sizes": row_beam_sizes}[args.row](args)
        out.update({"n_gpus": 1})
        print(json.dumps(out), flush=True)
        return

    import numpy as np

    import sbs_amd as S
    from sbs_amd import autotune, scenes

    if args.copy_calibration:
        copy_calibration()

    make_scene = {"beam_1m": scenes.beam_1m, "beam_6m": scenes.beam_6m, "beam_540": scenes.beam_540,
                  "batch_512": scenes.batch_512, "batch_512_large": scenes.batch_512_large,
                  "batch_512_contacts": scenes.batch_512_contacts,
                  "batch_512_large_contacts": scenes.batch_512_large_contacts}
    scene = make_scene[args.workload]()
    log(f"workload {scene.name}: {scene.num_vertices} vertices, {scene.num_tets} tets")
    if args.profile_frames:
        print(json.dumps(profile_run(args, scene)), flush=True)
        return
    t0 = time.perf_counter()
    sim = scenes.build_device(scene)
    _, ncolors = sim.group_colors(0)
    log(f"committed in {time.perf_counter() - t0:.1f}s, {ncolors} colors")
    kind = {"green": S._lib.KERNEL_GREEN, "neohookean": S._lib.KERNEL_NEOHOOKEAN}[scene.groups[0].kind]
    warmable = scene.groups[0].kind == "neohookean" and not args.no_learned
    if warmable and scene.body_contacts is not None:
        # a warm-start trial that diverges stretches the surfaces over so many cells that every
        # triangle goes to the large list, which every vertex then tests (a frame takes minutes)
        warmable = False
        log("no warm-start legs on a scene with body contacts: the plain solver only")

    sim.use_graphs(not args.no_graphs)
    sim.set_contact_fusion(not args.no_contact_fusion)
    launch = autotune.tune_launch(sim, scene)
    log(f"launch config {launch}")
    for _ in range(args.warmup):
        sim.step(scene.dt, scene.iterations, scene.substeps)
    sim.synchronize()

    # Tuning horizon: the tune_frames frames after warm-up (state S0). The plain solver's frame
    # residuals over them screen the warm-start candidates, which then qualify per state on them.
    sim.snapshot(0)
    tune = args.tune_frames
    plain_tune = autotune.frame_residuals(sim, scene, tune, scene.iterations)
    sim.restore(0)
    limit = 2.0 * float(plain_tune.max())
    log(f"warm-up done ({args.warmup} frames); plain {scene.iterations}-iteration frame residual over the {tune} "
        f"tuning frames: mean {plain_tune.mean():.4g}, largest {plain_tune.max():.4g}")
    found, trials, nets, train_info = {}, [], {}, None
    if warmable:
        f_prev, t_prev = autotune.tune_warm_start(sim, scene, None, float(plain_tune.mean()), frames=tune,
                                                  modes=("previous",), all_iterations=True, log=log)
        found.update(f_prev)
        trials += t_prev
        # one network per iteration count from the previous start's count up: each is trained on
        # trajectories of the iteration count it runs with (all from the tuning start)
        train_cfg = autotune.training_config(args, found.get("previous"))
        k0 = int(train_cfg.get("iterations", 4))
        train_info = {}
        for k in range(k0, scene.iterations):
            train_cfg["iterations"] = k
            nets[k], train_info[k] = autotune.train_warm_start(sim, scene, log=log, **train_cfg)
        f_l, t_l = autotune.tune_warm_start(sim, scene, nets, float(plain_tune.mean()), frames=tune,
                                            iterations=tuple(range(k0, scene.iterations)), modes=("learned",),
                                            all_iterations=True, log=log)
        found.update(f_l)
        trials += t_l
    # the timed horizon starts where the plain solver leaves the tuning horizon (state S1)
    sim.restore(0)
    sim.set_warm_start("none")
    for _ in range(tune):
        sim.step(scene.dt, scene.iterations, scene.substeps)
    sim.synchronize()
    sim.snapshot(0)

    def configure(mode, cfg):
        # mode: a leg's name ("plain" runs without a warm start)
        sim.restore(0)
        sim.set_warm_start("none" if mode == "plain" else mode, nets[cfg["iterations"]] if mode == "learned" else None)
        if mode != "plain":
            sim.set_warm_start_scale(*autotune._scales(cfg["scale"]))
            sim.set_fallback(limit, scene.iterations)
        else:
            sim.set_fallback(0.0)

    legs = {"plain": {"iterations": scene.iterations}}
    for mode in ("previous", "learned"):
        if found.get(mode):
            legs[mode] = dict(found[mode])
    # timing: every leg from S1, `reps` repeats of the --steps timed frames, median
    def time_leg(mode, cfg):
        configure(mode, cfg)
        sim.step(scene.dt, cfg["iterations"], scene.substeps)  # graph capture
        f0 = sim.fallback_count()
        times = []
        for _ in range(args.reps):
            configure(mode, cfg)
            times.append(autotune.time_frames(sim, scene, args.steps, cfg["iterations"]))
        med = sorted(times)[len(times) // 2]
        cfg.update({"substeps_per_s": round(args.steps * scene.substeps / med, 2), "seconds": times,
                    "fallback_frames": sim.fallback_count() - f0})
        log(f"{mode}: median of {args.reps} x {args.steps} frames {1e3 * med:.2f} ms -> {cfg['substeps_per_s']:.1f} "
            f"substeps/s with {cfg['iterations']} iterations, {cfg['fallback_frames']} fallback frames")

    # equal residual on the timed frames, per state: along the warm-started trajectory from S1 (no
    # fallback), every frame's residual against the plain solver's from the same state
    def verify_leg(mode, cfg):
        configure(mode, cfg)
        sim.set_fallback(0.0)
        rw, rp = autotune.per_state_ratios(sim, scene, args.steps, cfg["iterations"])
        ratio = rw / rp
        above = int(np.sum(~(ratio <= 1.0)))  # a frame whose residual is not finite counts as above
        fin = ratio[np.isfinite(ratio)]

        def num(v):
            return float(v) if np.isfinite(v) else None

        cfg["timed_frames_per_state"] = {
            "ratio_max": num(ratio.max()) if len(fin) == len(ratio) else None,
            "ratio_median": num(np.median(fin)) if len(fin) else None,
            "ratio_min": num(fin.min()) if len(fin) else None, "frames_above_1": above,
            "frames_not_finite": int(len(ratio) - len(fin)), "frames": int(len(ratio)),
            "ratios": [round(float(r), 4) if np.isfinite(r) else None for r in ratio]}
        t = cfg["timed_frames_per_state"]
        log(f"{mode} per state over the {args.steps} timed frames with {cfg['iterations']} iterations: ratio max "
            f"{t['ratio_max']}, median {t['ratio_median']}, {above} frames above 1 ({t['frames_not_finite']} not "
            f"finite)")
        return above

    time_leg("plain", legs["plain"])
    plain_sps = legs["plain"]["substeps_per_s"]
    # a warm-started leg whose timed frames do not all reach the plain residual per state runs again
    # with one more iteration (up to one less than the plain solver's)
    for mode, cfg in legs.items():
        if mode == "plain":
            continue
        cfg["tuned_iterations"] = cfg["iterations"]
        by_k = cfg.get("by_iterations", {})
        while True:
            time_leg(mode, cfg)
            above = verify_leg(mode, cfg)
            if above == 0 or cfg["iterations"] + 1 >= scene.iterations:
                break
            # one more iteration, with the scale the tuner found best for that count on the tuning
            # frames (the current one where none qualified there)
            cfg["iterations"] += 1
            if cfg["iterations"] in by_k:
                cfg["scale"] = by_k[cfg["iterations"]]["scale"]
        cfg["equal_residual"] = above == 0
        cfg["speedup_vs_plain"] = round(cfg["substeps_per_s"] / plain_sps, 3)
    sim.set_warm_start("none")
    sim.set_fallback(0.0)
    sim.restore(0)
    if legs.get("learned", {}).get("equal_residual"):
        solver, head = "learned", legs["learned"]
    else:
        solver, head = "plain", legs["plain"]
    substeps_per_s, iterations = head["substeps_per_s"], head["iterations"]
    log(f"value: {solver} ({substeps_per_s:.1f} substeps/s, {substeps_per_s / plain_sps:.2f}x plain)")
    overflow = sim.contact_overflow()
    assert overflow == 0, f"{overflow} contacts dropped"

    # instrumented region: events around every launch, stream launches (no graph)
    sim.use_graphs(False)
    sim.set_timing(True)
    sim.reset_stats()
    for _ in range(args.profile_steps):
        sim.step(scene.dt, scene.iterations, scene.substeps)
    launches, ms, units = sim.kernel_stats(kind)
    per_kernel = {}
    for k, name in enumerate(S.KERNEL_NAMES):
        n, t, u = sim.kernel_stats(k)
        if n:
            per_kernel[name] = {"launches": n, "avg_us": 1e3 * t / n, "share": 0.0}
    total = sum(v["avg_us"] * v["launches"] for v in per_kernel.values())
    for v in per_kernel.values():
        v["share"] = round(v["avg_us"] * v["launches"] / total, 4)
        v["avg_us"] = round(v["avg_us"], 3)
    sim.set_timing(False)
    avg_launch_s = ms / 1e3 / launches
    bytes_per_launch = BYTES_PER_TET * units / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9
    log(f"projection kernel: {launches} launches, avg {1e6 * avg_launch_s:.2f} us, {units / launches:.0f} tets/launch, "
        f"{achieved:.1f} GB/s algorithmic")
    res = float(sum(sim.residual(0, scene.dt / scene.substeps)) ** 0.5)
    sim.close()

    # HBM bytes per launch measured by the rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs recorded
    # under profiles/ (scaled by a device copy of known size); null for other workloads
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{'neohookean_hex' if kind == S._lib.KERNEL_NEOHOOKEAN else 'green'}_"
                                           f"{scene.name.replace('_neohookean', '')}_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath))["bytes_per_launch"]

    cpu, accuracy = None, None
    if not args.no_cpu_baseline:
        # the oracle: accuracy of the measured solvers from rest (the checker), then the baseline
        if args.accuracy_frames:
            acc_legs = [{"name": "plain", "iterations": scene.iterations, "mode": "none"}]
            for mode in ("previous", "learned"):
                if mode in legs:
                    acc_legs.append({"name": mode, "iterations": legs[mode]["iterations"], "mode": mode,
                                     "scale": legs[mode]["scale"], "net": nets.get(legs[mode]["iterations"])})
            accuracy = accuracy_legs(scene, acc_legs, launch, frames=args.accuracy_frames,
                                     check=(1, args.accuracy_frames))
            accuracy["qualifier"] = (
                "the warm-started legs exceed 1e-4 against the oracle running the same solver after about frame 30 "
                "because those trajectories amplify rounding-sized perturbations: two FP64 oracle runs whose initial "
                "positions differ by one FP32 rounding end about 1e-3 of the bbox apart at frame 60 "
                "(tests/test_gpu_accuracy.py measures that spread and holds the device to three times it); the "
                "plain solver stays within 1e-4, and every warm-started leg sits about 1.6e-2 from the plain FP64 "
                "solution at frame 60, as the oracle's own warm-started solution does")
            accuracy["definition"] = ("max over vertices of |x_device - x_oracle|, divided by the rest bounding-box "
                                      "diagonal, from rest; <leg>: the oracle runs the same solver (colored order, "
                                      "same warm start); <leg>_vs_plain: against the plain FP64 solver; no fallback")
            log(f"accuracy {accuracy}")
        sample = {"batch_512": scenes.batch, "batch_512_large": scenes.batch,
                  "batch_512_contacts": scenes.batch_512_contacts,
                  "batch_512_large_contacts": scenes.batch_512_contacts}.get(args.workload)
        cpu = cpu_baseline(scene, sample(16) if sample else None)
        log(f"cpu baseline {cpu['value']:.3f} substeps/s")

    others = {}
    for w in args.also:
        others[w] = plain_throughput(make_scene[w](), kind, args)

    out = {
        "metric": METRIC,
        "value": round(substeps_per_s, 2),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * scene.substeps / substeps_per_s, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": scene.name,
            "solver": ("learned multiplier warm start + colored Gauss-Seidel XPBD" if solver == "learned" else
                       "plain GPU XPBD (colored Gauss-Seidel, no warm start): " +
                       ("no learned configuration reached the plain residual on every timed frame per state"
                        if warmable else "no warm start tried (--no-learned, a Green group or body contacts)")),
            "tets": scene.num_tets,
            "vertices": scene.num_vertices,
            "substeps_per_step": scene.substeps,
            "iterations": iterations,
            "plain_iterations": scene.iterations,
            "colors": ncolors,
            "launch_config": launch,
            "obstacles": f"ground plane + {len(scene.spheres)} spheres",
            "launch": "hipGraph per frame",
            "parallelism": "single GPU",
            "positions": "fp64 storage, fp32 constraint arithmetic, fp64 residual norms",
        },
        "constraint_projections_per_s": round(substeps_per_s * iterations * scene.num_tets, 1),
        "plain_substeps_per_s": plain_sps,
        "speedup_vs_plain": round(substeps_per_s / plain_sps, 3),
        "legs": legs if warmable else None,
        "residual_definition": (
            f"FP64 frame residual sqrt(sum h^2), h = C + alpha~ lambda, after a frame's last substep. Tuning on the "
            f"{args.tune_frames} frames after the {args.warmup} warm-up frames: a candidate qualifies when on each of "
            f"them its residual is at most the plain {scene.iterations}-iteration solver's from the same state. "
            f"Timing on the next {args.steps} frames (disjoint), median of {args.reps} repeats, fallback check on "
            f"(limit {limit:.4g} = twice the largest plain tuning frame residual); timed_frames_per_state: the same "
            f"per-state comparison over the timed frames (no fallback); a leg with a timed frame above the plain "
            f"residual is timed and compared again with one more iteration and the scale the tuner found best for "
            f"that count on the tuning frames (by_iterations; tuned_iterations: the tuner's smallest count); "
            f"value: the learned leg when every timed frame reaches the plain residual, else the plain solver"),
        "training": train_info,
        "num_trials": len(trials),
        "contact_overflow": overflow,
        "accuracy": accuracy,
        "residual_l2_last_substep": res,
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{scene.groups[0].kind}" + ("_hex" if scene.groups[0].clusters is not None else ""),
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_unit": BYTES_PER_TET,
            "units_per_launch": round(units / launches, 1),
            "avg_launch_us": round(1e6 * avg_launch_s, 3),
        },
        "kernels": per_kernel,
        "other_workloads": others or None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def profile_run(args, scene):
    """A short run for a profiler (rocprofv3 --kernel-trace / --pmc): the launch configuration the
    bench would pick, `--warmup` plain frames, `--profile-frames` plain frames, then for a
    neo-Hookean workload a network trained on `--train-frames` captured frames (default 200) of the
    previous-multiplier start `--warm-config` (iterations:deviatoric,hydrostatic scales) and
    `--profile-frames` frames of the learned start with those iterations and scales. No tuning and no
    timing; one JSON line says what ran."""
    from sbs_amd import autotune, scenes

    sim = scenes.build_device(scene)
    sim.use_graphs(not args.no_graphs)
    launch = autotune.tune_launch(sim, scene, frames=2)
    for _ in range(args.warmup):
        sim.step(scene.dt, scene.iterations, scene.substeps)
    for _ in range(args.profile_frames):
        sim.step(scene.dt, scene.iterations, scene.substeps)
    sim.synchronize()
    out = {"profile_run": scene.name, "launch_config": launch, "plain_frames": args.profile_frames,
           "iterations": scene.iterations}
    if scene.groups[0].kind == "neohookean" and not args.no_learned and scene.body_contacts is None:
        k, scales = args.warm_config.split(":")
        k, scale = int(k), tuple(float(v) for v in scales.split(","))
        net, info = autotune.train_warm_start(sim, scene, iterations=k, scale=scale,
                                              frames=args.train_frames or 200, episodes=4, val_episodes=1,
                                              steps=500, log=log)
        sim.set_warm_start("learned", net)
        sim.set_warm_start_scale(*scale)
        sim.step(scene.dt, k, scene.substeps)  # graph capture
        for _ in range(args.profile_frames):
            sim.step(scene.dt, k, scene.substeps)
        sim.synchronize()
        out.update({"learned_frames": args.profile_frames, "learned_iterations": k, "learned_scale": scale,
                    "held_out_loss": info["held_out_loss"]})
    out["contact_overflow"] = sim.contact_overflow()
    sim.close()
    return out


def plain_throughput(scene, kind, args, warmup=3, frames=5, reps=3):
    """The plain solver on another workload (no tuning): median of `reps` x `frames` graph-replayed
    frames after `warmup`, and the projection kernel's average launch from an instrumented frame."""
    import sbs_amd as S
    from sbs_amd import autotune, scenes

    t0 = time.perf_counter()
    sim = scenes.build_device(scene)
    build = time.perf_counter() - t0
    sim.use_graphs(True)
    launch = autotune.tune_launch(sim, scene, frames=2)
    for _ in range(warmup):
        sim.step(scene.dt, scene.iterations, scene.substeps)
    sim.snapshot(0)
    times = []
    for _ in range(reps):
        sim.restore(0)
        times.append(autotune.time_frames(sim, scene, frames, scene.iterations))
    med = sorted(times)[len(times) // 2]
    sim.use_graphs(False)
    sim.set_timing(True)
    sim.reset_stats()
    sim.step(scene.dt, scene.iterations, scene.substeps)
    launches, ms, units = sim.kernel_stats(kind)
    _, detect_ms, _ = sim.kernel_stats(S._lib.KERNEL_DETECT)
    sim.set_timing(False)
    overflow = sim.contact_overflow()
    body_contacts = None
    if scene.body_contacts is not None:  # the contacts of the state after the timed frames, by kind
        import numpy as np
        ids, _ = sim.read_contacts()
        tri = sim.body_contact_setup()[4]
        nsdf = len(scene.planes) + len(scene.spheres) + len(scene.grids) + len(scene.meshes)  # the SDFs, in order
        t = ids[ids[:, 3] >= nsdf]
        other = int(np.sum(tri[t[:, 3] - nsdf, 3] != t[:, 0]))
        body_contacts = {"surface_triangles": int(len(tri)), "contacts_with_other_bodies": other,
                         "contacts_with_own_surface": int(len(t)) - other,
                         "contacts_with_obstacles": int(len(ids) - len(t))}
    sim.close()
    avg = ms / 1e3 / launches
    gbs = BYTES_PER_TET * units / launches / avg / 1e9
    sps = frames * scene.substeps / med
    log(f"{scene.name}: plain {sps:.1f} substeps/s, projection {1e6 * avg:.2f} us per launch ({gbs:.0f} GB/s), "
        f"built in {build:.1f}s")
    return {"workload": scene.name, "tets": scene.num_tets, "plain_substeps_per_s": round(sps, 2),
            "launch_config": launch, "projection_avg_launch_us": round(1e6 * avg, 3),
            "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "achieved_GBps": round(gbs, 1),
            "contact_overflow": overflow, "frames": f"median of {reps} x {frames} frames after {warmup}",
            "detect_ms_per_frame": round(detect_ms, 3),
            "body_contacts": body_contacts}


if __name__ == "__main__":
    main()
