#!/usr/bin/env python3
"""Benchmark of the MI355X XPBD time step on the 1M-tet cantilever beam.

One "step" is one timestep_t::step() call (one frame = `substeps` substeps of `iterations`
Gauss-Seidel sweeps over the 997,920 stable neo-Hookean tets, contacts against the ground plane and
16 spheres). Timed regions replay the frame as a hipGraph; inputs are resident in HBM.

The metric compares warm-started solvers against the plain solver at equal residual. From one
device snapshot the plain solver runs the timed horizon (`--steps` frames); its mean FP64 frame
residual over those frames is the target and twice its largest frame residual the fallback limit.
The autotuner screens warm-start configurations (previous multipliers, then a network trained on the
device from captured frames; iterations and a pair of multiplier scales) on the first 8 frames and
verifies the cheapest over the whole horizon: its mean must not exceed the plain solver's and no
frame may exceed the limit. Every qualifying configuration is timed from the snapshot with its
per-frame residual check and plain fallback active; `value` is the fastest (the plain solver's when
none qualifies).

After the timed regions an instrumented region (HIP events around every launch on the library's
stream) measures the average duration of the tet projection kernel for the roofline figure. The
CPU baseline times the FP64 oracle in the reference's insertion order on a bounded sample (frames of
the same scene, or of a 16-body batch for the batches, until about 10 s of CPU work).
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


def cpu_baseline(scene, sample=None, min_seconds=10.0, max_frames=8):
    """FP64 oracle, reference (insertion) order, one thread: frames of the same scene (or of a
    sample scene, scaled to the full scene by tet count) until min_seconds of CPU work or
    max_frames frames."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    run = sample or scene
    t0 = time.perf_counter()
    sim = O.Simulation()
    run.apply(sim)
    sim.set_order(False)
    setup = time.perf_counter() - t0
    log(f"cpu baseline: {run.name} built in {setup:.1f}s, timing frames of {run.substeps} substeps "
        f"for at least {min_seconds:.0f}s")
    frames, el = 0, 0.0
    while frames < max_frames and (frames == 0 or el < min_seconds):
        t0 = time.perf_counter()
        sim.step(run.dt, run.iterations, run.substeps)
        el += time.perf_counter() - t0
        frames += 1
    scale = scene.num_tets / run.num_tets
    what = (f"{frames} frames = {frames * run.substeps} substeps x {run.iterations} iterations of {run.name} "
            f"({run.num_tets} tets)" + (f", scaled by tet count to {scene.num_tets} tets" if sample else ""))
    return {"value": frames * scene.substeps / (el * scale), "unit": "substeps/s", "cores": 1, "kind": "port",
            "sample": f"{what}; FP64 oracle in reference insertion order, single thread, {el:.1f}s on {cpu_model()}",
            "projections_per_s": run.num_tets * run.iterations * run.substeps * frames / el}


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
    ap.add_argument("--workload", default="beam_1m", choices=["beam_1m", "beam_6m", "beam_540", "batch_512", "batch_512_large"])
    ap.add_argument("--no-learned", action="store_true", help="plain solver only")
    ap.add_argument("--no-graphs", action="store_true", help="stream launches instead of hipGraph replays")
    ap.add_argument("--no-contact-fusion", action="store_true",
                    help="the first iteration's contacts in their own launch, not in the warm-start gather")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run device copies of known size (to scale PMC byte counters)")
    ap.add_argument("--row", default="step", choices=["step", "boundary", "grid_sdf", "ply"],
                    help="step: the time step (the headline metric); the others measure the surface, grid SDF "
                         "and PLY components on their own workloads")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the time step runs on a single MI355X; sharding the bodies or the "
              "beam across GPUs is not implemented, so --gpus must be 1.", file=sys.stderr)
        sys.exit(2)
    if args.row != "step":
        out = {"boundary": row_boundary, "grid_sdf": row_grid_sdf, "ply": row_ply}[args.row](args)
        out.update({"n_gpus": 1})
        print(json.dumps(out), flush=True)
        return

    import sbs_amd as S
    from sbs_amd import scenes

    if args.copy_calibration:
        copy_calibration()

    scene = {"beam_1m": scenes.beam_1m, "beam_6m": scenes.beam_6m, "beam_540": scenes.beam_540,
             "batch_512": scenes.batch_512, "batch_512_large": scenes.batch_512_large}[args.workload]()
    log(f"workload {scene.name}: {scene.num_vertices} vertices, {scene.num_tets} tets")
    t0 = time.perf_counter()
    sim = scenes.build_device(scene)
    _, ncolors = sim.group_colors(0)
    log(f"committed in {time.perf_counter() - t0:.1f}s, {ncolors} colors")
    kind = {"green": S._lib.KERNEL_GREEN, "neohookean": S._lib.KERNEL_NEOHOOKEAN}[scene.groups[0].kind]

    from sbs_amd import autotune

    sim.use_graphs(not args.no_graphs)
    sim.set_contact_fusion(not args.no_contact_fusion)
    launch = autotune.tune_launch(sim, scene)
    log(f"launch config {launch}")
    for _ in range(args.warmup):
        sim.step(scene.dt, scene.iterations, scene.substeps)
    sim.synchronize()
    sim.snapshot()
    # the plain solver's frame residuals over the timed horizon from the snapshot: their mean is the
    # residual a warm start must reach over the same frames, twice their largest is the fallback
    # limit of a timed warm-started frame; the first eval_frames of them screen the tuner's trials
    eval_frames = 8
    plain_frames = autotune.frame_residuals(sim, scene, args.steps, scene.iterations)
    sim.restore()
    target = float(plain_frames[:eval_frames].mean())
    horizon_target = float(plain_frames.mean())
    limit = 2.0 * float(plain_frames.max())
    log(f"warmup done ({args.warmup} frames); plain {scene.iterations}-iteration frame residual: mean {horizon_target:.4g} "
        f"over {args.steps} frames (first {eval_frames}: {target:.4g}), largest {plain_frames.max():.4g}")

    learned = None
    if not args.no_learned:
        # the previous-multiplier start first; the network then learns, from that configuration's own
        # trajectory (or a 24-iteration one when it never reaches the target), the multipliers the
        # iterations end with, and is tuned the same way
        found, trials = autotune.tune_warm_start(sim, scene, None, target, frames=eval_frames, modes=("previous",),
                                                 verify_frames=args.steps, verify_target=horizon_target, limit=limit,
                                                 log=log)
        prev = found["previous"]
        cap = (prev["iterations"], prev["scale"]) if prev else (24, 0.5)
        net = autotune.train_warm_start(sim, scene, iterations=cap[0], scale=cap[1], log=log)
        found_l, trials_l = autotune.tune_warm_start(sim, scene, net, target, frames=eval_frames, modes=("learned",),
                                                     verify_frames=args.steps, verify_target=horizon_target,
                                                     limit=limit, log=log)
        found.update(found_l)
        # per (mode, iterations): the trial with the smallest mean residual
        best_trials = {}
        for t in trials + trials_l:
            k = (t["mode"], t["iterations"])
            if k not in best_trials or t["residual"] < best_trials[k]["residual"]:
                best_trials[k] = t
        learned = {"found": found, "best_trials": sorted(best_trials.values(), key=lambda t: (t["mode"], t["iterations"])),
                   "num_trials": len(trials) + len(trials_l), "capture": {"iterations": cap[0], "scale": cap[1]}}
        sim.restore()

    log(f"timing {args.steps} frames, plain")
    elapsed_plain = autotune.time_frames(sim, scene, args.steps, scene.iterations)
    plain_sps = args.steps * scene.substeps / elapsed_plain
    log(f"plain: {elapsed_plain:.3f}s -> {plain_sps:.1f} substeps/s")
    # every warm-start configuration that reached the plain residual is timed with its per-frame
    # residual check and plain fallback on; the fastest configuration overall is the result
    best = {"solver": "plain", "substeps_per_s": plain_sps, "elapsed": elapsed_plain, "iterations": scene.iterations}
    if learned:
        learned["timed"] = {}
        for mode, cfg in learned["found"].items():
            if not cfg:
                continue
            sim.restore()
            sim.set_warm_start(mode, net if mode == "learned" else None)
            sim.set_warm_start_scale(*autotune._scales(cfg["scale"]))
            sim.set_fallback(limit, scene.iterations)
            sim.step(scene.dt, cfg["iterations"], scene.substeps)  # graph capture
            sim.restore()
            f0 = sim.fallback_count()
            el = autotune.time_frames(sim, scene, args.steps, cfg["iterations"])
            sps = args.steps * scene.substeps / el
            nfb = sim.fallback_count() - f0
            learned["timed"][mode] = {"substeps_per_s": round(sps, 2), "iterations": cfg["iterations"],
                                      "scale": cfg["scale"], "fallback_frames": nfb, "fallback_limit": limit,
                                      "speedup_vs_plain": round(sps / plain_sps, 3)}
            log(f"{mode} warm start: {el:.3f}s -> {sps:.1f} substeps/s ({sps / plain_sps:.2f}x plain), "
                f"{nfb} fallback frames")
            if sps > best["substeps_per_s"]:
                best = {"solver": mode, "substeps_per_s": sps, "elapsed": el, "iterations": cfg["iterations"]}
            sim.set_fallback(0.0)
            sim.set_warm_start("none")
    sim.restore()
    substeps_per_s, elapsed, solver, iterations = best["substeps_per_s"], best["elapsed"], best["solver"], best["iterations"]
    log(f"fastest at the plain residual: {solver} ({substeps_per_s:.1f} substeps/s)")

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
    # residual after the last substep (FP64), a correctness anchor for later comparisons
    res = float(sum(sim.residual(0, scene.dt / scene.substeps)) ** 0.5)

    # HBM bytes per launch measured by the rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs recorded
    # under profiles/ (scaled by a device copy of known size); null for other workloads
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{'neohookean_hex' if kind == S._lib.KERNEL_NEOHOOKEAN else 'green'}_"
                                           f"{scene.name.replace('_neohookean', '')}_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath))["bytes_per_launch"]

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, scenes.batch(16) if args.workload in ("batch_512", "batch_512_large") else None)
        log(f"cpu baseline {cpu['value']:.3f} substeps/s")

    out = {
        "metric": METRIC,
        "value": round(substeps_per_s, 2),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": scene.name,
            "solver": {"learned": "learned multiplier warm start + colored Gauss-Seidel XPBD",
                       "previous": "previous-multiplier warm start + colored Gauss-Seidel XPBD",
                       "plain": "plain GPU XPBD (colored Gauss-Seidel, no warm start)"}[solver]
                      + ", fastest configuration reaching the plain solver's mean frame residual",
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
        },
        "constraint_projections_per_s": round(substeps_per_s * iterations * scene.num_tets, 1),
        "plain_substeps_per_s": round(plain_sps, 2),
        "residual_target": horizon_target,
        "residual_definition": (f"mean FP64 frame residual sqrt(sum h^2), h = C + alpha~ lambda after each frame's last "
                                f"substep, over the {args.steps} timed frames from the same state; a warm-started "
                                f"configuration qualifies when its mean is <= the plain solver's and no frame exceeds "
                                f"twice the plain solver's largest (its per-frame fallback limit)"),
        "plain_frame_residual_max": float(plain_frames.max()),
        "learned": learned,
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
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
