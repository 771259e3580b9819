"""Host-side tuning of a device simulation: launch configuration, and the learned multiplier warm
start (training data capture, training, and the choice of warm-start scale and iteration count
that reaches the plain solver's residual).

Every trial starts from the same device snapshot (sbs_gpu_snapshot / sbs_gpu_restore), so the
candidates are compared on identical states. Residuals are the FP64 XPBD residual norms
sqrt(sum h^2), h = C + alpha~ lambda, over the neo-Hookean groups after each frame's last substep.
"""
import math
import time

import numpy as np

from ._lib import Mlp


def frame_residual(sim, scene):
    dt = scene.dt / scene.substeps
    total = 0.0
    for gi, g in enumerate(scene.groups):
        if g.kind == "neohookean":
            total += float(np.sum(sim.residual(gi, dt)))
    return math.sqrt(total)


def frame_residuals(sim, scene, frames, iterations):
    """Frame residuals over `frames` frames from the current state."""
    r = []
    for _ in range(frames):
        sim.step(scene.dt, iterations, scene.substeps)
        r.append(frame_residual(sim, scene))
    return np.asarray(r)


def mean_residual(sim, scene, frames, iterations):
    """Mean frame residual over `frames` frames from the current state (inf if any is not finite)."""
    r = frame_residuals(sim, scene, frames, iterations)
    return float(r.mean()) if np.all(np.isfinite(r)) else math.inf


def per_state_ratios(sim, scene, frames, iterations, plain_iterations=None, slot=1):
    """Frame by frame from the current state, with the warm start as configured: the state X_f is
    copied (snapshot slot `slot`), the plain solver runs one frame from it (plain_iterations, the
    scene's count by default) and gives r_plain(X_f); X_f is restored and the warm-started solver runs
    the frame (`iterations`) and gives r_warm(X_f), continuing the warm-started trajectory. Returns
    (r_warm, r_plain) over the frames: equal residual on the same state means r_warm <= r_plain."""
    plain_iterations = scene.iterations if plain_iterations is None else plain_iterations
    mode, net = sim.warm_mode, sim.warm_net
    rw, rp = [], []
    for _ in range(frames):
        sim.snapshot(slot)
        sim.set_warm_start("none")
        sim.step(scene.dt, plain_iterations, scene.substeps)
        rp.append(frame_residual(sim, scene))
        sim.restore(slot)
        sim.set_warm_start(mode, net)
        sim.step(scene.dt, iterations, scene.substeps)
        rw.append(frame_residual(sim, scene))
    return np.asarray(rw), np.asarray(rp)


def time_frames(sim, scene, frames, iterations):
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        sim.step(scene.dt, iterations, scene.substeps)
    sim.synchronize()
    return time.perf_counter() - t0


STREAMED = 65  # SBS_CLUSTER_BLOCK_STREAMED


def tune_launch(sim, scene, frames=5, blocks=(128, 256), cluster_blocks=(64, 128, STREAMED)):
    """Fastest (block, cluster_block) over a few graph-replayed frames; leaves it set."""
    best = None
    sim.snapshot()
    for b in blocks:
        for cb in cluster_blocks:
            sim.restore()
            sim.set_launch_config(b, cb)
            sim.step(scene.dt, scene.iterations, scene.substeps)  # capture
            t = time_frames(sim, scene, frames, scene.iterations)
            if best is None or t < best[0]:
                best = (t, b, cb)
    sim.restore()
    sim.set_launch_config(best[1], best[2])
    return {"block": best[1], "cluster_block": best[2], "ms_per_frame": 1e3 * best[0] / frames}


def _scales(s):
    """A warm-start scale: one factor, or a (deviatoric, hydrostatic) pair."""
    return tuple(s) if isinstance(s, (tuple, list)) else (s,)


# Warm-start scales the tuner tries: pairs (deviatoric, hydrostatic) on a 0.05 grid. The stable
# neo-Hookean pair responds differently to carried-over multipliers: on the 1M beam a deviatoric
# factor near 0.7 with a hydrostatic one near 0.65 reaches the plain residual with fewer iterations
# than any single factor.
SCALE_GRID = tuple((d, h) for d in (0.5, 0.55, 0.6, 0.65, 0.7, 0.75) for h in (0.5, 0.55, 0.6, 0.65, 0.7, 0.75))


WARM_NET_INPUTS, WARM_NET_OUTPUTS = 39, 12  # per hex cluster: features in, multiplier changes out


def _random_direction(rng):
    v = rng.standard_normal(3)
    return v / np.linalg.norm(v)


def capture_episodes(sim, scene, net, episodes, frames, iterations, scale, stride, rng, impulse=(0.5, 2.0),
                     radius=(0.1, 0.5), log=print):
    """Training rows from `episodes` trajectories of `frames` frames each, all starting at the state in
    snapshot slot 1, with the previous-multiplier warm start (`iterations`, `scale`): each episode
    under gravity of the scene's magnitude in a random direction, and with an impulse load at its
    start (a velocity change of U[impulse] m/s in a random direction added to the movable particles
    within U[radius] m of a random particle of a random body). Rows go to net's dataset (capture
    stride `stride`). The scene's gravity is restored at the end; the state is left as the last
    episode ends. Returns per episode its range of dataset rows and its largest frame residual."""
    g = float(np.linalg.norm(scene.gravity))
    record = []
    for e in range(episodes):
        sim.restore(1)
        d = _random_direction(rng)
        sim.set_params(gravity=tuple(g * d), gravity_is_force=scene.gravity_is_force,
                       collision_compliance=scene.collision_compliance)
        b = int(rng.integers(len(scene.bodies)))
        x, v = sim.read_body(b)
        c = x[int(rng.integers(len(x)))]
        r = float(rng.uniform(*radius))
        sel = np.linalg.norm(x - c, axis=1) <= r
        if scene.bodies[b].mass is not None:
            sel &= np.asarray(scene.bodies[b].mass) > 0
        v = v.copy()
        v[sel] += float(rng.uniform(*impulse)) * _random_direction(rng)
        sim.write_body(b, None, v)
        sim.set_warm_start("previous")
        sim.set_warm_start_scale(*_scales(scale))
        first = net.dataset_size()
        sim.capture(net, stride)
        largest = 0.0
        for _ in range(frames):
            sim.step(scene.dt, iterations, scene.substeps)
            r = frame_residual(sim, scene)
            largest = r if not np.isfinite(r) else max(largest, r)
        sim.capture(None, 0)
        record.append((first, net.dataset_size(), largest))
    sim.set_params(gravity=scene.gravity, gravity_is_force=scene.gravity_is_force,
                   collision_compliance=scene.collision_compliance)
    sim.set_warm_start("none")
    return record


def diverged_episodes(record, factor=20.0):
    """The (first row, end row) ranges of the episodes in `record` ((first, end, largest frame
    residual) each) whose trajectory diverged: a largest frame residual that is not finite or above
    `factor` times the median over the episodes with a finite one."""
    largest = np.array([r for _, _, r in record], dtype=np.float64)
    finite = largest[np.isfinite(largest)]
    if not len(finite):
        return [(a, b) for a, b, _ in record]
    limit = factor * float(np.median(finite))
    return [(a, b) for a, b, r in record if not (np.isfinite(r) and r <= limit)]


def drop_diverged_episodes(net, record, factor=20.0):
    """Removes from net's dataset the rows of the episodes diverged_episodes names (the warm start
    they were captured along blew up, and its multipliers would set the network's normalization).
    Returns the number of episodes removed."""
    bad = diverged_episodes(record, factor)
    if not bad:
        return 0
    x, t = net.dataset()
    keep = np.ones(len(x), dtype=bool)
    for a, b in bad:
        keep[a:b] = False
    net.clear()
    if keep.any():
        net.append(x[keep], t[keep])
    return len(bad)


def train_warm_start(sim, scene, iterations=4, scale=(0.7, 0.65), frames=2000, episodes=40, val_episodes=4,
                     max_rows=4_000_000, steps=3000, batch=16384, lr=1e-3, seed=0, log=print):
    """The learned warm start's network, trained on the device on trajectories the solver generates:
    `frames` frames in `episodes` episodes from the current state under random gravity directions and
    impulse loads (capture_episodes), captured with the previous-multiplier warm start (`iterations`,
    `scale`); one row per sampled hex cluster and substep (the capture stride keeps about max_rows
    rows). `val_episodes` more episodes with their own loads are held out: the normalized MSE on them
    is reported beside the training loss. Adam on minibatches of `batch` rows for `steps` steps.
    Returns the network and a record of the training; the state is restored."""
    clusters = sum(len(g.clusters) - 1 for g in scene.groups if g.kind == "neohookean" and g.clusters is not None)
    per_episode = max(1, frames // episodes)
    stride = max(1, -(-clusters * scene.substeps * per_episode * episodes // max_rows))
    rng = np.random.default_rng(seed)
    net = Mlp(WARM_NET_INPUTS, WARM_NET_OUTPUTS, seed=seed)
    net.reserve(max_rows)
    val = Mlp(WARM_NET_INPUTS, WARM_NET_OUTPUTS, seed=seed)
    val.reserve(max(1, max_rows * val_episodes // max(episodes, 1)))
    sim.snapshot(1)
    t0 = time.perf_counter()
    rec = capture_episodes(sim, scene, net, episodes, per_episode, iterations, scale, stride, rng, log=log)
    rec_val = capture_episodes(sim, scene, val, val_episodes, per_episode, iterations, scale, stride, rng, log=log)
    dropped = drop_diverged_episodes(net, rec)
    # the held-out episodes are measured against the training episodes' typical residual
    typical = float(np.median([r for _, _, r in rec if np.isfinite(r)] or [np.inf]))
    dropped_val = drop_diverged_episodes(val, rec_val + [(0, 0, typical)] * len(rec))
    cap_s = time.perf_counter() - t0
    sim.restore(1)
    n = net.dataset_size()
    net.fit_normalization()
    t0 = time.perf_counter()
    loss = net.train(steps, batch, lr, seed=seed + 1)
    train_s = time.perf_counter() - t0
    xv, tv = val.dataset()
    val_loss, _ = net.loss_grad(xv, tv)
    # the same error for the network that predicts no change (the previous-multiplier start)
    zero = float(np.mean((tv / net.get_normalization()[2]) ** 2))
    val.close()
    if dropped or dropped_val:
        log(f"warm-start network: {dropped} training and {dropped_val} held-out episodes diverged and were dropped")
    log(f"warm-start network: {n} rows from {episodes} x {per_episode} frames (stride {stride}, capture {cap_s:.1f}s), "
        f"{steps} Adam steps of {batch} in {train_s:.1f}s, loss {loss:.4f}; held out ({len(xv)} rows of "
        f"{val_episodes} episodes): {val_loss:.4f} (no change predicted: {zero:.4f})")
    return net, {"rows": n, "stride": stride, "episodes": episodes, "frames_per_episode": per_episode,
                 "frames": episodes * per_episode, "captured_with": {"iterations": iterations, "scale": scale},
                 "steps": steps, "batch": batch, "capture_s": round(cap_s, 2), "train_s": round(train_s, 2),
                 "train_loss": loss, "held_out_rows": int(len(xv)), "held_out_loss": val_loss,
                 "held_out_loss_no_change": zero, "episodes_dropped": dropped,
                 "held_out_episodes_dropped": dropped_val,
                 "loads": "gravity of the scene's magnitude in a random direction per episode, and a velocity "
                          "impulse U[0.5, 2] m/s on the particles within U[0.1, 0.5] m of a random particle"}


def training_config(args, found=None):
    """Keyword arguments of train_warm_start from the bench's command line and the tuned
    previous-multiplier configuration (captured trajectories use it)."""
    out = {}
    if getattr(args, "train_frames", None):
        out["frames"] = args.train_frames
    if found:
        out["iterations"], out["scale"] = found["iterations"], found["scale"]
    return out


def tune_warm_start(sim, scene, net, target, frames=8, scales=SCALE_GRID, iterations=(2, 3, 4, 5, 6, 7),
                    modes=("learned", "previous"), screen_slack=1.15, all_iterations=False, log=print):
    """For each mode, the cheapest (iterations, scale) that reaches the plain solver's residual on the
    same states over `frames` frames from the current state. Each candidate is screened by its mean
    frame residual along its own trajectory (<= screen_slack x target, the plain solver's mean over
    those frames); the ones that pass are qualified per state (per_state_ratios): every frame's
    warm-started residual must be at most the plain solver's from the same state. Among the scales
    that qualify at the smallest iteration count, the one with the smallest largest ratio is taken.
    A scale is one factor or a (deviatoric, hydrostatic) pair; net is one network or a dict of networks
    by iteration count (the learned start at k iterations uses net[k]; counts without one are
    skipped). Returns {mode: {"iterations", "scale", "screen_mean", "max_ratio", "mean_ratio"} or None}
    and the trials, and leaves the simulation at the state it started from (snapshot slot 0), warm
    start off. With all_iterations every count is tried and the result carries "by_iterations": the
    best qualifying scale of each count that has one."""
    sim.snapshot(0)
    found = {}
    trials = []

    def net_for(k):
        return net.get(k) if isinstance(net, dict) else net

    def start(mode, s, k):
        sim.restore(0)
        sim.set_warm_start(mode, net_for(k) if mode == "learned" else None)
        sim.set_warm_start_scale(*_scales(s))

    for mode in modes:
        best, by_k = None, {}
        for k in iterations:
            if mode == "learned" and net_for(k) is None:
                continue
            best_k = None
            passing = []
            for s in scales:
                start(mode, s, k)
                r = mean_residual(sim, scene, frames, k)
                trials.append({"mode": mode, "iterations": k, "scale": s, "residual": r})
                log(f"  {mode} x{s} {k} iterations: mean residual {r:.4g} (target {target:.4g})")
                if r <= screen_slack * target:
                    passing.append((r, s))
            for r, s in sorted(passing, key=lambda p: p[0]):
                start(mode, s, k)
                rw, rp = per_state_ratios(sim, scene, frames, k)
                ratio = rw / rp
                ok = bool(np.all(np.isfinite(ratio))) and float(ratio.max()) <= 1.0
                log(f"  {mode} x{s} {k} iterations per state over {frames} frames: ratio max {ratio.max():.4g} "
                    f"mean {ratio.mean():.4g}{'' if ok else ': rejected'}")
                if ok and (best_k is None or ratio.max() < best_k["max_ratio"]):
                    best_k = {"iterations": k, "scale": s, "screen_mean": r, "max_ratio": float(ratio.max()),
                              "mean_ratio": float(ratio.mean())}
            if best_k is not None:
                by_k[k] = best_k
                if best is None:
                    best = dict(best_k)  # the smallest iteration count that reaches the target
                if not all_iterations:
                    break
        if best is not None and all_iterations:
            best["by_iterations"] = by_k
        found[mode] = best
    sim.set_warm_start("none")
    sim.restore(0)
    return found, trials
