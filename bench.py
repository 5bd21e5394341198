"""Throughput of the XPBD substep pipeline on one MI355X: body-substeps/s.

Default workload (--workload combined): 262,144 dynamic bodies in 1,024 independent
scenes of 64-512 bodies (convex hulls from a shared library, boxes, 8-segment ragdolls on
revolute + spherical joints with 60 deg limits) in one world, dt = 1/120 s, 8 substeps,
FP32 state, XPBD joint rows.  The timed leg is the learned + batched pipeline: GJK/EPA
warm-started by the network (trained on the GJK/EPA pairs the engine logs in the first
untimed frames), one workgroup per scene for colouring and solve, substeps replayed from
a hipGraph or launched directly as the autotuner picks.  A "step" is one World::step (one
frame); --steps frames are timed after the untimed frames, inputs resident in HBM.
value = dynamic bodies * substeps * steps / seconds.

Untimed frames before the timed region: max(--warmup, 60) frames with the training log
open, then the autotuner's trial frames and settling frames up to 60 more.  The plain leg
steps the same frames, so both legs time the same frames of the same scene.  If no GJK/EPA
pair reaches the log (every hull pair rejected exactly before GJK), the learned leg keeps
the plain seeds and says so.

Beside the value:
  plain        — the plain GPU pipeline (plain GJK/EPA seeds, grid-wide colouring,
                 per-colour launches, autotuned too) on a fresh world over the same frames.
  roofline     — the dominant kernel's algorithmic bytes per launch / its average launch
                 duration (HIP events on the world's stream around every launch of that
                 kernel in every 10th timed frame), against 8 TB/s HBM; `solve` the same
                 for the per-scene position solve, over frames after the timed region.
  accuracy     — GPU (FP32, the batched per-scene path) against the reference engine
                 (FP64, one World per scene, tests/golden/accuracy_resting_layers.npz)
                 after 300 steps of a well-posed layer of the small plain configuration.
  cpu_baseline — the oracle restatement of the reference step (FP64, the reference's
                 sequential order), one World per scene on all host threads (the single
                 pile: one core), started from the GPU's state at the first timed frame, for
                 a bounded time; restatement/reference speed measured on the small plain
                 configuration (tests/golden/cpu_ratio_pile_small.json).
"""
import argparse
import json
import os
import subprocess
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "body-substeps/s on 1 MI355X (learned+batched vs plain), pos err ≤1e-4 m"
WORKLOAD = "plain GPU pipeline, large pile: 65,536 mixed convex bodies, 10 substeps, 500 steps"
COMBINED_WORKLOAD = ("combined at scale: 262,144 bodies in 1,024 uneven scenes (64-512 bodies) of convex hulls, "
                     "boxes and 8-segment ragdolls (revolute + spherical joints, limits +-60 deg), dt 1/120 s, "
                     "8 substeps, XPBD joint rows, capsule-capsule normals from A to B, learned GJK/EPA warm start")
HULL_WORKLOAD = ("GJK/EPA-heavy: 32,768 random convex hulls (16-64 vertices on ellipsoids with semi-axes "
                 "U(0.2,0.8), triangulated) in a dense pile over a plane, 10 substeps, plain GJK/EPA")
LOG_FRAMES = 60      # frames with the training log open (at least --warmup)
SETTLE_FRAMES = 60   # autotuner trials and settling frames after them
CPU_BUDGET_S = 10.0


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------- host CPU baseline
def host_cpu():
    """(model name, logical CPUs, threads this run may use).  Threads: the CPUs this
    process may run on, at most 16 (one GPU's share of a shared host)."""
    model, ncpu = None, os.cpu_count()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name" and model is None:
                model = v.strip()
            elif k.strip() == "CPU(s)":
                ncpu = int(v.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), None)
        except OSError:
            pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, ncpu, max(1, min(16, avail))


def cpu_ratio():
    path = os.path.join(ROOT, "tests", "golden", "cpu_ratio_pile_small.json")
    try:
        with open(path) as f:
            return json.load(f)
    except OSError:
        return None


def cpu_baseline(parts, states, threads, joint_mode=0, contact_mode=0, budget_s=CPU_BUDGET_S, note=""):
    """The oracle restatement (FP64, reference order), one World per scene, scenes dealt to
    `threads` threads that step their scenes a frame at a time, round robin, until the
    budget is spent.  Each World starts from the given state (the GPU's at the first timed
    frame).  Set-up is not timed; the step calls release the GIL."""
    from oracle import orc
    worlds = []
    for sc, st in zip(parts, states):
        w = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
        w.set_joint_mode(joint_mode)
        w.set_contact_mode(contact_mode)
        if st is not None:
            w.set_state(st)
        worlds.append((w, sc.n_dynamic * sc.substeps))
    work = [0] * threads
    frames = [0] * threads
    start = threading.Barrier(threads + 1)
    deadline = [0.0]

    def run(t):
        mine = worlds[t::threads]
        start.wait()
        while mine and time.perf_counter() < deadline[0]:
            for w, units in mine:
                w.step(1)
                work[t] += units
                frames[t] += 1
                if time.perf_counter() >= deadline[0]:
                    break

    ths = [threading.Thread(target=run, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    t0 = time.perf_counter()
    deadline[0] = t0 + budget_s
    start.wait()
    for th in ths:
        th.join()
    dt = time.perf_counter() - t0
    model, ncpu, _ = host_cpu()
    ratio = cpu_ratio()
    out = {"value": round(sum(work) / dt, 1), "unit": "body-substeps/s", "cores": threads, "kind": "port",
           "cpu_model": model, "host_cpus": ncpu,
           "sample": (f"{len(worlds)} scene World(s) x {sum(frames)} frames in total over {dt:.1f} s on {threads} "
                      f"thread(s), one World per scene, FP64, the reference's sequential order, started from the "
                      f"GPU state at the first timed frame{note}")}
    if ratio:
        out["restatement_over_reference"] = ratio["port_over_reference"]
        out["reference_equivalent_value"] = round(out["value"] / ratio["port_over_reference"], 1)
    return out


def split_states(sc, state, parts):
    """Per-scene rows of a batch's body state, in the order of `parts` (scenes.split)."""
    out = []
    scene = sc.bodies["scene"]
    for p in parts:
        sel = np.nonzero(scene == p.meta["scene"])[0]
        out.append({k: v[sel[0]:sel[-1] + 1] for k, v in state.items()})
    return out


# ---------------------------------------------------------------- accuracy
def accuracy(world, scenes, options):
    """GPU vs the reference engine on the well-posed accuracy batch (50 resting layers of
    the small plain configuration's bodies, 1,000 bodies): position error per body at the
    fixture's frames, the batched per-scene path with the timed leg's options."""
    path = os.path.join(ROOT, "tests", "golden", "accuracy_resting_layers.npz")
    if not os.path.exists(path):
        return None
    fx = np.load(path, allow_pickle=False)
    sc = scenes.accuracy_batch()
    if not (np.array_equal(sc.bodies, fx["bodies"]) and np.array_equal(sc.colliders, fx["colliders"])):
        return {"error": "the accuracy scene generator no longer matches the fixture"}
    w = world.World(sc.timestep, sc.substeps).load(sc)
    for k, v in options.items():
        w.set_option(k, v)
    dyn = sc.bodies["type"] == 1
    is_box = np.zeros(len(sc.bodies), bool)
    is_box[sc.colliders["body_id"][sc.colliders["shape"] == 1]] = True
    curve, done, e = {}, 0, None
    for f in (int(x) for x in fx["frames"]):
        w.step(f - done)
        done = f
        e = np.linalg.norm(w.state()["position"] - fx[f"position_{f}"], axis=1)
        curve[str(f)] = float(e[dyn].max())
    ed = e[dyn]
    st = w.stats()
    w.close()
    return {"scene": f"{sc.name}: 50 independent layers of 20 bodies (unit boxes start flat and slide, spheres "
                     "drop), dt 1/60 s, 20 substeps; reference = the reference engine, FP64, one World per scene",
            "frames": done, "bodies": int(dyn.sum()), "max_m": float(ed.max()), "p99_m": float(np.quantile(ed, 0.99)),
            "median_m": float(np.median(ed)), "within_1e-4": float((ed <= 1e-4).mean()),
            "spheres_max_m": float(e[dyn & ~is_box].max()), "boxes_max_m": float(e[dyn & is_box].max()),
            "max_m_by_frame": curve, "scene_kernels": st["scene_kernels"]}


# ---------------------------------------------------------------- GPU legs
def run_frames(w, steps, label=""):
    """Steps `steps` frames (in chunks of 50 only to print progress) and returns the
    seconds between two synchronisations of the world's stream."""
    w.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(50, steps - done)
        w.step(k)
        done += k
        log(f"{label + ' ' if label else ''}{done}/{steps} frames")
    w.synchronize()
    return time.perf_counter() - t0


def tune_and_settle(w, frames_per_trial=3, settle=SETTLE_FRAMES):
    """Autotune (result-neutral options, real frames), then step up to `settle` frames
    after the call, so that differently tuned legs time the same frames."""
    rep = w.autotune(frames_per_trial)
    if rep["frames"] < settle:
        w.step(settle - rep["frames"])
    return rep


def train_learned(w, log_frames):
    """Steps `log_frames` frames with the training log open, then trains the network on
    the logged GJK/EPA pairs and switches the world to the learned seeds.  Returns the
    report; with no logged sample the world keeps the plain seeds (value null)."""
    w.mlp_init(5)
    w.mlp_log(4_000_000)
    w.step(log_frames)
    w.synchronize()
    logged = w.mlp_logged(stop=True)
    if logged == 0:
        return {"value": None, "logged_samples": 0, "log_frames": log_frames,
                "reason": f"no GJK/EPA pair reached the training log in frames 1..{log_frames} (the exact rejection "
                          "tests removed every hull pair before GJK); the timed leg keeps the plain seeds"}
    t1 = time.perf_counter()
    w.mlp_train(600, 8192, 2e-3)
    loss = w.mlp_train(600, 8192, 5e-4)
    w.synchronize()
    w.mlp_mode("gjk")
    return {"logged_samples": logged, "log_frames": log_frames, "train_s": round(time.perf_counter() - t1, 3),
            "train_loss": round(loss, 4), "mode": "gjk"}


def plain_leg(world, sc, args, log_frames, pos_learned, value):
    """The plain GPU pipeline on a fresh world of the same scene: plain GJK/EPA, grid-wide
    colouring and per-colour solve launches, the same untimed and timed frames."""
    w = world.World(sc.timestep, sc.substeps).load(sc)
    w.set_option("joint_mode", 1)
    w.set_option("contact_mode", 1)
    w.set_option("scene_kernels", 0)   # grid-wide colouring rounds and per-colour launches
    w.step(log_frames)
    rep = tune_and_settle(w)
    dt = run_frames(w, args.steps, "plain")
    plain = sc.n_dynamic * sc.substeps * args.steps / dt
    d = np.abs(w.state()["position"] - pos_learned).max(axis=1)
    out = {"value": round(plain, 1), "unit": "body-substeps/s", "ms_per_step": round(dt / args.steps * 1e3, 4),
           "speedup": round(value / plain, 4), "tuned": rep,
           "pos_diff_vs_learned_m": {"max": float(d.max()), "p99": float(np.quantile(d, 0.99)),
                                     "median": float(np.median(d)), "frame": w.stats()["frames"]}}
    w.close()
    return out


def measured_traffic(kernel, workload="pile"):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy's readings), or None."""
    name = {"pile": f"{kernel}_pile_mixed_65k",
            "combined": {"gjk": "gjk_epa_rows_combined_1024_scenes",
                         "solve_pos": "scene_solve_pos_combined_1024_scenes"}.get(kernel)}.get(workload)
    if name is None:
        return None
    path = os.path.join(ROOT, "profiles", f"{name}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return round(json.load(f)["traffic_bytes_per_launch"], 1)


def roofline_of(tm, kernel):
    if not tm["launches"]:
        return None
    avg_s = tm["ms"] / 1e3 / tm["launches"]
    bytes_per_launch = tm["bytes"] / tm["launches"]
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "kernel": kernel, "avg_launch_us": round(avg_s * 1e6, 3),
            "launches": tm["launches"], "algorithmic_bytes_per_launch": round(bytes_per_launch, 1)}


def solve_roofline(w, frames=20):
    """The per-scene (or per-colour) position solve timed over `frames` frames after the
    timed region: contact bytes (296 B per point) per launch / launch duration."""
    w.set_option("timing_stride", 1)
    w.set_timing("solve_pos")
    w.step(frames)
    r = roofline_of(w.timing(), "k_scene_solve<pos> / k_solve_pos")
    w.set_timing("off")
    if r:
        r["traffic"] = measured_traffic("solve_pos", "combined")
        r["note"] = (f"frames after the timed region ({frames}); algorithmic bytes = 296 B per contact point "
                     "solved (contact row, two bodies gathered and written, lambda); traffic: rocprofv3 PMC "
                     "of the timed frames (it also carries the scenes' joint rows)")
    return r


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20)
    x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(10):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 10}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed frames (default: combined 1000, pile and hulls 500, batched 100)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed frames before the tuner (default 10; combined: at least 60, with the "
                         "training log open)")
    ap.add_argument("--timing-kernel", default=None,
                    choices=["off", "solve_pos", "solve_vel", "narrowphase", "integrate", "update_vel", "broadphase",
                             "colouring", "solve_pos_runs", "solve_vel_runs", "joints", "gjk"],
                    help="kernel family timed for the roofline (default: gjk for hulls and combined, else solve_pos)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the accuracy run against the reference fixture")
    ap.add_argument("--no-learned", action="store_true",
                    help="combined: time the plain GJK/EPA seeds (no training log)")
    ap.add_argument("--no-plain", action="store_true", help="combined: skip the plain leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every substep directly instead of replaying a captured hipGraph "
                         "(kernel-trace profiles: rocprofv3's kernel trace does not survive graph replays here)")
    ap.add_argument("--workload", default="combined", choices=["pile", "batched", "hulls", "combined"],
                    help="combined (default): 1,024 scenes of hulls, boxes and ragdolls (262,144 bodies), "
                         "learned + batched, with the plain pipeline beside it; pile: the 65,536-body mixed "
                         "pile; batched: 4,096 independent box-stack + capsule-chain scenes (1.1M bodies) in "
                         "one world; hulls: 32,768 random convex hulls in a dense pile (GJK/EPA)")
    ap.add_argument("--no-tune", action="store_true", help="combined: skip the autotuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB ten times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a --pmc run to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(independent scenes would shard by scene ranges with no per-substep traffic)", file=sys.stderr)
        sys.exit(2)

    from rbps_amd import scenes, world
    combined = args.workload == "combined"
    if args.timing_kernel is None:
        args.timing_kernel = "gjk" if args.workload in ("hulls", "combined") else "solve_pos"
    if args.steps is None:
        args.steps = {"combined": 1000, "batched": 100}.get(args.workload, 500)
    if args.warmup is None:
        args.warmup = 10
    if args.workload == "hulls":
        scene_fn = lambda: scenes.hull_pile(32768, seed=4)   # noqa: E731
        workload = HULL_WORKLOAD
    elif args.workload == "batched":
        scene_fn = scenes.batched_stack_chains
        workload = ("batched scenes: 4,096 independent scenes of 32-512 bodies (box stack + 16-capsule chain "
                    "on spherical joints, compliance 1e-6), 10 substeps, XPBD joint mode")
    elif combined:
        scene_fn = scenes.combined_batch
        workload = COMBINED_WORKLOAD
    else:
        scene_fn = scenes.pile_mixed
        workload = WORKLOAD
    log(f"generating the {args.workload} scene")
    sc = scene_fn()
    w = world.World(sc.timestep, sc.substeps).load(sc)
    options = {}
    if args.workload in ("batched", "combined"):
        options["joint_mode"] = 1
    if combined:   # capsule-capsule normals from A to B (ragdoll limbs)
        options["contact_mode"] = 1
    for k, v in options.items():
        w.set_option(k, v)
    if args.no_graph:
        w.set_option("graph", 0)

    learned, tuned = None, None
    log_frames = max(args.warmup, LOG_FRAMES) if combined else args.warmup
    if combined and not args.no_learned:
        log(f"{log_frames} untimed frames with the training log open")
        try:
            learned = train_learned(w, log_frames)
        except world.RbgpuError as e:   # the timed leg then runs the plain seeds
            w.mlp_mode("plain")
            learned = {"value": None, "reason": f"training failed: {e}"}
    else:
        w.step(log_frames)
    if combined and not args.no_tune:
        log("autotuning")
        tuned = tune_and_settle(w)
        if args.no_graph:
            w.set_option("graph", 0)
    untimed = w.stats()["frames"]
    state0 = w.state() if not args.no_cpu_baseline else None
    # per-launch HIP events on every 10th frame (those frames run without the substep graph):
    # events between launches cost time themselves
    w.set_option("timing_stride", 10)
    w.set_timing(args.timing_kernel)
    label = "learned" if learned and learned.get("mode") else ""
    dt = run_frames(w, args.steps, label)
    tm = w.timing()
    w.set_timing("off")
    st = w.stats()
    value = sc.n_dynamic * sc.substeps * args.steps / dt

    roof = roofline_of(tm, args.timing_kernel) if args.timing_kernel != "off" else None
    if roof:
        roof["traffic"] = measured_traffic(args.timing_kernel, args.workload)
    if roof and args.timing_kernel == "gjk":
        roof["kernel"] = "k_gjk_epa_rows"
        roof["note"] = ("GJK, then EPA and the manifold for the overlapping pairs, on the candidates that pass the "
                        "exact rejection tests, one pair per 16-lane row; algorithmic bytes = candidates (device "
                        "count) x (2 x (16 B x mean hull vertices + 164 B) + 40 B), the inputs (the contacts written "
                        "for the overlapping pairs are not counted); VALU-issue bound, not HBM bound")
    out = {
        "metric": METRIC,
        "value": round(value, 1), "unit": "body-substeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": workload, "scene": sc.name, "bodies": sc.n_dynamic, "substeps": sc.substeps,
                   "timestep": sc.timestep, "untimed_frames": untimed,
                   "shapes": "convex hulls + static plane" if args.workload == "hulls" else "box/sphere/capsule + static plane",
                   "last_frame": st},
        "roofline": roof,
    }
    if args.workload in ("batched", "combined") and args.timing_kernel != "solve_pos":
        out["solve_roofline"] = solve_roofline(w)
    if combined:
        out["config"]["shapes"] = "convex hulls, boxes, ragdolls (boxes, sphere, capsules) + static planes"
        out["config"]["scenes"] = int(sc.meta["scenes"])
        out["config"]["joints"] = int(len(sc.joints))
        out["learned"] = learned
        out["tuned"] = tuned
        if not args.no_plain:
            log("plain leg")
            out["plain"] = plain_leg(world, sc, args, log_frames, w.state()["position"], value)
    w.close()
    if not args.no_accuracy:
        log("accuracy run (300 frames of the well-posed layers)")
        out["accuracy"] = accuracy(world, scenes, {k: v for k, v in options.items()})
    if not args.no_cpu_baseline:
        log("CPU baseline")
        _, _, threads = host_cpu()
        if args.workload in ("batched", "combined"):
            parts = scenes.split(sc, only=range(min(int(sc.meta["scenes"]), 8 * threads)))
            out["cpu_baseline"] = cpu_baseline(parts, split_states(sc, state0, parts), threads,
                                               joint_mode=options.get("joint_mode", 0),
                                               contact_mode=options.get("contact_mode", 0),
                                               note=" (same options as the GPU legs: XPBD joint rows"
                                                    + (", capsule-capsule normals from A to B" if combined else "")
                                                    + "; the reference itself has no spherical joint)")
        else:
            out["cpu_baseline"] = cpu_baseline([sc], [state0], 1, note=" (one World: the reference is serial)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
