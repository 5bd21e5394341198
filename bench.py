"""Throughput of the XPBD substep pipeline on one MI355X: body-substeps/s.

Default workload (--workload combined): 262,144 dynamic bodies in 1,024 independent
scenes of 64-512 bodies (convex hulls from a shared library, boxes, 8-segment ragdolls on
revolute + spherical joints with 60 deg limits) in one world, dt = 1/120 s, 8 substeps,
FP32 state, XPBD joint rows.  The timed leg is the learned + batched pipeline: GJK/EPA
warm-started by the network (trained on the GJK/EPA pairs of the warm-up frames), one
workgroup per scene for colouring and solve, each substep replayed from a hipGraph, launch
options picked by the autotuner during warm-up.  A "step" is one World::step (one frame);
--steps frames are timed after the warm-up, inputs resident in HBM.
value = dynamic bodies * substeps * steps / seconds.  Beside it ("plain"): the plain GPU
pipeline (plain GJK/EPA, grid-wide colouring, per-colour launches, autotuned too) on a
fresh world of the same scene over the same frames.

Other workloads: pile (65,536 mixed convex bodies, the plain pipeline), batched (4,096
box-stack + chain scenes), hulls (32,768 random hulls, GJK/EPA).

Also reported:
  roofline     — the dominant kernel's algorithmic bytes per launch / its average
                 launch duration (HIP events on the world's stream around every
                 launch of that kernel in every 10th frame of the timed region),
                 against 8 TB/s HBM.
  cpu_baseline — the reference engine (compiled from its sources into
                 oracle/_ref/libref.so; FP64, serial, 1 core) or, where it cannot run the
                 scene (spherical joints), the oracle restatement, on a bounded sample.
"""
import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "body-substeps/s on 1 MI355X (learned+batched vs plain), pos err \u22641e-4 m"
WORKLOAD = "plain GPU pipeline, large pile: 65,536 mixed convex bodies, 10 substeps, 500 steps"
COMBINED_WORKLOAD = ("combined at scale: 262,144 bodies in 1,024 uneven scenes (64-512 bodies) of convex hulls, "
                     "boxes and 8-segment ragdolls (revolute + spherical joints, limits +-60 deg), dt 1/120 s, "
                     "8 substeps, XPBD joint rows, capsule-capsule normals from A to B, learned GJK/EPA warm start")
HULL_WORKLOAD = ("GJK/EPA-heavy: 32,768 random convex hulls (16-64 vertices on ellipsoids with semi-axes "
                 "U(0.2,0.8), triangulated) in a dense pile over a plane, 10 substeps, plain GJK/EPA")


def _cpu_world(sc):
    """The reference engine where it can build the scene, else the oracle restatement in
    the reference's order (FP64); the reference has no spherical joint."""
    from oracle import orc
    kind = "reference" if orc.ref_available() and not (sc.joints["type"] == 3).any() else "port"
    w = orc.RefWorld(sc.timestep, sc.substeps) if kind == "reference" else orc.OracleWorld(sc.timestep, sc.substeps)
    return w.load(sc), kind


def cpu_baseline(scene_fn, budget_s=10.0, max_frames=30):
    sc = scene_fn()
    w, kind = _cpu_world(sc)
    t0 = time.perf_counter()
    frames = 0
    while frames < max_frames:
        w.step(1)
        frames += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": sc.n_dynamic * sc.substeps * frames / dt, "unit": "body-substeps/s", "cores": 1, "kind": kind,
            "sample": f"frames 1..{frames} of the same {sc.n_dynamic}-body scene ({dt:.1f} s, FP64, serial)"}


def cpu_baseline_batched(budget_s=10.0, frames=10, combined=False):
    """The reference runs one World per scene: scenes of the same batch (same seeds),
    each for `frames` frames, one after another on one core, until the budget is spent."""
    from rbps_amd import scenes
    if combined:
        counts = scenes.combined_counts()
        lib = scenes.hull_library()
        make = lambda k: scenes.combined_scene(int(counts[k]), seed=2026 + 1 + k, library=lib)   # noqa: E731
    else:
        counts = np.random.default_rng(77).integers(32, 513, 4096)
        make = lambda k: scenes.stack_chain(int(counts[k]), seed=77 + 1 + k)   # noqa: E731
    t0 = time.perf_counter()
    work = done = 0
    while time.perf_counter() - t0 < budget_s and done < len(counts):
        sc = make(done)
        w, kind = _cpu_world(sc)
        if combined:
            w.set_joint_mode(1)
            w.set_contact_mode(1)
        w.step(frames)
        work += sc.n_dynamic * sc.substeps * frames
        done += 1
    dt = time.perf_counter() - t0
    return {"value": work / dt, "unit": "body-substeps/s", "cores": 1, "kind": kind,
            "sample": f"scenes 0..{done - 1} of the batch, frames 1..{frames} each, one World per scene "
                      f"({dt:.1f} s incl. set-up, FP64, serial; the oracle restatement in the "
                      f"reference's order because the reference has no spherical joint"
                      f"{'; XPBD joint rows and capsule-capsule normals from A to B as on the GPU' if combined else ''})"}


def learned_report(w, sc, train_frames=20, capacity=4_000_000, steps=1200, batch=8192):
    """The learned warm start on the hull workload (after the timed plain run): log the
    GJK pairs of `train_frames` frames, train, compare on the current state's GJK pairs
    the iterations and contacts of the plain and learned seeds, then time frames with the
    learned seeds.  (Errors against FP64 are measured by tests/test_gpu_mlp.py.)"""
    t0 = time.perf_counter()
    w.mlp_init(5)
    w.mlp_log(capacity)
    w.step(train_frames)
    logged = w.mlp_logged(stop=True)
    t1 = time.perf_counter()
    w.mlp_train(steps // 2, batch, 2e-3)
    loss = w.mlp_train(steps - steps // 2, batch, 5e-4)
    w.synchronize()
    t2 = time.perf_counter()
    pairs = w.sap_pairs()
    k = sc.colliders["shape"][pairs]
    pairs = pairs[(k == 9).any(1) & ~(k == 6).any(1)]
    w.mlp_mode("plain")
    rp = w.collide_pairs(pairs)
    w.mlp_mode("gjk")
    rl = w.collide_pairs(pairs)
    both = (rp["n"] > 0) & (rl["n"] > 0)
    cos = np.einsum("ij,ij->i", rp["normal"][both], rl["normal"][both])
    dd = np.abs(rp["depth"][both, 0] - rl["depth"][both, 0])
    frames = 20
    res = {}
    for mode in ("plain", "gjk"):
        w.mlp_mode(mode)
        w.synchronize()
        t3 = time.perf_counter()
        w.step(frames)
        w.synchronize()
        res[mode] = sc.n_dynamic * sc.substeps * frames / (time.perf_counter() - t3)
    w.mlp_mode("plain")
    it = lambda r: float((r["gjk_iters"] + r["epa_iters"]).mean())   # noqa: E731
    return {"value": round(res["gjk"], 1), "plain_same_frames": round(res["plain"], 1), "unit": "body-substeps/s",
            "frames_timed": frames, "logged_samples": logged, "log_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
            "train_loss": round(loss, 4), "gjk_pairs": int(len(pairs)),
            "iterations_per_pair": {"plain": round(it(rp), 4), "learned": round(it(rl), 4)},
            "fallbacks": int(rl["fallback"].sum()), "contacts_agree": float(np.mean((rp["n"] > 0) == (rl["n"] > 0))),
            "normal_deg_max_vs_plain": round(float(np.degrees(np.arccos(np.clip(cos.min(), -1, 1)))), 5) if both.any() else None,
            "depth_abs_max_vs_plain": float(dd.max()) if both.any() else None}


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
        print(f"bench: {label + ' ' if label else ''}{done}/{steps} frames", file=sys.stderr, flush=True)
    w.synchronize()
    return time.perf_counter() - t0


def tune_and_align(w, start, frames_per_trial=3, settle=60):
    """Autotune (result-neutral options, real frames), then step up to frame start + settle
    so that differently tuned legs time the same frames."""
    rep = w.autotune(frames_per_trial)
    if rep["frames"] < settle:
        w.step(settle - rep["frames"])
    return rep


def plain_leg(world, sc, args, pos_learned, value):
    """The plain GPU pipeline on a fresh world of the same scene: plain GJK/EPA, grid-wide
    colouring and per-colour solve launches, the same warm-up and timed frames."""
    w = world.World(sc.timestep, sc.substeps).load(sc)
    w.set_option("joint_mode", 1)
    w.set_option("contact_mode", 1)
    w.set_option("scene_kernels", 0)   # grid-wide colouring rounds and per-colour launches
    w.step(args.warmup)
    rep = tune_and_align(w, args.warmup)
    dt = run_frames(w, args.steps, "plain")
    plain = sc.n_dynamic * sc.substeps * args.steps / dt
    d = np.abs(w.state()["position"] - pos_learned).max(axis=1)
    return {"value": round(plain, 1), "unit": "body-substeps/s", "ms_per_step": round(dt / args.steps * 1e3, 4),
            "speedup": round(value / plain, 4), "tuned": rep,
            "pos_diff_vs_learned_m": {"max": float(d.max()), "p99": float(np.quantile(d, 0.99)),
                                      "median": float(np.median(d)), "frame": w.stats()["frames"]}}


def measured_traffic(kernel, workload="pile"):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy's readings), or None."""
    name = {"pile": f"{kernel}_pile_mixed_65k", "combined": "gjk_pass_combined_1024_scenes"}.get(workload)
    if name is None or (workload == "combined" and kernel != "gjk"):
        return None
    path = os.path.join(ROOT, "profiles", f"{name}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return round(json.load(f)["traffic_bytes_per_launch"], 1)


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
                    help="untimed frames before (default: combined 60 + tuning, others 10)")
    ap.add_argument("--timing-kernel", default=None,
                    choices=["off", "solve_pos", "solve_vel", "narrowphase", "integrate", "update_vel", "broadphase", "colouring",
                             "solve_pos_runs", "solve_vel_runs", "joints", "gjk"],
                    help="kernel family timed for the roofline (default: gjk for hulls and combined, else solve_pos)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="hulls: skip the learned warm-start report; combined: time the plain GJK/EPA")
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
    if args.timing_kernel is None:
        args.timing_kernel = "gjk" if args.workload in ("hulls", "combined") else "solve_pos"
    if args.steps is None:
        args.steps = {"combined": 1000, "batched": 100}.get(args.workload, 500)
    if args.warmup is None:
        args.warmup = 60 if args.workload == "combined" else 10
    if args.workload == "hulls":
        scene_fn = lambda: scenes.hull_pile(32768, seed=4)   # noqa: E731
        workload = HULL_WORKLOAD
    elif args.workload == "batched":
        scene_fn = scenes.batched_stack_chains
        workload = ("batched scenes: 4,096 independent scenes of 32-512 bodies (box stack + 16-capsule chain "
                    "on spherical joints, compliance 1e-6), 10 substeps, XPBD joint mode")
    elif args.workload == "combined":
        scene_fn = scenes.combined_batch
        workload = COMBINED_WORKLOAD
    else:
        scene_fn = scenes.pile_mixed
        workload = WORKLOAD
    sc = scene_fn()
    w = world.World(sc.timestep, sc.substeps).load(sc)
    if args.workload in ("batched", "combined"):
        w.set_option("joint_mode", 1)
    if args.workload == "combined":   # capsule-capsule normals from A to B (ragdoll limbs)
        w.set_option("contact_mode", 1)
    if args.no_graph:
        w.set_option("graph", 0)
    learned = args.workload == "combined" and not args.no_learned
    if learned:   # train on the GJK/EPA pairs of the warm-up frames, then warm-start from the network
        w.mlp_init(5)
        w.mlp_log(4_000_000)
    w.step(args.warmup)
    w.synchronize()
    train = None
    if learned:
        logged = w.mlp_logged(stop=True)
        t1 = time.perf_counter()
        w.mlp_train(600, 8192, 2e-3)
        loss = w.mlp_train(600, 8192, 5e-4)
        w.synchronize()
        train = {"logged_samples": logged, "train_s": round(time.perf_counter() - t1, 3), "train_loss": round(loss, 4)}
        w.mlp_mode("gjk")
    tuned = None
    if args.workload == "combined" and not args.no_tune:
        tuned = tune_and_align(w, args.warmup)
        if args.no_graph:
            w.set_option("graph", 0)
    # per-launch HIP events on every 10th frame (those frames run without the substep graph):
    # events between launches cost time themselves
    w.set_option("timing_stride", 10)
    w.set_timing(args.timing_kernel)
    dt = run_frames(w, args.steps, "learned" if learned else "")
    tm = w.timing()
    st = w.stats()
    value = sc.n_dynamic * sc.substeps * args.steps / dt

    roof = None
    if args.timing_kernel != "off" and tm["launches"]:
        avg_s = tm["ms"] / 1e3 / tm["launches"]
        bytes_per_launch = tm["bytes"] / tm["launches"]
        achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(args.timing_kernel, args.workload),
                "kernel": args.timing_kernel,
                "avg_launch_us": round(avg_s * 1e6, 3), "launches": tm["launches"],
                "algorithmic_bytes_per_launch": round(bytes_per_launch, 1)}

    if roof and args.timing_kernel == "gjk":
        roof["kernel"] = "k_gjk_pass"
        roof["note"] = ("GJK on the candidates that pass the exact rejection tests, one wavefront per pair; "
                        "algorithmic bytes = candidates (device count) x (2 x (16 B x mean hull vertices + 164 B) + 40 B)")
    out = {
        "metric": METRIC,
        "value": round(value, 1), "unit": "body-substeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": workload, "scene": sc.name, "bodies": sc.n_dynamic, "substeps": sc.substeps, "timestep": sc.timestep,
                   "shapes": "convex hulls + static plane" if args.workload == "hulls" else "box/sphere/capsule + static plane",
                   "last_frame": st},
        "roofline": roof if args.workload in ("pile", "combined") else dict(roof or {}, traffic=None),
    }

    if args.workload == "hulls" and not args.no_learned:
        out["learned"] = learned_report(w, sc)
    if args.workload == "combined":
        out["config"]["shapes"] = "convex hulls, boxes, ragdolls (boxes, sphere, capsules) + static planes"
        out["config"]["scenes"] = int(sc.meta["scenes"])
        out["config"]["joints"] = int(len(sc.joints))
        out["learned"] = train
        out["tuned"] = tuned
        if not args.no_plain:
            out["plain"] = plain_leg(world, sc, args, w.state()["position"], value)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = (cpu_baseline_batched(combined=args.workload == "combined")
                               if args.workload in ("batched", "combined") else cpu_baseline(lambda: sc))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
