"""TreePM force-evaluation benchmark on one MI355X.

Workload (default): 512^3 particles on a 512^3 mesh in a periodic 100 Mpc/h box, Zel'dovich
ICs at z = 49 from a seeded Gaussian field with P(k) ~ k^-2 exp(-k / 2 h/Mpc) and a 0.3-cell
3D displacement RMS, split r_s = 1.25 cells, r_cut = 4.5 r_s, theta = 0.35, softening L/(30*512).
A step is one full TreePM force evaluation (deposit -> r2c -> Green/gradient -> c2r -> CIC
interpolation -> Morton sort/octree -> short-range walk) on positions resident in HBM.

Prints one JSON line: particle-force evaluations/s, force error vs an FP64 TreePM judge
(theta = 0.2) on a seeded 16,384-particle subsample, the roofline of the dominant kernel and a
CPU baseline (the oracle's restatement of the reference algorithm on this host's cores).
"""
import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cosmological-particle-simulation_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import numpy as np  # noqa: E402

OMEGA_M = 0.315
BOX = 100.0
Z_INIT = 49.0
K_C = 2.0
FLOP_PER_INTERACTION = 68  # project convention for one pair / node interaction
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n_side=128, theta=0.35, threads=None):
    """Times the oracle (reference algorithm restated in C, OpenMP where the reference uses
    rayon, serial where it is serial) on one force evaluation of the 64^3 configuration."""
    import oracle
    import treepm

    threads = threads or len(os.sched_getaffinity(0))
    threads = min(threads, 16)
    oracle.set_threads(threads)
    with treepm.Solver(n_side, BOX, OMEGA_M, max_particles=n_side ** 3) as ic:
        pos, _ = ic.zeldovich(1, K_C, 0.3 * BOX / n_side)
    pos = pos.astype(np.float32)
    split = oracle.force_split(BOX, n_side, n_side, 1 / 30, theta)
    a = 1.0 / (1.0 + Z_INIT)
    t0 = time.perf_counter()
    oracle.accelerations(pos, n_side, BOX, OMEGA_M, a, split)
    dt = time.perf_counter() - t0
    return {"value": n_side ** 3 / dt, "unit": "particle-force evals/s", "cores": threads, "kind": "port",
            "sample": f"one full TreePM force evaluation, {n_side}^3 particles on a {n_side}^3 mesh "
                      f"(same generator and split; per-particle walk cost is size-independent at "
                      f"mesh = particle grid), {dt:.2f} s wall"}


def walk_traffic(n, workload, tuning):
    """HBM bytes per k_group_walk launch from the committed rocprofv3 counter summary of this
    workload (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads for a copy of known
    size), or None when this run is not that workload."""
    path = os.path.join(REPO, "profiles", "k_group_walk_pmc_zeldovich_512cubed.md")
    if n != 512 or workload != "zeldovich" or tuning is None or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def judge_error(solver, pos, acc, npart, a, m=16384, seed=12345):
    """RMS / p99 relative error of acc (device) against the FP64 judge on a seeded subsample."""
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(npart, size=min(m, npart), replace=False)).astype(np.uint32)
    host_pos = pos.numpy().reshape(-1, 3)
    host_acc = acc.numpy().reshape(-1, 3)
    ref = solver.judge(host_pos, idx, a, theta=0.2)
    d = host_acc[idx].astype(np.float64) - ref
    rel = np.sqrt((d ** 2).sum(1)) / np.sqrt((ref ** 2).sum(1))
    return {"rms": float(np.sqrt(np.mean(rel ** 2))), "p99": float(np.percentile(rel, 99)),
            "median": float(np.median(rel)), "sample": int(len(idx))}


def learned_vs_plain(args):
    """Learned short-range correction on the clustered workload: train on walk-generated targets,
    then compare the learned form (MLP + walk fallback above tol) with the plain TreePM."""
    import treepm

    n = args.learned_size
    npart = n ** 3
    a = 1.0
    split = treepm.force_split(BOX, n, n, 1 / 30, args.theta)
    s = treepm.Solver(n, BOX, OMEGA_M, split=split, max_particles=npart)
    pos = treepm.DeviceArray(3 * npart, np.float64)
    acc = treepm.DeviceArray(3 * npart, np.float32)
    s.nfw_halos_dev(treepm.NfwSpec.clustered(seed=1), npart, pos.ptr)
    s.sync()
    log(f"[bench] learned: training on {args.learned_samples} walk samples, {args.learned_epochs} epochs")
    s.mlp_init(1)
    t0 = time.perf_counter()
    hist = s.mlp_train_dev(pos.ptr, npart, a, args.learned_samples, args.learned_epochs, batch=8192, lr=1e-3, seed=1)
    s.sync()
    train_s = time.perf_counter() - t0
    log(f"[bench] learned: trained in {train_s:.1f} s, loss {list(hist)}")
    # plain TreePM, with the knobs tuned on this input (the learned form runs with the same knobs)
    s.accelerations_dev(pos.ptr, npart, a, acc.ptr)
    s.sync()
    tuning = s.autotune(pos.ptr, npart, a, acc.ptr).as_dict()
    s.sync()
    log(f"[bench] learned: plain TreePM tuning {tuning}")
    t0 = time.perf_counter()
    s.accelerations_dev(pos.ptr, npart, a, acc.ptr)
    s.sync()
    ms_plain = 1000 * (time.perf_counter() - t0)
    err_plain = judge_error(s, pos, acc, npart, a)
    # learned form
    s.accelerations_learned_dev(pos.ptr, npart, a, args.learned_tol, acc.ptr)
    s.sync()
    t0 = time.perf_counter()
    nf = s.accelerations_learned_dev(pos.ptr, npart, a, args.learned_tol, acc.ptr)
    s.sync()
    ms_learned = 1000 * (time.perf_counter() - t0)
    err_learned = judge_error(s, pos, acc, npart, a)
    s.close()
    pos.free()
    acc.free()
    return {"workload": f"treepm_nfw20000_{n}cubed_mesh{n}_z0", "train_samples": int(min(args.learned_samples, npart)),
            "epochs": args.learned_epochs, "train_s": train_s, "loss_history": [float(v) for v in hist],
            "tol": args.learned_tol, "fallback_fraction": nf / npart, "ms_plain": ms_plain,
            "ms_learned": ms_learned, "speedup": ms_plain / ms_learned, "tuning": tuning, "error_plain": err_plain,
            "error_learned": err_learned}


def ensemble_bench(args):
    """Batched ensemble: independent periodic boxes with sides drawn from {32,48,64,96,128}
    (mesh = particle grid), per-box seeded Zel'dovich ICs at z = 49 (momenta a^2 f E Psi), the
    initial force plus KDK steps to z = 9, autotuned, run by one HIP graph launch."""
    import treepm

    rng = np.random.default_rng(4)
    sides = [int(v) for v in rng.choice([32, 48, 64, 96, 128], size=args.ensemble_boxes)]
    total = sum(s ** 3 for s in sides)
    a0, a1 = 1.0 / (1.0 + Z_INIT), 0.1
    steps = args.ensemble_steps
    log(f"[bench] ensemble: {len(sides)} boxes, {total} particles, {steps} KDK steps")
    t0 = time.perf_counter()
    e = treepm.Ensemble(sides, box_size=BOX, omega_m=OMEGA_M, softening_scale=1 / 30, theta=args.ensemble_theta)
    e.ics([1000 + b for b in range(len(sides))], k_c=K_C, disp_cells=0.3,
          mom_factor=treepm.zeldovich_momentum_factor(a0))
    setup_s = time.perf_counter() - t0
    ms_eager_default = e.run(a0, a1, steps, graph=False)
    log(f"[bench] ensemble: eager run (default knobs) {ms_eager_default:.1f} ms")
    t0 = time.perf_counter()
    tuning = e.autotune(a0)
    tune_s = time.perf_counter() - t0
    ms_eager = e.run(a0, a1, steps, graph=False)
    log(f"[bench] ensemble: eager run (tuned) {ms_eager:.1f} ms; tuning {tune_s:.1f} s")
    t0 = time.perf_counter()
    e.run(a0, a1, steps, graph=True)
    capture_s = time.perf_counter() - t0
    ms_graph = min(e.run(a0, a1, steps, graph=True) for _ in range(2))
    log(f"[bench] ensemble: graph run {ms_graph:.1f} ms (first launch incl. capture {capture_s:.1f} s)")
    choices = {}
    for side, t in zip(sides, tuning):
        choices[str(side)] = {"fft_path": t["fft_path"], "fft_cap": t["fft_cap"], "tree_depth": t["tree_depth"],
                              "group_level_shift": t["group_level_shift"], "small_node_direct": t["small_node_direct"],
                              "ms_default": round(t["ms_default"], 3), "ms_tuned": round(t["ms_tuned"], 3)}
    e.close()
    evals = total * (steps + 1)
    return {"workload": f"ensemble_{len(sides)}boxes_zeldovich_z49_to_z9_{steps}kdk",
            "boxes": len(sides), "sides": {str(v): sides.count(v) for v in sorted(set(sides))},
            "particles": total, "kdk_steps": steps, "theta": args.ensemble_theta,
            "force_evals": evals, "ms_graph": ms_graph, "ms_eager_tuned": ms_eager,
            "ms_eager_default": ms_eager_default, "evals_per_s_graph": evals / (ms_graph / 1000.0),
            "graph_vs_eager_default": ms_eager_default / ms_graph, "setup_s": setup_s, "tune_s": tune_s,
            "first_graph_launch_s": capture_s, "tuning": choices}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512, help="particles per side (= mesh cells per side)")
    ap.add_argument("--workload", choices=["zeldovich", "nfw"], default="zeldovich",
                    help="zeldovich: lattice ICs at z = 49; nfw: 20,000 NFW halos + 30%% background at z = 0")
    ap.add_argument("--theta", type=float, default=0.35)
    ap.add_argument("--judge-sample", type=int, default=16384)
    ap.add_argument("--no-judge", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-size", type=int, default=128)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-size", type=int, default=512)
    ap.add_argument("--learned-samples", type=int, default=20_000_000)
    ap.add_argument("--learned-epochs", type=int, default=3)
    ap.add_argument("--learned-tol", type=float, default=0.005)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default knobs)")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer (PCIe-inclusive) evaluation")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-boxes", type=int, default=96)
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--ensemble-theta", type=float, default=0.5)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device three times (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.copy_probe:
        import treepm
        src = treepm.DeviceArray(args.copy_probe, np.uint8)
        dst = treepm.DeviceArray(args.copy_probe, np.uint8)
        for _ in range(3):
            if treepm.load().tpm_memcpy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), args.copy_probe, 2) != 0:
                raise RuntimeError("device copy failed")
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "copies": 3}), flush=True)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import treepm

    n = args.size
    npart = n ** 3
    a = 1.0 / (1.0 + Z_INIT) if args.workload == "zeldovich" else 1.0
    split = treepm.force_split(BOX, n, n, 1 / 30, args.theta)
    log(f"[bench] {npart} particles, mesh {n}^3, theta {args.theta}")

    solver = treepm.Solver(n, BOX, OMEGA_M, split=split, max_particles=npart)
    pos = treepm.DeviceArray(3 * npart, np.float64)
    acc = treepm.DeviceArray(3 * npart, np.float32)
    if args.workload == "zeldovich":
        solver.zeldovich_dev(1, K_C, 0.3 * BOX / n, 0.0, pos.ptr)
    else:
        solver.nfw_halos_dev(treepm.NfwSpec.clustered(seed=1), npart, pos.ptr)
    solver.sync()
    log("[bench] ICs generated")

    for i in range(args.warmup):
        solver.accelerations_dev(pos.ptr, npart, a, acc.ptr)
        solver.sync()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    tuning = None
    if not args.no_tune:
        tuning = solver.autotune(pos.ptr, npart, a, acc.ptr).as_dict()
        solver.sync()
        log(f"[bench] autotune: {tuning}")

    solver.set_timing(True)
    # One evaluation with the walk's interaction counting on gives the interaction total (a function
    # of the input alone); the timed evaluations run without the per-pair counter.
    solver.accelerations_dev(pos.ptr, npart, a, acc.ptr)
    counted = solver.stats()
    log(f"[bench] counted evaluation: {counted['interactions']:.4e} interactions, walk {counted['ms_walk_kernel']:.1f} ms")
    solver.set_counting(False)
    stats = []
    solver.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        solver.accelerations_dev(pos.ptr, npart, a, acc.ptr)
        stats.append(solver.stats())
        log(f"[bench] step {i + 1}/{args.steps}: {stats[-1]['ms_total']:.1f} ms")
    solver.sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    solver.set_timing(False)
    solver.set_counting(True)

    def mean(key):
        return float(np.mean([s[key] for s in stats]))

    stages = {k: round(mean(k), 3) for k in ("ms_sort", "ms_deposit", "ms_fft", "ms_interp", "ms_tree", "ms_walk",
                                               "ms_total")}
    inter = float(counted["interactions"])
    walk_s = mean("ms_walk_kernel") / 1000.0  # the k_group_walk launch alone (HIP events on its stream)
    achieved = inter * FLOP_PER_INTERACTION / walk_s / 1e12 if walk_s > 0 else 0.0
    traffic = walk_traffic(n, args.workload, tuning)

    result = {
        "metric": "TreePM particle-force evals/s at 512^3 + RMS force err vs FP64; learned vs plain",
        "value": npart / (ms_per_step / 1000.0),
        "unit": "particle-force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"treepm_zeldovich_{n}cubed_mesh{n}_z49" if args.workload == "zeldovich"
                                else f"treepm_nfw20000_{n}cubed_mesh{n}_z0"), "particles": npart, "mesh": n,
                   "box": BOX, "theta": args.theta, "r_split_cells": 1.25, "r_cut_over_r_split": 4.5,
                   "softening": split["softening"], "positions": "f64", "pair_and_mesh_arithmetic": "f32"},
        "tuning": tuning,
        "stages_ms": stages,
        "interactions_per_particle": inter / npart,
        "walk_sources_per_group": {"groups": mean("groups"), "listed": mean("listed") / max(mean("groups"), 1),
                                   "kept": mean("kept") / max(mean("groups"), 1)},
        "roofline": {"bound": "mfma", "kernel": "k_group_walk (FP32 VALU pair arithmetic)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "walk_kernel_ms": round(mean("ms_walk_kernel"), 3)},
    }

    if not args.no_host:
        # The host-buffer entry point (the Rust binding's path): positions copied in over PCIe,
        # accelerations copied back, same evaluation. Reported beside `value`, never as it.
        host_pos = pos.numpy().reshape(-1, 3)
        solver.accelerations(host_pos, a)
        t0 = time.perf_counter()
        solver.accelerations(host_pos, a)
        host_s = time.perf_counter() - t0
        result["host_buffers"] = {"evals_per_s": npart / host_s, "ms_per_eval": 1000.0 * host_s,
                                  "entry": "tpm_accelerations_f64 (FP64 positions in, FP32 accelerations out, "
                                           "pageable host memory)"}
        log(f"[bench] host-buffer evaluation {1000.0 * host_s:.1f} ms")
        del host_pos

    if not args.no_judge:
        m = min(args.judge_sample, npart)
        rng = np.random.default_rng(12345)
        idx = np.sort(rng.choice(npart, size=m, replace=False)).astype(np.uint32)
        host_pos = pos.numpy().reshape(-1, 3)
        host_acc = acc.numpy().reshape(-1, 3)
        log("[bench] FP64 judge ...")
        ref = solver.judge(host_pos, idx, a, theta=0.2)
        d = host_acc[idx].astype(np.float64) - ref
        rel = np.sqrt((d ** 2).sum(1)) / np.sqrt((ref ** 2).sum(1))
        result["force_error"] = {"rms": float(np.sqrt(np.mean(rel ** 2))), "p99": float(np.percentile(rel, 99)),
                                 "median": float(np.median(rel)), "sample": int(m),
                                 "judge": "FP64 TreePM, theta 0.2, per-particle walk"}
        log(f"[bench] force error rms {result['force_error']['rms']:.3e} p99 {result['force_error']['p99']:.3e}")

    solver.close()
    pos.free()
    acc.free()

    if not args.no_learned:
        try:
            result["learned_vs_plain"] = learned_vs_plain(args)
            lv = result["learned_vs_plain"]
            log(f"[bench] learned: fallback {lv['fallback_fraction']:.3f}, speedup {lv['speedup']:.3f}, "
                f"error {lv['error_learned']['rms']:.3e} vs plain {lv['error_plain']['rms']:.3e}")
        except Exception as e:  # reported, not hidden: the line still carries the plain result
            result["learned_vs_plain"] = {"error": repr(e)}

    if not args.no_ensemble:
        try:
            result["ensemble"] = ensemble_bench(args)
        except Exception as e:  # reported, not hidden: the line still carries the plain result
            result["ensemble"] = {"error": repr(e)}

    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(args.cpu_size, theta=args.theta)
        log(f"[bench] CPU baseline {result['cpu_baseline']['value']:.3e} evals/s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
