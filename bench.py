"""Throughput of the Monte Carlo fault-tree engine on one MI355X.

Workload: plain Monte Carlo on the large synthetic layered tree (25,000 basic
events with p log-uniform in [1e-6, 1e-2]; 40,000 gates in 30 layers, fan-in
2-8, 48 % AND / 47 % OR / 5 % ATLEAST with n up to 16, 10 % negated inputs),
every gate tallied, 2^18 uint64 words = 2^24 trials per pass. One step is one
pass (Philox sampling of every basic event, every gate layer, tallies);
2^32 trials are 256 steps.

Prints one JSON line: gate-evals/s (gates x trials / s), the HBM roofline of
the dominant kernel measured with HIP events on the engine's stream, and the
CPU oracle timed on the host on a bounded sample of the same tree (with a
bit-exact tally comparison against the GPU on that sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(pd, n_bp: int) -> dict:
    """Bytes the materialised dataflow moves per pass: each node word written
    once, each gate input word read once, each tallied word read once."""
    n_be, n_g, f = pd.n_vars, pd.n_gates, pd.n_args
    return {"sampler": 8 * n_be * n_bp,
            "gates": 8 * (n_g + f) * n_bp,
            "pass": 8 * (n_be + n_g + f + n_g) * n_bp}


def cpu_baseline(pd, target_s: float, max_words: int):
    from oracle import binding as orc
    tally = pd.gate_index.astype(np.int32)
    t = time.perf_counter()
    orc.run(pd, 256, tally=tally)
    per_word = (time.perf_counter() - t) / 256
    words = 256
    while words * 2 <= max_words and per_word * words * 2 <= target_s:
        words *= 2
    t = time.perf_counter()
    ones, _ = orc.run(pd, words, tally=tally)
    dt = time.perf_counter() - t
    return words, ones, dt, orc.max_threads()


def host_cpu() -> dict:
    """The host's CPU model and logical CPU count (the CPU baseline's machine)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count()}


def measured_traffic(tree: str, n_bp: int):
    """HBM bytes per launch of pass_layer from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (scaled by the copy
    calibration recorded there), or None when no summary matches."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"pass_layer_{tree}_tree_hbm_counters.json")
    try:
        with open(path) as f:
            s = json.load(f)
    except OSError:
        return None
    if s.get("words_per_pass") != n_bp or s.get("tree") != tree:
        return None
    return s["traffic_bytes_per_launch"]


def copy_calibration(n_bytes: int) -> dict:
    """One device-to-device copy of n_bytes through the HIP runtime's blit
    kernel: the known-size reference that scales HBM byte counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n_bytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(n_bytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(dst, src, C.c_size_t(n_bytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    return {"read": n_bytes, "write": n_bytes}


def bench_fused(args, mc) -> int:
    """Batched uneven trees (or the small tree alone) in the persistent
    whole-tree-in-LDS kernel; one step = one pass over every tree."""
    from oracle import binding as orc
    if args.workload == "batched":
        trees = mc.batched_trees(seed=3, n_trees=args.trees)
        words = 1 << 20  # 2^26 trials per tree per pass
        name = f"batched plain MC, {args.trees} uneven trees (50-5000 gates), every gate tallied"
    else:
        trees = [mc.small_tree(1)]
        words = 16 * 1024  # 2^20 trials
        name = "plain MC, small layered tree, every gate tallied, fused kernel"
    tallies = [[int(g) for g in t.gate_index] for t in trees]
    cfg = tuned_batch_config(args.workload)
    b = mc.Batch(trees, words, tallies, **cfg)
    info = b.info()
    if args.warmup:
        b.time_passes(args.warmup)
    b.wait()
    t0 = time.perf_counter()
    dev_ms = b.time_passes(args.steps)
    wall = time.perf_counter() - t0
    n_g = sum(t.n_gates for t in trees)
    n_be = sum(t.n_vars for t in trees)
    trials = args.steps * words * 64
    value = n_g * trials / wall
    # VALU roofline of the sampler: Philox words at the measured issue cost.
    out = {
        "metric": "bit-level gate-evals/s", "value": value, "unit": "gate-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": name, "trees": len(trees), "basic_events": n_be, "gates": n_g,
                   "words_per_pass": words, "trials_per_tree_per_pass": words * 64,
                   "batch_config": cfg, "grid": info["grid"], "lds_bytes": info["lds_bytes"],
                   "work_items": info["num_items"], "fused_trees": info["num_fused_trees"],
                   "layered_trees": info["num_layered_trees"],
                   "layered_launches_per_pass": info["layered_launches"]},
        "node_evals_per_s": (n_g + n_be) * trials / wall,
        "device_ms_per_pass": dev_ms / args.steps,
        "philox_words_per_s": n_be * words * args.steps / wall,
        "hbm": "fused trees: tallies only (the tree lives in LDS); layered trees: "
               "materialised rows, next tree sampled beside each tree's gates",
    }
    if not args.no_cpu:
        # CPU oracle on a fixed subset of the trees, bit-exact check beside it.
        sub = list(range(0, len(trees), max(1, len(trees) // 8)))[:8]
        sw = 1 << 10 if args.workload == "batched" else words
        t = time.perf_counter()
        cpu = [orc.run(trees[i], sw, tally=tallies[i])[0] for i in sub]
        dt = time.perf_counter() - t
        with mc.Batch([trees[i] for i in sub], sw, [tallies[i] for i in sub], **cfg) as small:
            small.pass_(1)
            ok = all([x.num_one_bits for x in small.collect(j)] == [int(v) for v in cpu[j]]
                     for j in range(len(sub)))
        out["cpu_baseline"] = {
            "value": sum(trees[i].n_gates for i in sub) * sw * 64 / dt, "unit": "gate-evals/s",
            "cores": orc.max_threads(), "kind": "port",
            "sample": f"oracle, trees {sub}, one pass of {sw} words each, {dt:.1f} s",
            "parity_bit_exact": ok, **host_cpu()}
    b.close()
    print(json.dumps(out), flush=True)
    return 0


def bench_ua(args) -> int:
    """Surrogate-screened uncertainty analysis against plain batched GPU Monte
    Carlo over the same held-out parameter samples.

    ua: the 2,000-event / 3,000-gate tree (lognormal basic events, median
    log-uniform in [1e-5, 1e-2], error factor uniform in [3, 10]); training
    pairs and outer samples at 2^22 trials, 16,384 outer samples.
    ua-large: the large tree's gate 64134 (956 events, 210 gates; the tree's
    own top event is certain), lognormal around the tree's probabilities with
    error factor uniform in [3, 10]; training pairs at 2^22 trials, 100,000
    outer samples at 2^24 trials, plain Monte Carlo over all of them."""
    from mcscram_amd import screening as scr
    from mcscram_amd import ua
    large = args.workload == "ua-large"
    if large:
        pd, model = ua.large_tree_uncertainty()
        outer = args.ua_outer or 100_000
        words = args.ua_words or 1 << 18
        name = (f"surrogate-screened UA, large tree's gate {ua.LARGE_TREE_UA_TOP} "
                f"({pd.n_vars} events / {pd.n_gates} gates)")
    else:
        pd, model = ua.uncertainty_tree()
        outer = args.ua_outer or 16384
        words = args.ua_words or 1 << 16
        name = "surrogate-screened UA, 2,000 events / 3,000 gates"
    train_words = args.ua_train_words or 1 << 16
    epochs = args.ua_epochs or (60 if large else 30)
    say = lambda m: print(f"[ua] {m}", file=sys.stderr, flush=True)  # noqa: E731
    say(f"training pairs: {args.ua_train} x {train_words * 64} trials")
    ens, info = scr.train_surrogate(pd, model, args.ua_train, train_words, seed=3,
                                    epochs=epochs,
                                    progress=lambda d, n: say(f"training pairs {d}/{n}"))
    say(f"trained: {info}")
    res = scr.screened_uncertainty(pd, model, ens, outer, words, seed=1234)
    say(f"screened: {res.seconds}, routed {res.routed.mean():.3f}")
    plain, t_plain = scr.plain_uncertainty(pd, model, outer, words, seed=1234, chunk=8192,
                                           progress=lambda d, n: say(f"plain Monte Carlo {d}/{n}"))
    say(f"plain Monte Carlo: {t_plain:.1f} s")
    cmp = scr.compare(res, plain)
    t_scr = res.seconds["total"]
    out = {
        "metric": "UA parameter-samples/s at <=2% rel err vs plain MC",
        "value": outer / t_scr, "unit": "samples/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": t_scr * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 network, u64 Monte Carlo", "data": "synthetic",
        "config": {"workload": name, "training_pairs": args.ua_train,
                   "trials_per_training_pair": train_words * 64, "outer_samples": outer,
                   "trials_per_mc_run": words * 64,
                   "network": f"{pd.n_vars}-512-512-512-1 GELU x2", "epochs": epochs,
                   "disagreement": 0.05, "audit": 0.05},
        "plain_mc_samples_per_s": outer / t_plain,
        "speedup_vs_plain_mc": t_plain / t_scr,
        # The one-off cost of the training pairs and the fit, charged to this
        # one set of outer samples (it is shared by every later analysis of
        # the same tree and uncertainty model).
        "training_seconds": info["mc_seconds"] + info["fit_seconds"],
        "speedup_including_training": t_plain / (t_scr + info["mc_seconds"] + info["fit_seconds"]),
        # The random audit sends this fraction to Monte Carlo whatever the
        # ensemble says, so the speedup cannot exceed 1 / audit.
        "speedup_cap_from_audit": 1.0 / 0.05,
        "comparison": cmp, "audit": res.audit_rel_error,
        "training": info,
        "within_2pct": bool(cmp["mean"]["rel_error"] <= 0.02 and cmp["p95"]["rel_error"] <= 0.02),
    }
    print(json.dumps(out), flush=True)
    return 0


def tuned_batch_config(workload: str) -> dict:
    """Batch configuration chosen by the autotuner for this workload, if one
    was recorded (mcscram_amd/tuned.json), else the defaults."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mcscram_amd", "tuned.json")
    try:
        with open(path) as f:
            return dict(json.load(f).get(workload, {}).get("config", {}))
    except OSError:
        return {}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed passes (default: 2^32 trials for the layered tree, i.e. 256 "
                         "passes at 2^18 words; 3 for batched; 32 for small-fused)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--words", type=int, default=1 << 18, help="uint64 words per pass")
    ap.add_argument("--tree", choices=["large", "small"], default="large")
    ap.add_argument("--workload", choices=["layered", "batched", "small-fused", "ua", "ua-large"],
                    default="layered",
                    help="layered: the large tree through the layered engine (default); "
                         "batched: 256 uneven trees, 2^26 trials each, one persistent launch "
                         "per pass; small-fused: the small tree, 2^20 trials, fused kernel; "
                         "ua: surrogate-screened uncertainty analysis vs plain GPU Monte Carlo; "
                         "ua-large: the same on the large tree's gate 64134, 100,000 outer "
                         "samples at 2^24 trials")
    ap.add_argument("--ua-train", type=int, default=65536, help="training pairs (ua)")
    ap.add_argument("--ua-outer", type=int, default=0,
                    help="outer parameter samples (ua: 16,384; ua-large: 100,000)")
    ap.add_argument("--ua-words", type=int, default=0,
                    help="words per outer Monte Carlo run (ua: 2^16; ua-large: 2^18)")
    ap.add_argument("--ua-train-words", type=int, default=0,
                    help="words per training Monte Carlo run (2^16)")
    ap.add_argument("--ua-epochs", type=int, default=0,
                    help="training epochs (ua: 30; ua-large: 60, which cuts the members' "
                         "disagreement on its 956 inputs from 0.7 % to 0.2 % of samples)")
    ap.add_argument("--trees", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device-to-device copy (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine's sharding over trials is a "
              "design note, not a launch path; run with --gpus 1", file=sys.stderr)
        return 2

    if args.steps <= 0:
        if args.workload == "layered":
            args.steps = max(1, (1 << 32) // (args.words * 64))
        else:
            args.steps = 3 if args.workload == "batched" else 32
    import mcscram_amd as mc
    if args.workload in ("ua", "ua-large"):
        return bench_ua(args)
    if args.workload != "layered":
        return bench_fused(args, mc)
    pd = mc.large_tree() if args.tree == "large" else mc.small_tree(1)
    tally = [int(g) for g in pd.gate_index]
    n_bp = args.words
    eng = mc.LayerManager(pd, n_bp, tally=tally)
    info = eng.info
    # The first pass also runs the first pass's sampler alone (one launch):
    # its extra time is the sampler's VALU-bound cost, the pass's floor.
    first = eng.time_passes(1)
    if args.warmup:
        eng.time_passes(args.warmup)
    eng.wait()
    t0 = time.perf_counter()
    dev = eng.time_passes(args.steps)  # synchronises the engine's stream
    wall = time.perf_counter() - t0
    tallies = eng.collect_tallies()
    eng.close()

    trials = args.steps * n_bp * 64
    gate_evals = pd.n_gates * trials
    sampler_ms = first["kernel_ms"] - dev["kernel_ms"] / args.steps
    value = gate_evals / wall
    ab = algorithmic_bytes(pd, n_bp)
    # pass_layer is the only kernel: every launch carries one gate layer and
    # a slice of the next pass's sampling.
    launch_s = dev["kernel_ms"] / 1e3 / dev["launches"]
    launches_per_pass = dev["launches"] / args.steps
    # Algorithmic bytes of the materialised dataflow, 8 (N_be + N_g + F + T)
    # per word, tallied words' reads included (the reference's tally kernel
    # re-reads every tallied node word; here the popcount is fused at
    # production): the launches of a pass together process one pass of words.
    bytes_per_launch = ab["pass"] / launches_per_pass
    achieved = bytes_per_launch / launch_s / 1e9
    pass_gbs = ab["pass"] / (wall / args.steps) / 1e9

    out = {
        "metric": "bit-level gate-evals/s",
        "value": value,
        "unit": "gate-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"plain MC, {args.tree} layered tree, every gate tallied",
                   "basic_events": pd.n_vars, "gates": pd.n_gates, "gate_inputs": pd.n_args,
                   "layers": info["num_layers"], "words_per_pass": n_bp,
                   "trials_per_pass": n_bp * 64, "trials_timed": trials},
        "node_evals_per_s": (pd.n_vars + pd.n_gates) * trials / wall,
        # The metric's second part: the dominant kernel's share of HBM peak
        # (the roofline's achieved / peak, in percent).
        "hbm_peak_pct": 100.0 * achieved / HBM_PEAK_GBS,
        "roofline": {"kernel": "pass_layer", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.tree, n_bp),
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": launch_s * 1e3, "launches_per_pass": launches_per_pass},
        "pass_roofline": {"bound": "hbm", "algorithmic_bytes_per_pass": ab["pass"],
                          "achieved": pass_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": pass_gbs / HBM_PEAK_GBS},
        # The sampler alone issues VALU instructions at the chip's peak rate
        # (rocprofv3 SQ_INSTS_VALU); a pass cannot be shorter than it.
        "valu_floor": {"sampler_only_ms": sampler_ms, "pass_ms": dev["kernel_ms"] / args.steps,
                       "frac": sampler_ms / (dev["kernel_ms"] / args.steps)},
        "device_ms_per_pass": {"kernels": dev["kernel_ms"] / args.steps,
                               "first_to_last": dev["pass_ms"] / args.steps},
        "event_buffers": info["event_buffers"],
        "root_estimate": next(t.mean for t in tallies if t.index == pd.root),
    }

    if not args.no_cpu:
        words, ones, dt, threads = cpu_baseline(pd, args.cpu_seconds, 1 << 15)
        with mc.LayerManager(pd, words, tally=tally) as small:
            small.pass_(1)
            gpu_ones = [t.num_one_bits for t in small.collect_tallies()]
        out["cpu_baseline"] = {
            "value": pd.n_gates * words * 64 / dt, "unit": "gate-evals/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle, same tree, one pass of {words} words ({words * 64} trials), "
                      f"every gate tallied, {dt:.1f} s on {threads} OpenMP threads",
            "parity_bit_exact": [int(x) for x in ones] == gpu_ones,
            # Linear in trials: the same oracle over the whole 2^32 trials.
            "extrapolated_seconds_2e32_trials": dt * float(1 << 32) / (words * 64), **host_cpu()}
    if args.copy_calibration:
        out["copy_calibration_bytes"] = copy_calibration(1 << 30)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
