"""Throughput of the Monte Carlo fault-tree engine on one MI355X.

Workload: plain Monte Carlo on the large synthetic layered tree (25,000 basic
events with p log-uniform in [1e-6, 1e-2]; 40,000 gates in 30 layers, fan-in
2-8, 48 % AND / 47 % OR / 5 % ATLEAST with n up to 16, 10 % negated inputs),
every gate tallied, 2^18 uint64 words = 2^24 trials per pass. One step is one
pass (Philox sampling of every basic event, every gate layer, tallies);
2^32 trials are 256 steps.

Prints one JSON line: gate-evals/s (gates x trials / s), the roofline of
the dominant kernel (its launch time measured with HIP events on the
engine's stream; its VALU instructions and HBM bytes per launch from the
committed rocprofv3 counter summaries of the same kernel source), and the
CPU oracle timed on the host on a bounded sample of the same tree (with a
bit-exact tally comparison against the GPU on that sample).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# VALU issue peak: one wave64 instruction per CU per clock (four SIMDs of
# 16 lanes, four cycles each), 256 CUs at the 2.4 GHz peak engine clock:
# 39.3 T int32 lane operations per second.
CUS, PEAK_CLOCK_GHZ = 256, 2.4
VALU_PEAK_TOPS = CUS * 64 * PEAK_CLOCK_GHZ / 1e3
ROOT = os.path.dirname(os.path.abspath(__file__))
KERNEL_SOURCES = ("mcscram_amd/csrc/kernels.hip", "mcscram_amd/csrc/kernels.h",
                  "mcscram_amd/csrc/device_ops.h")


def kernel_source_sha16() -> str:
    """Hash of pass_layer's source files: counter summaries recorded with the
    same hash describe the kernel that runs."""
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


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


def counter_summary(kind: str, tree: str, n_bp: int):
    """The committed rocprofv3 counter summary of pass_layer on this workload
    (kind "hbm": FETCH_SIZE / WRITE_SIZE scaled by the copy calibration;
    "valu": SQ_INSTS_VALU and the clock), or None when none matches the
    workload or the kernel source."""
    path = os.path.join(ROOT, "profiles", f"pass_layer_{tree}_tree_{kind}_counters.json")
    try:
        with open(path) as f:
            s = json.load(f)
    except OSError:
        return None
    if s.get("words_per_pass") != n_bp or s.get("tree") != tree:
        return None
    if s.get("kernel_source_sha16") != kernel_source_sha16():
        return None
    return s


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
        if args.workload == "batched":
            # Words per tree doubled from 1,024 while the sample stays near
            # --cpu-seconds (the oracle's time is linear in words).
            sw = 1 << 10
            t = time.perf_counter()
            for i in sub:
                orc.run(trees[i], sw, tally=tallies[i])
            per_word = (time.perf_counter() - t) / sw
            while sw * 2 <= words and per_word * sw * 2 <= args.cpu_seconds:
                sw *= 2
        else:
            sw = words
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
            "build": os.path.basename(orc.library_variant()),
            "parity_bit_exact": ok, **host_cpu()}
    b.close()
    print(json.dumps(out), flush=True)
    return 0


def bench_ua(args) -> int:
    print(json.dumps(run_ua(args)), flush=True)
    return 0


def run_ua(args) -> dict:
    """Surrogate-screened uncertainty analysis against plain batched GPU Monte
    Carlo.

    ua: the 2,000-event / 3,000-gate tree (lognormal basic events, median
    log-uniform in [1e-5, 1e-2], error factor uniform in [3, 10]); 65,536
    training pairs and 16,384 outer samples at 2^22 trials; plain Monte Carlo
    over every outer sample is the yardstick.
    ua-scale: the large tree's family (25,000 events, 40,002 gates; top
    AND(A, B) over the last layer, p about 0.04) with lognormal uncertainty
    (error factor uniform in [3, 10]) on every event, all 25,000 the network's
    inputs; 65,536 training pairs at 2^20 trials; 100,000 outer samples, routed ones
    (disagreement over 0.05 in logit, plus a random 5 % audit) at 2^24 trials.
    Plain Monte Carlo over all 100,000 would take hours: its rate is the
    routed runs' (same engine, same shape) and the surrogate's error is
    measured paired, prediction against Monte Carlo, on the random audit."""
    from mcscram_amd import screening as scr
    from mcscram_amd import ua
    scale = args.workload == "ua-scale"
    if scale:
        pd, model = ua.at_scale_uncertainty()
        outer = args.ua_outer or 100_000
        words = args.ua_words or 1 << 18
        train = args.ua_train or 65536
        train_words = args.ua_train_words or 1 << 14
        name = (f"surrogate-screened UA at scale: {pd.n_vars} events / {pd.n_gates} gates "
                f"(large tree's family, top AND of the last layer's rare and likely gates)")
    else:
        pd, model = ua.uncertainty_tree()
        outer = args.ua_outer or 16384
        words = args.ua_words or 1 << 16
        train = args.ua_train or 65536
        train_words = args.ua_train_words or 1 << 16
        name = "surrogate-screened UA, 2,000 events / 3,000 gates"
    epochs = args.ua_epochs or (60 if scale else 30)
    say = lambda m: print(f"[ua] {m}", file=sys.stderr, flush=True)  # noqa: E731
    if args.ua_load:
        # A surrogate trained and saved by an earlier run of this workload
        # (--ua-save): its training cost is the one that run recorded.
        from mcscram_amd.surrogate import Ensemble
        ens, info = Ensemble.load(args.ua_load, input_map=scr.lognormal_input_map(model))
        say(f"loaded {args.ua_load}: {info}")
    else:
        say(f"training pairs: {train} x {train_words * 64} trials")
        ens, info = scr.train_surrogate(pd, model, train, train_words, seed=3, epochs=epochs,
                                        chunk=4096 if scale else 16384,
                                        progress=lambda d, n: say(f"training pairs {d}/{n}"))
        say(f"trained: {info}")
        if args.ua_save:
            ens.save(args.ua_save, info)
        if args.ua_train_only:
            return {"workload": name, "phase": "trained", "training": info,
                    "saved": args.ua_save or None}
    res = scr.screened_uncertainty(pd, model, ens, outer, words, seed=1234,
                                   mc_chunk=1024 if scale else 1 << 30,
                                   progress=lambda d, n: say(f"routed Monte Carlo {d}/{n}"))
    say(f"screened: {res.seconds}, routed {res.routed.mean():.4f}")
    t_scr = res.seconds["total"]
    out = {
        "metric": "UA parameter-samples/s at <=2% rel err vs plain MC",
        "value": outer / t_scr, "unit": "samples/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": t_scr * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 network, u64 Monte Carlo", "data": "synthetic",
        "config": {"workload": name, "training_pairs": train,
                   "trials_per_training_pair": train_words * 64, "outer_samples": outer,
                   "trials_per_mc_run": words * 64,
                   "network": f"{pd.n_vars}-512-512-512-1 GELU x2", "epochs": epochs,
                   "disagreement_logit": 0.05, "audit": 0.05},
        # The one-off cost of the training pairs and the fit, charged to this
        # one set of outer samples (it is shared by every later analysis of
        # the same tree and uncertainty model).
        "training_seconds": info["mc_seconds"] + info["fit_seconds"],
        # The random audit sends this fraction to Monte Carlo whatever the
        # ensemble says, so the speedup cannot exceed 1 / audit.
        "speedup_cap_from_audit": 1.0 / 0.05,
        "routed_fraction": float(res.routed.mean()),
        "disagreed_fraction": float(res.disagreed.mean()),
        "audit_paired": res.audit_rel_error,
        "screened_statistics": scr._quantiles(res.estimates),
        "training": info,
    }
    if scale:
        per = res.seconds["mc_per_sample"]
        t_plain = per * outer
        out["plain_mc_samples_per_s"] = 1.0 / per
        out["plain_mc"] = ("rate of the routed runs (%d samples, same engine and shape); all "
                           "%d samples would take %.0f s" % (int(res.routed.sum()), outer, t_plain))
        cmp = res.audit_rel_error
    else:
        plain, t_plain = scr.plain_uncertainty(pd, model, outer, words, seed=1234, chunk=8192,
                                               progress=lambda d, n: say(f"plain Monte Carlo {d}/{n}"))
        say(f"plain Monte Carlo: {t_plain:.1f} s")
        cmp = scr.compare(res, plain)
        out["plain_mc_samples_per_s"] = outer / t_plain
        out["comparison"] = cmp
    out["speedup_vs_plain_mc"] = t_plain / t_scr
    out["speedup_including_training"] = t_plain / (t_scr + out["training_seconds"])
    out["within_2pct"] = bool(cmp["mean"]["rel_error"] <= 0.02 and cmp["p95"]["rel_error"] <= 0.02)
    return out


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
    ap.add_argument("--workload", choices=["layered", "batched", "small-fused", "ua", "ua-scale"],
                    default="layered",
                    help="layered: the large tree through the layered engine (default); "
                         "batched: 256 uneven trees, 2^26 trials each, one persistent launch "
                         "per pass; small-fused: the small tree, 2^20 trials, fused kernel; "
                         "ua: surrogate-screened uncertainty analysis vs plain GPU Monte Carlo; "
                         "ua-scale: the same on the large tree's family, 100,000 outer "
                         "samples at 2^24 trials")
    ap.add_argument("--ua-train", type=int, default=0,
                    help="training pairs (ua and ua-scale: 65,536)")
    ap.add_argument("--ua-outer", type=int, default=0,
                    help="outer parameter samples (ua: 16,384; ua-scale: 100,000)")
    ap.add_argument("--ua-words", type=int, default=0,
                    help="words per outer Monte Carlo run (ua: 2^16; ua-scale: 2^18)")
    ap.add_argument("--ua-train-words", type=int, default=0,
                    help="words per training Monte Carlo run (ua: 2^16; ua-scale: 2^14)")
    ap.add_argument("--ua-epochs", type=int, default=0,
                    help="training epochs (ua: 30; ua-scale: 60)")
    ap.add_argument("--ua-save", default="",
                    help="write the trained surrogate ensemble to this .npz file")
    ap.add_argument("--ua-load", default="",
                    help="screen with an ensemble written by --ua-save instead of training")
    ap.add_argument("--ua-train-only", action="store_true",
                    help="stop after training (with --ua-save: the first half of a run "
                         "split in two)")
    ap.add_argument("--trees", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=24.0,
                    help="bound on the CPU baseline's sample (words double while the oracle's "
                         "time stays under it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ua", action="store_true",
                    help="skip the screened uncertainty analysis (2,000 events / 3,000 gates) "
                         "that the default layered run also measures")
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
    if args.workload in ("ua", "ua-scale"):
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
    # Counters of the same kernel source on the same workload: HBM bytes
    # and VALU instructions per launch, over this run's launch time.
    hbm_c = counter_summary("hbm", args.tree, n_bp)
    valu_c = counter_summary("valu", args.tree, n_bp)
    traffic = hbm_c["traffic_bytes_per_launch"] if hbm_c else None
    roofline = {"kernel": "pass_layer", "bound": "hbm", "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic}
    hbm = {"algorithmic_bytes_per_launch": bytes_per_launch, "algorithmic_achieved": achieved,
           "algorithmic_frac": achieved / HBM_PEAK_GBS, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    if traffic:
        hbm["counter_achieved"] = traffic / launch_s / 1e9
        hbm["counter_frac"] = hbm["counter_achieved"] / HBM_PEAK_GBS
    if valu_c:
        instr = valu_c["valu_instructions_per_launch"]
        tops = instr * 64 / launch_s / 1e12
        clock = valu_c["clock_ghz"]
        valu = {"instructions_per_launch": instr, "achieved": tops, "peak": VALU_PEAK_TOPS,
                "unit": "TOP/s", "frac": tops / VALU_PEAK_TOPS,
                # SQ_INSTS_VALU / (256 CUs x the clock the counters measured x time)
                "achieved_valu": instr / (CUS * clock * 1e9 * launch_s), "clock_ghz": clock}
        # The binding roofline is the one the counters put closest to its peak.
        if valu["frac"] >= hbm.get("counter_frac", hbm["algorithmic_frac"]):
            roofline = {"kernel": "pass_layer", "bound": "valu", "achieved": tops,
                        "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": tops / VALU_PEAK_TOPS,
                        "traffic": traffic}
        roofline["valu"] = valu
        roofline["achieved_valu"] = valu["achieved_valu"]
    roofline["hbm"] = hbm
    roofline.update({"avg_launch_ms": launch_s * 1e3, "launches_per_pass": launches_per_pass,
                     "counters": "profiles/pass_layer_%s_tree_{hbm,valu}_counters.json" % args.tree
                     if (hbm_c and valu_c) else "no summary for this kernel source",
                     "kernel_source_sha16": kernel_source_sha16()})

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
        # on the materialised dataflow's bytes, and on the bytes it moves.
        "hbm_peak_pct": 100.0 * achieved / HBM_PEAK_GBS,
        "hbm_peak_pct_measured_traffic": 100.0 * hbm["counter_frac"] if traffic else None,
        "roofline": roofline,
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
        words, ones, dt, threads = cpu_baseline(pd, args.cpu_seconds, 1 << 17)
        with mc.LayerManager(pd, words, tally=tally) as small:
            small.pass_(1)
            gpu_ones = [t.num_one_bits for t in small.collect_tallies()]
        out["cpu_baseline"] = {
            "value": pd.n_gates * words * 64 / dt, "unit": "gate-evals/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle, same tree, one pass of {words} words ({words * 64} trials), "
                      f"every gate tallied, {dt:.1f} s on {threads} OpenMP threads",
            "build": os.path.basename(__import__("oracle.binding", fromlist=["x"]).library_variant()),
            "parity_bit_exact": [int(x) for x in ones] == gpu_ones,
            # Linear in trials: the same oracle over the whole 2^32 trials.
            "extrapolated_seconds_2e32_trials": dt * float(1 << 32) / (words * 64), **host_cpu()}
    if args.copy_calibration:
        out["copy_calibration_bytes"] = copy_calibration(1 << 30)
    if not args.no_ua and args.tree == "large":
        # The metric's UA half in the same run: the 2,000-event / 3,000-gate
        # workload at its specified size, after the timed passes above. A
        # failure here is reported in the line instead of losing it.
        ua_args = argparse.Namespace(**{**vars(args), "workload": "ua"})
        try:
            u = run_ua(ua_args)
            out["ua"] = {k: u[k] for k in ("metric", "value", "unit", "plain_mc_samples_per_s",
                                           "speedup_vs_plain_mc", "within_2pct", "routed_fraction",
                                           "disagreed_fraction", "comparison", "audit_paired",
                                           "training_seconds", "speedup_including_training",
                                           "config")}
        except Exception as e:  # noqa: BLE001
            out["ua"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
