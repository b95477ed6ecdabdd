"""Autotuner for the batch engine's launch shape.

Sweeps, on the GPU, the knobs that decide how a batch of trees is executed:

* ``max_tile_words`` — words (bitpacks) per LDS tile of the fused launch;
* ``min_fused_tile_words`` — the narrowest tile a tree may get and still be
  fused whole in LDS; narrower trees run layer by layer through the layered
  pipeline (one layer per launch instead of all layers in one pass);
* ``lds_bytes`` — LDS per workgroup of the fused launch;
* ``workgroup_lanes`` — workgroup shape of the fused launch;
* ``layered_streams`` — layered pipelines run side by side (each with its
  own words arena, so it is swept at full size only).

Each configuration is timed on a proxy of the workload (the same trees at
fewer words per pass); the fastest few are re-timed at full size with each
``layered_streams`` value, and the winner is written to ``mcscram_amd/tuned.json`` under the workload's name,
with the sweep that chose it. Every configuration gives bit-identical
tallies, so the choice is about speed only.

Run: ``python -m mcscram_amd.autotune --workload batched``
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "tuned.json")

SPACE = {
    "max_tile_words": [16, 32, 64],
    "min_fused_tile_words": [8, 16, 32, 128],
    "lds_bytes": [32 << 10, 64 << 10, 96 << 10],
    "workgroup_lanes": [256, 512, 1024],
}
STREAMS = [1, 2, 3]  # layered_streams, at full size
# One small tree alone (320 nodes, 2^20 trials) is a single short launch:
# what decides its time is how evenly its tiles fill the workgroups and the
# workgroups the CUs, so narrower tiles, smaller LDS budgets and the work
# item size are swept instead of the fused/layered split (it always fits).
SMALL_SPACE = {
    "max_tile_words": [4, 8, 16, 32],
    "min_fused_tile_words": [1],
    "lds_bytes": [16 << 10, 32 << 10, 48 << 10, 64 << 10],
    "workgroup_lanes": [128, 256, 512, 1024],
    "item_words": [0, 32, 64],
}


def workload(name: str):
    import mcscram_amd as mc
    if name == "batched":
        return mc.batched_trees(seed=3, n_trees=256), 1 << 20
    if name == "small-fused":
        return [mc.small_tree(1)], 16 * 1024
    raise ValueError(name)


def time_config(trees, words, cfg, passes=1) -> float:
    import mcscram_amd as mc
    tallies = [[int(g) for g in t.gate_index] for t in trees]
    with mc.Batch(trees, words, tallies, **cfg) as b:
        b.time_passes(1)  # warm-up; primes the layered pipelines
        return b.time_passes(passes) / passes


def sweep(name: str, proxy_div: int = 16, keep: int = 4, log=print) -> dict:
    trees, words = workload(name)
    small = name == "small-fused"
    # The small tree's pass is a fraction of a millisecond: time it at full
    # size over many passes instead of on a proxy.
    proxy_words = words if small else max(words // proxy_div, 1024)
    passes = 50 if small else 1
    space = SMALL_SPACE if small else SPACE
    results = []
    keys = list(space)
    for values in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, values))
        if cfg["min_fused_tile_words"] > cfg["max_tile_words"] and cfg["min_fused_tile_words"] != 128:
            continue  # no tree could be fused: same as 128
        try:
            ms = time_config(trees, proxy_words, cfg, passes)
        except Exception as e:  # a shape the device rejects (LDS, occupancy)
            log(f"skip {cfg}: {e}")
            continue
        results.append({"config": cfg, "proxy_ms": ms})
        log(f"proxy {proxy_words} words  {ms:9.3f} ms  {cfg}")
    results.sort(key=lambda r: r["proxy_ms"])
    finals = []
    for r in results[:keep]:
        for ls in ([0] if small else STREAMS):  # no layered tree in the small workload
            cfg = dict(r["config"], layered_streams=ls) if ls else dict(r["config"])
            ms = time_config(trees, words, cfg, 4 * passes)
            finals.append({"config": cfg, "ms_per_pass": ms})
            log(f"full  {words} words  {ms:9.3f} ms  {cfg}")
    finals.sort(key=lambda r: r["ms_per_pass"])
    return {"config": finals[0]["config"], "ms_per_pass": finals[0]["ms_per_pass"],
            "words_per_pass": words, "proxy_words_per_pass": proxy_words,
            "finals": finals, "proxy": results, "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S")}


def main() -> int:
    ap = argparse.ArgumentParser(description="Autotune the batch engine's launch shape on the GPU")
    ap.add_argument("--workload", choices=["batched", "small-fused"], default="batched")
    ap.add_argument("--out", default=TUNED)
    ap.add_argument("--proxy-div", type=int, default=16)
    args = ap.parse_args()
    res = sweep(args.workload, args.proxy_div, log=lambda s: print(s, flush=True))
    data = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            data = json.load(f)
    data[args.workload] = res
    with open(args.out, "w") as f:
        json.dump(data, f, indent=1)
    print(json.dumps({"workload": args.workload, "config": res["config"],
                      "ms_per_pass": res["ms_per_pass"]}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
