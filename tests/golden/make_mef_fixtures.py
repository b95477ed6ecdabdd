"""Writes the MEF fixtures under tests/golden/mef/ from the reference's data files.

* mef/aralia/<model>.xml.gz: the Aralia fault trees the reference's benchmark
  table lists (input/Aralia/*.xml), gzip-compressed.
* mef/core/<model>.xml: the reference's small known-answer inputs
  (tests/input/core/*.xml) whose features the level-0 loader supports.
* mef/aralia.json: per model, the exact top-event probability of the
  reference's benchmark table (benchmarks/input.csv, column oracle_p; -1 means
  unknown) and the reference's level-0 PDAG statistics for it
  (benchmarks/compiler.csv, rows with req_compile = C0).

Run where the reference tree exists: python tests/golden/make_mef_fixtures.py
"""
import csv
import gzip
import json
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "mef")

# Core inputs with no house events, constants, CCF groups or cardinality.
CORE = ["abc", "ab_bc", "ab_or_not_ac", "atleast", "a_or_not_a", "a_or_not_b", "a_and_not_a",
        "a_and_not_b", "a_or_not_ab", "multiple_parent_negative_gate", "nand", "nor",
        "nand_or_equality", "not_and_or_equality", "xor", "iff", "imply", "null_a",
        "single_exponential"]

STATS = ["model_num_basic_events", "model_num_gates", "pdag_num_gates_total",
         "pdag_num_variables_total", "pdag_num_gates_and", "pdag_num_gates_or",
         "pdag_num_gates_atleast", "pdag_num_gates_xor", "pdag_num_gates_not",
         "pdag_num_gates_nand", "pdag_num_gates_nor", "pdag_num_gates_null",
         "pdag_num_gates_neg_indices", "pdag_max_gate_fan_in", "pdag_avg_gate_fan_in",
         "pdag_complement_graph", "pdag_coherent_graph", "pdag_normalized_graph",
         "pdag_num_gates_pos_and_neg", "pdag_num_variables_neg_indices",
         "pdag_num_variables_pos_and_neg", "pdag_num_constants", "pdag_has_constants",
         "pdag_has_null_gates", "pdag_num_substitutions"]


def main():
    os.makedirs(os.path.join(OUT, "aralia"), exist_ok=True)
    os.makedirs(os.path.join(OUT, "core"), exist_ok=True)
    with open(os.path.join(REF, "benchmarks", "input.csv"), encoding="utf-8-sig") as f:
        table = {r["input_file"]: float(r["oracle_p"]) for r in csv.DictReader(f)}
    stats = {}
    with open(os.path.join(REF, "benchmarks", "compiler.csv"), encoding="utf-8-sig") as f:
        for r in csv.DictReader(f):
            name = os.path.basename(r["input_file"])
            if r["req_compile"] == "C0" and name in table and name not in stats:
                stats[name] = {k: float(r[k]) for k in STATS}
    out = {}
    for name, p in table.items():
        src = os.path.join(REF, "input", "Aralia", name)
        with open(src, "rb") as fi, gzip.GzipFile(os.path.join(OUT, "aralia", name + ".gz"), "wb",
                                                 mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        out[name[:-4]] = {"oracle_p": p, "c0": stats.get(name)}
    with open(os.path.join(OUT, "aralia.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    for name in CORE:
        shutil.copyfile(os.path.join(REF, "tests", "input", "core", name + ".xml"),
                        os.path.join(OUT, "core", name + ".xml"))
    print(f"{len(out)} Aralia models, {len(CORE)} core models -> {OUT}")


if __name__ == "__main__":
    main()
