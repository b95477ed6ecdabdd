"""Monte Carlo on the GPU against exact probabilities (MEF models and the
small synthetic tree).

Exact values: the reference's benchmark table for the Aralia models
(benchmarks/input.csv; for das9204 the BDD instead, see tests/test_mef.py),
the reference's core known answers (tests/bench_core_tests.cc), and the BDD
of the small tree. The engine's stream is a pure function of the PDAG and the
sample shape (fixed Philox key), so each check below passes or fails the same
way on every run.

Tolerances: |mean - p| <= z * sqrt(p (1 - p) / N) with z = 3.2905 (two-sided
99.9 %), and the printed 99 % interval itself for the small tree at the
workload's 2^20 trials. Where p N < 30 the count of true trials is checked
against a Poisson bound instead.
"""
import gzip
import json
import math
import os

import pytest

import mcscram_amd as mc
from mcscram_amd import mef
from oracle import binding as orc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mef")
with open(os.path.join(GOLD, "aralia.json")) as _f:
    ARALIA = json.load(_f)
DAS9204_EXACT = 2.1694e-11  # BDD of the level-0 PDAG (tests/test_mef.py)

WORDS = 1 << 20  # words per pass: 2^26 trials


def trials_for(p: float) -> int:
    """Enough trials for about 4,000 expected hits, 2^26 .. 2^34."""
    want = 4000.0 / max(p, 1e-300)
    return 1 << min(34, max(26, math.ceil(math.log2(want))))


def check(ones: int, total: int, p: float, z: float) -> None:
    lam = p * total
    if lam < 30:
        bound = lam + 6.0 * math.sqrt(lam) + 6.0
        assert ones <= bound, (ones, lam)
        return
    mean = ones / total
    se = math.sqrt(p * (1 - p) / total)
    assert abs(mean - p) <= z * se, (mean, p, (mean - p) / se)


def run(pd, trials):
    passes = trials // (WORDS * 64)
    with mc.LayerManager(pd, WORDS) as eng:
        t = eng.pass_wait_collect(passes, passes_between_waits=16)[0]
    assert t.total_bits == trials
    return t


@pytest.mark.parametrize("name", sorted(n for n in ARALIA if ARALIA[n]["oracle_p"] >= 0))
def test_aralia_models_within_interval_of_exact(gpu, name):
    pd = mef.load_pdag(gzip.open(os.path.join(GOLD, "aralia", name + ".xml.gz")))
    p = DAS9204_EXACT if name == "das9204" else ARALIA[name]["oracle_p"]
    n = trials_for(p)
    t = run(pd, n)
    check(t.num_one_bits, t.total_bits, p, 3.2905)


def test_das9204_table_value_is_not_the_estimate(gpu):
    # The table's 6.07651e-08 would give about 1,044 true trials in 2^34;
    # the engine finds a count consistent with the exact 2.17e-11 instead.
    pd = mef.load_pdag(gzip.open(os.path.join(GOLD, "aralia", "das9204.xml.gz")))
    t = run(pd, 1 << 34)
    assert t.num_one_bits < 30


CORE = {"abc": 0.496, "ab_bc": 0.074, "ab_or_not_ac": 0.29, "atleast": 0.098,
        "a_or_not_a": 1.0, "a_or_not_b": 0.82, "a_and_not_a": 0.0, "a_and_not_b": 0.08,
        "a_or_not_ab": 0.28, "multiple_parent_negative_gate": 0.9, "xor": 0.404,
        "iff": 0.404, "imply": 0.986, "null_a": 0.5, "nand": 0.98, "nor": 0.72}


@pytest.mark.parametrize("name", sorted(CORE))
def test_core_models_through_loader(gpu, name):
    pd = mef.load_pdag(os.path.join(GOLD, "core", name + ".xml"))
    p = CORE[name]
    t = run(pd, 1 << 26)
    if p in (0.0, 1.0):
        assert t.num_one_bits == (t.total_bits if p == 1.0 else 0)
    else:
        check(t.num_one_bits, t.total_bits, p, 3.2905)


def test_small_tree_within_ci99_of_exact_bdd(gpu):
    """The small synthetic tree (200 events, 120 gates) at 2^20 trials: the
    exact BDD probability lies inside the estimate's printed 99 % interval."""
    pd = mc.small_tree(1)
    p, _ = orc.bdd_probability(pd)
    with mc.LayerManager(pd, (16, 1024)) as eng:
        t = eng.pass_wait_collect(1)[0]
    assert t.total_bits == 1 << 20
    lo99, hi99 = t.ci[2], t.ci[3]
    assert lo99 <= p <= hi99, (t.mean, p, t.ci)
    # And at 2^32 trials, where the interval is 64 times narrower.
    t = run(pd, 1 << 32)
    check(t.num_one_bits, t.total_bits, p, 3.2905)


def test_probability_analysis_entry_on_a_loaded_model(gpu):
    """The reference's entry point (fixed N) on baobab2, whose exact value the
    reference's README quotes next to its own GPU estimate."""
    pd = mef.load_pdag(gzip.open(os.path.join(GOLD, "aralia", "baobab2.xml.gz")))
    s = mc.Settings(num_trials=1 << 30, early_stop=False)
    p_total, tallies, passes = mc.probability_analysis(pd, s, WORDS)
    assert passes == (1 << 30) // (WORDS * 64)
    check(tallies[0].num_one_bits, tallies[0].total_bits, 7.13018e-4, 3.2905)
    assert p_total == tallies[0].mean


def test_convergence_csv_row_of_a_gpu_run(gpu, tmp_path):
    """A fixed-N run on baobab2 appends one convergence.csv row whose tally
    and run columns are the run's own."""
    from mcscram_amd import report
    pd = mef.load_pdag(gzip.open(os.path.join(GOLD, "aralia", "baobab2.xml.gz")))
    path = str(tmp_path / "convergence.csv")
    s = mc.Settings(num_trials=1 << 28, early_stop=False)
    for _ in range(2):
        p, tallies, passes = report.probability_analysis_logged(pd, s, WORDS, path,
                                                                input_file="baobab2.xml")
    lines = open(path).read().splitlines()
    assert len(lines) == 3
    header = lines[0].split(",")
    row = dict(zip(header, lines[2].split(",")))
    assert int(row["one_bits"]) == tallies[0].num_one_bits
    assert int(row["total_bits"]) == 1 << 28 == tallies[0].total_bits
    assert int(row["shaper_num_iterations"]) == passes == 4
    assert row["device_type"] == "gpu" and int(row["max_compute_units"]) > 0
    assert float(row["mean"]) == pytest.approx(p, rel=1e-7)


ALL_ARALIA = sorted(f[:-7] for f in os.listdir(os.path.join(GOLD, "aralia")) if f.endswith(".xml.gz"))


@pytest.mark.parametrize("name", ALL_ARALIA)
def test_aralia_model_tallies_bit_exact(gpu, name):
    # Every basic event and gate of the reference's own models, loaded from
    # their MEF files: the engine's one-bit counts equal the oracle's over
    # two passes of a ragged word count (ATLEAST gates, complemented inputs,
    # wide gates split into partial folds, fused single-consumer gates).
    pd = mef.load_pdag(gzip.open(os.path.join(GOLD, "aralia", name + ".xml.gz")))
    tally = [int(v) for v in pd.var_index] + [int(g) for g in pd.gate_index]
    with mc.LayerManager(pd, 1000, tally=tally) as eng:
        eng.pass_(2)
        got = {t.index: t.num_one_bits for t in eng.collect_tallies()}
    want, _ = orc.run(pd, 1000, n_iters=2, tally=tally)
    bad = [(i, got.get(i), int(w)) for i, w in zip(tally, want) if got.get(i) != int(w)]
    assert not bad, bad[:5]
