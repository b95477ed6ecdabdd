"""convergence.csv rows (host only).

Pins: the column order and names of the reference's loggers
(src/mc/logger/log_settings.h, log_working_set.h, log_layers.h,
log_sample_shaper.h, log_pdag.h, log_tally.h, log_convergence.h, in the order
src/mc/core/probability_analysis.cc:148-193 concatenates them); the value
format of src/mc/logger/csv.h; the header-once append of
src/mc/logger/log_benchmark.h; and the PDAG columns against the values the
reference itself logged for the Aralia models at compilation level 0
(benchmarks/compiler.csv).
"""
import gzip
import json
import os

import pytest

import mcscram_amd as mc
from mcscram_amd import mef, report
from mcscram_amd.engine import Tally

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mef")

SETTINGS = ["input_file", "req_trials", "req_burnin_trials", "req_conf_lvl", "req_delta",
            "req_stop_on_conv", "req_conv_pol", "req_no_kn", "req_no_xor", "req_compile",
            "req_ccf", "req_overhead_ratio"]
DEVICE = ["device_type", "device_name", "vendor_id", "vendor", "driver_version", "profile",
          "version", "opencl_c_version", "max_compute_units", "max_clock_frequency",
          "max_work_item_dimensions", "max_work_item_sizes_1d", "max_work_item_sizes_2d_x",
          "max_work_item_sizes_2d_y", "max_work_item_sizes_3d_x", "max_work_item_sizes_3d_y",
          "max_work_item_sizes_3d_z", "work_item_independent_forward_progress",
          "max_work_group_size", "max_num_sub_groups", "sub_group_independent_forward_progress",
          "sub_group_sizes", "preferred_vector_width_char", "max_mem_alloc_size",
          "global_mem_cache_line_size", "global_mem_size", "global_mem_cache_size",
          "global_mem_cache_type", "local_mem_type", "local_mem_size"]
LAYERS = ["num_nodes", "num_layers", "num_kernels", "avg_kernels_per_layer",
          "device_basic_event_blocks", "device_gate_blocks", "device_atleast_gate_blocks",
          "device_tally_blocks", "device_total_blocks", "avg_blocks_per_kernel",
          "tally_events_tracked", "num_variables", "num_gates"] + [
    "shaper_" + k for k in ["requested_num_trials", "num_nodes", "max_device_bytes",
                            "max_device_bits", "bits_in_bitpack", "num_trials_rounded",
                            "total_bits_to_sample", "target_bits_per_iteration",
                            "bits_per_iteration", "num_iterations", "sample_shape_batch_size",
                            "sample_shape_bitpacks_per_batch", "sample_shape_num_bitpacks",
                            "total_iterations"]]
PDAG = ["pdag_root_index", "pdag_complement_graph", "pdag_coherent_graph",
        "pdag_normalized_graph", "pdag_num_gates_total", "pdag_num_modules",
        "pdag_num_gates_neg_indices", "pdag_num_gates_pos_and_neg", "pdag_num_variables_total",
        "pdag_num_variables_neg_indices", "pdag_num_variables_pos_and_neg"] + [
    "pdag_num_gates_" + t for t in ("and", "or", "atleast", "xor", "not", "nand", "nor", "null")
] + ["pdag_num_constants", "pdag_has_constants", "pdag_has_null_gates", "pdag_max_gate_fan_in",
     "pdag_avg_gate_fan_in", "pdag_num_substitutions"]
TALLY = ["one_bits", "total_bits", "std_err", "p01", "p05", "mean", "p95", "p99", "eps",
         "log10_eps", "target_eps", "target_log10_eps", "target_bits_eps",
         "target_bits_log10_eps", "cum_info_gain"]
RUN = ["convergence_time_ms", "bits_per_iteration", "total_iterations", "iterations_per_second",
       "bits_per_second"]


def _tally(ones, total):
    t = Tally(index=5, num_one_bits=ones, total_bits=total, mean=0.0, std_err=0.0,
              ci=(0.0, 0.0, 0.0, 0.0))
    from oracle import binding as orc
    e = orc.point_estimates(ones, total)
    t.mean, t.std_err, t.ci = float(e[0]), float(e[1]), (e[2], e[3], e[4], e[5])
    return t


def test_csv_string_follows_the_reference_format():
    assert report.csv_string(3) == "3"
    assert report.csv_string(True) == "1" and report.csv_string(False) == "0"
    assert report.csv_string(0.000123456789012) == "1.23456789e-04"
    assert report.csv_string(0.99) == "9.90000000e-01"
    assert report.csv_string(None) == "" and report.csv_string("C0") == "C0"


def test_row_has_the_reference_columns_in_order():
    pd = mc.small_tree(1)
    s = mc.Settings(num_trials=1 << 22, early_stop=False)
    row = report.convergence_row(pd, s, (16, 1024), _tally(790000, 1 << 22), 4, 1.25,
                                 input_file="small.xml")
    assert [k for k, _ in row] == SETTINGS + DEVICE + LAYERS + PDAG + TALLY + RUN
    kv = dict(row)
    assert kv["req_conv_pol"] == "bayes" and kv["req_stop_on_conv"] == "0"
    assert kv["shaper_bits_per_iteration"] == str(16 * 1024 * 64)
    assert kv["shaper_total_iterations"] == "4" and kv["shaper_num_iterations"] == "4"
    assert kv["convergence_time_ms"] == "1250"
    # log_convergence.h: iterations = bits / bits per iteration; bits_per_second
    # divides by the time in milliseconds.
    assert float(kv["total_iterations"]) == pytest.approx(4.0)
    assert float(kv["iterations_per_second"]) == pytest.approx(4.0 / 1.25)
    assert float(kv["bits_per_second"]) == pytest.approx((1 << 22) / 1250.0)
    assert float(kv["eps"]) > 0 and kv["cum_info_gain"] == ""


@pytest.mark.parametrize("name", ["baobab1", "das9601", "edf9203", "isp9605", "nus9601"])
def test_pdag_columns_equal_the_reference_log(name):
    with open(os.path.join(GOLD, "aralia.json")) as f:
        want = json.load(f)[name]["c0"]
    pd = mef.load_pdag(gzip.open(os.path.join(GOLD, "aralia", name + ".xml.gz")))
    got = dict(report.pdag_pairs(pd))
    compared = 0
    for k, v in want.items():
        if k in got and got[k] is not None:
            assert float(got[k]) == pytest.approx(v, abs=5e-6), k
            compared += 1
    assert compared >= 20


def test_logger_writes_the_header_once(tmp_path):
    path = str(tmp_path / "convergence.csv")
    a = [("x", "1"), ("y", "2")]
    report.BenchmarkLogger(path).log_pairs(a)
    report.BenchmarkLogger(path).log_pairs([("y", "4"), ("x", "3")])
    lines = open(path).read().splitlines()
    assert lines == ["x,y", "1,2", "3,4"]
