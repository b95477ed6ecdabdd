"""convergence.csv rows in the reference's format.

After a probability analysis the reference appends one row to
``convergence.csv`` (src/mc/core/probability_analysis.cc:148-193): settings,
device, layer-manager, PDAG and root-tally columns, the wall time, and four
derived run metrics. This module writes the same columns in the same order,
with the same value formatting (src/mc/logger/csv.h: integers in decimal,
doubles in scientific notation with 8 digits, booleans as 1/0), and the
same file behaviour (src/mc/logger/log_benchmark.h: rows are appended, the
header is written only into an empty file, a column a row lacks is left
empty).

Columns and sources:

* settings (src/mc/logger/log_settings.h): input_file, req_trials,
  req_burnin_trials, req_conf_lvl, req_delta, req_stop_on_conv, req_conv_pol,
  req_no_kn, req_no_xor, req_compile, req_ccf, req_overhead_ratio;
* device (src/mc/logger/log_working_set.h): the SYCL device queries; here the
  HIP device's name, compute units, clock, memory and LDS sizes, the rest
  left empty;
* layers (src/mc/logger/log_layers.h, log_sample_shaper.h with the prefix
  "shaper_"): node, layer and kernel counts and the sample shape. This engine
  has no per-connective device blocks, so the block columns are empty, and
  num_kernels is its launches per pass;
* pdag (src/mc/logger/log_pdag.h): the graph's flags and counts; module
  detection is not part of the Monte Carlo path, so pdag_num_modules is
  empty;
* tally (src/mc/logger/log_tally.h) of the root, with the stopping rule's
  epsilons (src/mc/scheduler/convergence_policy.h); cum_info_gain (a
  Boost.Math Beta-entropy diagnostic) is empty;
* convergence_time_ms, then (src/mc/logger/log_convergence.h)
  bits_per_iteration, total_iterations, iterations_per_second and
  bits_per_second. As in the reference, bits_per_second divides the total
  bits by the time in milliseconds.
"""
from __future__ import annotations

import os

import numpy as np

from .pdag import AND, ATLEAST, NAND, NOR, NOT, NULL, OR, XOR, Pdag

_CONN_TAGS = ("and", "or", "atleast", "xor", "not", "nand", "nor", "null")


def csv_string(v) -> str:
    """src/mc/logger/csv.h: decimal integers, %.8e doubles, strings as is."""
    if v is None:
        return ""
    if isinstance(v, (bool, np.bool_)):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return f"{float(v):.8e}"
    return str(v)


def settings_pairs(settings, input_file: str = "", compilation_level: int = 0,
                   overhead_ratio: float = 0.05):
    return [("input_file", input_file),
            ("req_trials", settings.num_trials),
            ("req_burnin_trials", settings.burn_in_trials),
            ("req_conf_lvl", float(settings.ci_confidence)),
            ("req_delta", float(settings.ci_rel_margin)),
            ("req_stop_on_conv", settings.num_trials == 0 or bool(settings.early_stop)),
            ("req_conv_pol", settings.ci_policy),
            ("req_no_kn", False),
            ("req_no_xor", False),
            ("req_compile", f"C{compilation_level}"),
            ("req_ccf", False),
            ("req_overhead_ratio", float(overhead_ratio))]


_DEVICE_KEYS = ("device_type", "device_name", "vendor_id", "vendor", "driver_version", "profile",
                "version", "opencl_c_version", "max_compute_units", "max_clock_frequency",
                "max_work_item_dimensions", "max_work_item_sizes_1d", "max_work_item_sizes_2d_x",
                "max_work_item_sizes_2d_y", "max_work_item_sizes_3d_x", "max_work_item_sizes_3d_y",
                "max_work_item_sizes_3d_z", "work_item_independent_forward_progress",
                "max_work_group_size", "max_num_sub_groups",
                "sub_group_independent_forward_progress", "sub_group_sizes",
                "preferred_vector_width_char", "max_mem_alloc_size", "global_mem_cache_line_size",
                "global_mem_size", "global_mem_cache_size", "global_mem_cache_type",
                "local_mem_type", "local_mem_size")


def device_pairs(device: dict | None = None):
    """device: any of the keys above (e.g. from ``hip_device_info``)."""
    device = device or {}
    return [(k, device.get(k)) for k in _DEVICE_KEYS]


def hip_device_info(device: int = 0) -> dict:
    """The HIP device's properties under the reference's column names (via
    torch, which only reads them; empty when no device is visible)."""
    try:
        import torch
        if not torch.cuda.is_available():
            return {}
        p = torch.cuda.get_device_properties(device)
    except Exception:
        return {}
    info = {"device_type": "gpu", "device_name": getattr(p, "gcnArchName", p.name),
            "vendor": "AMD", "max_compute_units": int(p.multi_processor_count),
            "global_mem_size": int(p.total_memory), "max_work_group_size": 1024,
            "sub_group_sizes": "64", "local_mem_type": "local"}
    for attr, key in (("L2_cache_size", "global_mem_cache_size"),
                      ("shared_memory_per_block", "local_mem_size")):
        if hasattr(p, attr):
            info[key] = int(getattr(p, attr))
    return info


def pdag_pairs(pd: Pdag):
    t = pd.gate_type
    neg = pd.args < 0
    gate_ids = set(int(g) for g in pd.gate_index)
    a = np.abs(pd.args)
    is_gate_arg = np.fromiter((int(x) in gate_ids for x in a), dtype=bool, count=a.shape[0])
    gates_neg = set(a[neg & is_gate_arg].tolist())
    gates_pos = set(a[~neg & is_gate_arg].tolist())
    vars_neg = set(a[neg & ~is_gate_arg].tolist())
    vars_pos = set(a[~neg & ~is_gate_arg].tolist())
    coherent = not neg.any() and not np.isin(t, (XOR, NOT, NAND, NOR)).any()
    normal = bool(np.isin(t, (AND, OR)).all())
    fan = pd.fan_in()
    pairs = [("pdag_root_index", int(pd.root)),
             ("pdag_complement_graph", False),
             ("pdag_coherent_graph", bool(coherent)),
             ("pdag_normalized_graph", normal),
             ("pdag_num_gates_total", pd.n_gates),
             ("pdag_num_modules", None),
             ("pdag_num_gates_neg_indices", len(gates_neg)),
             ("pdag_num_gates_pos_and_neg", len(gates_neg & gates_pos)),
             ("pdag_num_variables_total", pd.n_vars),
             ("pdag_num_variables_neg_indices", len(vars_neg)),
             ("pdag_num_variables_pos_and_neg", len(vars_neg & vars_pos))]
    for c, tag in enumerate(_CONN_TAGS):
        pairs.append((f"pdag_num_gates_{tag}", int((t == c).sum())))
    pairs += [("pdag_num_constants", 0), ("pdag_has_constants", False),
              ("pdag_has_null_gates", bool((t == NULL).any())),
              ("pdag_max_gate_fan_in", int(fan.max()) if fan.size else 0),
              ("pdag_avg_gate_fan_in", float(fan.mean()) if fan.size else 0.0),
              ("pdag_num_substitutions", 0)]
    return pairs


def layer_pairs(pd: Pdag, num_layers: int, launches_per_pass: int, n_tallies: int,
                requested_trials: int, rounded_trials: int, shape, passes_planned: int,
                passes_run: int):
    batch, bpb = shape if isinstance(shape, tuple) else (1, int(shape))
    n_bp = batch * bpb
    nodes = pd.n_vars + pd.n_gates
    pairs = [("num_nodes", nodes), ("num_layers", num_layers), ("num_kernels", launches_per_pass),
             ("avg_kernels_per_layer", launches_per_pass / max(num_layers, 1)),
             ("device_basic_event_blocks", None), ("device_gate_blocks", None),
             ("device_atleast_gate_blocks", None), ("device_tally_blocks", None),
             ("device_total_blocks", None), ("avg_blocks_per_kernel", None),
             ("tally_events_tracked", n_tallies), ("num_variables", pd.n_vars),
             ("num_gates", pd.n_gates)]
    shaper = [("requested_num_trials", requested_trials), ("num_nodes", nodes),
              ("max_device_bytes", None), ("max_device_bits", None), ("bits_in_bitpack", 64),
              ("num_trials_rounded", rounded_trials), ("total_bits_to_sample", rounded_trials),
              ("target_bits_per_iteration", n_bp * 64), ("bits_per_iteration", n_bp * 64),
              ("num_iterations", passes_run), ("sample_shape_batch_size", batch),
              ("sample_shape_bitpacks_per_batch", bpb), ("sample_shape_num_bitpacks", n_bp),
              ("total_iterations", passes_planned)]
    return pairs + [("shaper_" + k, v) for k, v in shaper]


def tally_pairs(tally, convergence: dict | None = None):
    c = convergence or {}
    return [("one_bits", tally.num_one_bits), ("total_bits", tally.total_bits),
            ("std_err", float(tally.std_err)), ("p01", float(tally.ci[2])),
            ("p05", float(tally.ci[0])), ("mean", float(tally.mean)),
            ("p95", float(tally.ci[1])), ("p99", float(tally.ci[3])),
            ("eps", c.get("epsilon")), ("log10_eps", c.get("epsilon_log10")),
            ("target_eps", c.get("target_epsilon")),
            ("target_log10_eps", c.get("target_epsilon_log10")),
            ("target_bits_eps", c.get("target_trials")),
            ("target_bits_log10_eps", c.get("target_trials_log10")),
            ("cum_info_gain", None)]


def convergence_pairs(pairs):
    """The reference's derived run metrics (src/mc/logger/log_convergence.h)."""
    kv = dict(pairs)
    bits_per_iteration = float(kv.get("shaper_bits_per_iteration") or 0)
    total_bits = float(kv.get("total_bits") or 0)
    total_time = float(kv.get("convergence_time_ms") or 0)
    total_iterations = total_bits / bits_per_iteration if bits_per_iteration else 0.0
    its = total_iterations / total_time * 1000.0 if total_time else 0.0
    bps = total_bits / total_time if total_time else 0.0
    return [("bits_per_iteration", bits_per_iteration), ("total_iterations", total_iterations),
            ("iterations_per_second", its), ("bits_per_second", bps)]


def convergence_row(pd: Pdag, settings, shape, root_tally, passes_run: int, seconds: float,
                    input_file: str = "", device: dict | None = None,
                    launches_per_pass: int | None = None, n_tallies: int = 1):
    """One convergence.csv row, as [(column, text)]."""
    from .engine import convergence_update, normal_quantile_two_sided, pdag_gate_depths, plan_passes
    depths = pdag_gate_depths(pd)
    num_layers = int(depths.max()) + 1 if depths.size else 1
    if settings.num_trials:
        rounded, planned = plan_passes(settings.num_trials, shape)
    else:
        rounded, planned = 0, 0
    conv = convergence_update(settings.ci_policy, root_tally.num_one_bits, root_tally.total_bits,
                              settings.ci_rel_margin,
                              normal_quantile_two_sided(settings.ci_confidence))
    pairs = (settings_pairs(settings, input_file) + device_pairs(device) +
             layer_pairs(pd, num_layers, launches_per_pass or num_layers - 1, n_tallies,
                         settings.num_trials, rounded, shape, planned, passes_run) +
             pdag_pairs(pd) + tally_pairs(root_tally, conv) +
             [("convergence_time_ms", int(seconds * 1000.0))])
    pairs += convergence_pairs(pairs)
    return [(k, csv_string(v)) for k, v in pairs]


class BenchmarkLogger:
    """Appends rows to a CSV file; the header goes only into an empty file,
    and a row's columns follow the file's header (missing ones left empty)."""

    def __init__(self, path: str):
        self.path = path
        self.header = None

    def log_pairs(self, pairs) -> None:
        keys = [k for k, _ in pairs]
        new_file = not os.path.exists(self.path) or os.path.getsize(self.path) == 0
        if self.header is None:
            # An existing file's header orders the columns (the reference
            # assumes it equals the row's own order, which holds for rows it
            # writes itself).
            if new_file:
                self.header = keys
            else:
                with open(self.path) as f:
                    self.header = f.readline().rstrip("\n").split(",")
        values = dict(pairs)
        with open(self.path, "a") as f:
            if new_file:
                f.write(",".join(self.header) + "\n")
            f.write(",".join(values.get(k, "") for k in self.header) + "\n")


def probability_analysis_logged(pd: Pdag, settings, shape, csv_path: str, input_file: str = "",
                                observe=(), device: dict | None = None):
    """``probability_analysis`` timed as the reference times it (engine
    build to last collect), with one row appended to ``csv_path``.
    -> (p_total, tallies, passes)."""
    import time

    from .engine import probability_analysis
    t = time.perf_counter()
    p, tallies, passes = probability_analysis(pd, settings, shape, observe)
    seconds = time.perf_counter() - t
    row = convergence_row(pd, settings, shape, tallies[0], passes, seconds, input_file,
                          device if device is not None else hip_device_info(),
                          n_tallies=len(tallies))
    BenchmarkLogger(csv_path).log_pairs(row)
    return p, tallies, passes
