"""bench.py's contract pieces that run without a GPU: the one-GPU guard and
the algorithmic byte count of a pass (8 bytes per word for each basic-event
write, gate write, gate-input read and tallied-word read)."""
import os
import subprocess
import sys

import mcscram_amd as mc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_algorithmic_bytes_count_every_word_once():
    pd = mc.small_tree(1)
    n_bp = 1000
    fan_in = sum(len(pd.gate_args(s)) for s in range(pd.n_gates))
    b = bench.algorithmic_bytes(pd, n_bp)
    assert b["sampler"] == 8 * pd.n_vars * n_bp
    assert b["gates"] == 8 * (pd.n_gates + fan_in) * n_bp
    # Every gate tallied: each gate word is read once more.
    assert b["pass"] == b["sampler"] + b["gates"] + 8 * pd.n_gates * n_bp


def test_host_cpu_names_the_machine():
    h = bench.host_cpu()
    assert h["logical_cpus"] == os.cpu_count()
    assert isinstance(h["cpu_model"], str)
