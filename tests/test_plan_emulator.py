"""The engine's host-side plan (schedule, wide-gate split, single-consumer
fusion, launch plan) replayed on the CPU with the kernel's gate semantics
(tests/plan_emulator.cpp), against the oracle's pass over the same basic-event
words (src/mc/kernel/basic_event.h, src/mc/kernel/gate.h as restated in
oracle/mc_oracle.cpp). The emulator checks every row, argument range and
fused reference the kernel would touch; a planning mistake (a gate scheduled
before an input, a fused gate under a NOT, a partial fold dropped, a tally
slot reused) fails here without a GPU.
"""
import gzip
import os
import struct
import subprocess

import numpy as np
import pytest

import mcscram_amd as mc
from oracle import binding as orc
from tests.models import core_models, small_random

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mcscram_amd", "csrc")
W = 64


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu") / "plan_emulator")
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           "-I", CSRC, "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "plan_emulator.cpp"),
           os.path.join(CSRC, "layering.cpp"), os.path.join(CSRC, "launch_plan.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _emulate(exe, tmp, pd, fuse: int, split: int):
    """-> (words per gate slot [n_gates, W], tallies by PDAG index, launches, fused)."""
    hdr = struct.pack("<iiqi", pd.n_vars, pd.n_gates, pd.n_args, int(pd.root))
    body = b"".join([pd.var_index.astype(np.int32).tobytes(), pd.var_p.astype(np.float64).tobytes(),
                     pd.gate_index.astype(np.int32).tobytes(), pd.gate_type.astype(np.uint8).tobytes(),
                     pd.gate_k.astype(np.int32).tobytes(), pd.arg_off.astype(np.int64).tobytes(),
                     pd.args.astype(np.int32).tobytes()])
    (tmp / "pd.bin").write_bytes(hdr + body)
    # Basic-event words of pass 1 from the oracle, in event_idx order.
    _, words = orc.run(pd, W, want_words=True)
    ev, _, _ = orc.layout(pd)
    n_ev = int((ev >= 0).sum())
    be = np.zeros((n_ev, W), dtype=np.uint64)
    for slot in range(pd.n_vars):
        if ev[slot] >= 0:
            be[ev[slot]] = words[slot]
    be.tofile(tmp / "be.bin")
    r = subprocess.run([exe, str(tmp / "pd.bin"), str(tmp / "be.bin"), str(fuse), str(split),
                        str(tmp / "out.bin")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    launches, fused = int(r.stdout.split()[1]), int(r.stdout.split()[3])
    out = np.fromfile(tmp / "out.bin", dtype=np.uint8)
    n_nodes = (out.size + n_ev * W * 8) // (W * 8 + 12)
    gw = (n_nodes - n_ev) * W * 8
    g = out[:gw].view(np.uint64).reshape(-1, W)
    ones = out[gw:gw + n_nodes * 8].view(np.uint64)
    idx = out[gw + n_nodes * 8:].view(np.int32)
    assert idx.size == n_nodes
    slot_of = {int(x): s for s, x in enumerate(pd.gate_index)}
    gate_words = np.zeros((pd.n_gates, W), dtype=np.uint64)
    reached = np.zeros(pd.n_gates, dtype=bool)
    for n in range(n_ev, n_nodes):
        s = slot_of[int(idx[n])]
        gate_words[s] = g[n - n_ev]
        reached[s] = True
    tallies = {int(idx[n]): int(ones[n]) for n in range(n_nodes)}
    return gate_words, reached, tallies, launches, fused


def _check(exe, tmp, pd, fuse, split):
    gate_words, reached, tallies, launches, fused = _emulate(exe, tmp, pd, fuse, split)
    want_idx = np.concatenate([pd.var_index, pd.gate_index]).astype(np.int32)
    want_ones, words = orc.run(pd, W, tally=want_idx, want_words=True)
    want = dict(zip(want_idx.tolist(), want_ones.tolist()))
    assert reached.any()
    bad = [int(pd.gate_index[s]) for s in np.flatnonzero(reached)
           if not np.array_equal(gate_words[s], words[pd.n_vars + s])]
    assert not bad, f"gate words differ for PDAG indices {bad[:10]}"
    bad = [i for i, t in tallies.items() if t != want[i]]
    assert not bad, f"tallies differ for PDAG indices {bad[:10]}"
    if not fuse:
        assert fused == 0
    return launches, fused


CASES = ([("small", lambda: mc.small_tree(1))]
         + [(f"random{s}", lambda s=s: small_random(s)) for s in range(8)]
         + [(f"allconn{s}", lambda s=s: mc.random_layered_tree(s, 60, 80, 6, all_connectives=True,
                                                             neg_frac=0.2)) for s in range(4)]
         + [(f"core_{k}", lambda k=k: core_models()[k][0]) for k in sorted(core_models())])


@pytest.mark.parametrize("fuse", [0, 1])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_plan_replays_oracle(emulator, tmp_path, name, make, fuse):
    _check(emulator, tmp_path, make(), fuse, 64)


@pytest.mark.parametrize("split", [2, 3, 5])
def test_plan_with_narrow_split_replays_oracle(emulator, tmp_path, split):
    # A tiny fan-in cap turns most gates into chains of partial folds.
    for s in range(4):
        pd = mc.random_layered_tree(100 + s, 40, 60, 5, all_connectives=True, neg_frac=0.3)
        for fuse in (0, 1):
            _check(emulator, tmp_path, pd, fuse, split)


ARALIA = os.path.join(ROOT, "tests", "golden", "mef", "aralia")


@pytest.mark.parametrize("name", sorted(f[:-7] for f in os.listdir(ARALIA) if f.endswith(".xml.gz")))
def test_aralia_model_replays_oracle(emulator, tmp_path, name):
    # MEF models with wide gates, ATLEAST gates and complemented inputs.
    pd = mc.mef.load_pdag(gzip.open(os.path.join(ARALIA, name + ".xml.gz")))
    _check(emulator, tmp_path, pd, 1, 64)


def test_large_tree_plan_replays_oracle(emulator, tmp_path):
    launches, fused = _check(emulator, tmp_path, mc.large_tree(), 1, 64)
    assert fused > 1000
    assert launches >= 2


def test_workgroup_split_covers_every_task_once(emulator):
    # The pass kernel's division-free map from workgroups to gate and sampler
    # tasks (kernels.h: xcd_share, interleave_frac, gates_before).
    r = subprocess.run([emulator, "--split"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("split ok")


@pytest.fixture(scope="module")
def sanitized_emulator(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu_asan") / "plan_emulator")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", "/opt/rocm/include",
           os.path.join(ROOT, "tests", "plan_emulator.cpp"), os.path.join(CSRC, "layering.cpp"),
           os.path.join(CSRC, "launch_plan.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitizer" in r.stderr.lower():
        pytest.skip("no sanitizer runtime for g++ here")
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def test_host_plan_code_under_address_and_ub_sanitizers(sanitized_emulator, tmp_path):
    # The schedule, split, fusion and launch-plan code built with ASan and
    # UBSan: any out-of-range access or undefined behaviour aborts the run.
    r = subprocess.run([sanitized_emulator, "--split"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    for pd in (mc.small_tree(1), small_random(3),
               mc.random_layered_tree(7, 60, 80, 6, all_connectives=True, neg_frac=0.2),
               mc.mef.load_pdag(gzip.open(os.path.join(ARALIA, "baobab1.xml.gz")))):
        for fuse, split in ((0, 64), (1, 64), (1, 3)):
            _check(sanitized_emulator, tmp_path, pd, fuse, split)
