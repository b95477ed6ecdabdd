"""ctypes binding of oracle/liboracle.so (test infrastructure only).

The oracle restates the reference's Monte Carlo algorithm on the CPU; see the
header of oracle/mc_oracle.cpp for the reference file:line each function
follows. The reference's own Philox header, built by oracle/Makefile into
oracle/_ref/libref_philox.so when the reference tree is present, is exposed
here too so tests can compare the restatement against it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


_V4_FLAGS = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}


def host_has_avx512() -> bool:
    """Whether this host's CPU flags include everything x86-64-v4 needs."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return _V4_FLAGS <= set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return False


def library_variant() -> str:
    """The oracle build this host runs: the AVX-512 one where it can."""
    v4 = os.path.join(_HERE, "liboracle_v4.so")
    return v4 if (os.path.exists(v4) and host_has_avx512()) else os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = library_variant()
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_threshold.argtypes = [C.c_double]
        L.orc_threshold.restype = C.c_uint64
        L.orc_sample_word.argtypes = [C.c_uint32] * 4 + [C.c_uint64]
        L.orc_sample_word.restype = C.c_uint64
        L.orc_gate_word.argtypes = [C.c_uint8, C.c_int32, C.c_void_p, C.c_int32, C.c_int32]
        L.orc_gate_word.restype = C.c_uint64
        L.orc_layout.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32] + [C.c_void_p] * 5 + \
            [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_layout.restype = C.c_int32
        L.orc_run.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32] + [C.c_void_p] * 5 + \
            [C.c_int32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p,
             C.c_void_p, C.c_int32]
        L.orc_run.restype = C.c_int32
        L.orc_run_keyed.argtypes = L.orc_run.argtypes + [C.c_uint32, C.c_uint32]
        L.orc_run_keyed.restype = C.c_int32
        L.orc_point_estimates.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
        L.orc_max_threads.restype = C.c_int32
        L.orc_bdd_probability.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32] + \
            [C.c_void_p] * 5 + [C.c_int32, C.c_uint64, C.c_void_p]
        L.orc_bdd_probability.restype = C.c_double
        L.orc_bdd_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's philox128.h behind a C shim, or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_philox.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_philox_generate.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.ref_pack_bernoulli_draws.argtypes = [C.c_void_p, C.c_uint64]
        L.ref_pack_bernoulli_draws.restype = C.c_uint64
        _REF = L
    return _REF


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def threshold(p: float) -> int:
    return int(lib().orc_threshold(float(p)))


def sample_word(pdag_idx, event_idx, sample_idx, iteration, thr) -> int:
    return int(lib().orc_sample_word(pdag_idx, event_idx, sample_idx, iteration, thr))


def gate_word(conn, k, vals, n_neg) -> int:
    v = np.ascontiguousarray(vals, dtype=np.uint64)
    return int(lib().orc_gate_word(conn, k, _p(v), v.shape[0], n_neg))


def _pdag_args(pd):
    a = (np.ascontiguousarray(pd.var_index, np.int32), np.ascontiguousarray(pd.var_p, np.float64),
         np.ascontiguousarray(pd.gate_index, np.int32), np.ascontiguousarray(pd.gate_type, np.uint8),
         np.ascontiguousarray(pd.gate_k, np.int32), np.ascontiguousarray(pd.arg_off, np.int64),
         np.ascontiguousarray(pd.args, np.int32))
    return a, [a[0].shape[0], _p(a[0]), _p(a[1]), a[2].shape[0], _p(a[2]), _p(a[3]), _p(a[4]),
               _p(a[5]), _p(a[6]), int(pd.root)]


def layout(pd):
    """-> (event_idx per variable slot, depth per gate slot, gate order)."""
    keep, args = _pdag_args(pd)
    ev = np.zeros(pd.n_vars, dtype=np.int32)
    depth = np.zeros(pd.n_gates, dtype=np.int32)
    order = np.zeros(pd.n_gates, dtype=np.int32)
    n = lib().orc_layout(*args, _p(ev), _p(depth), _p(order))
    if n < 0:
        raise ValueError(lib().orc_last_error().decode())
    return ev, depth, order[:n]


def run(pd, n_bp, first_iter=1, n_iters=1, tally=None, want_words=False, threads=0, key=None):
    """Monte Carlo on the CPU. Returns (ones per tallied index, words or None).

    words: [n_vars + n_gates, n_bp] uint64 of the last pass, rows = variable
    slots then gate slots (unreachable rows are zero). key: Philox key
    (default: the reference's fixed key)."""
    keep, args = _pdag_args(pd)
    tally = np.ascontiguousarray(tally if tally is not None else [pd.root], dtype=np.int32)
    ones = np.zeros(tally.shape[0], dtype=np.uint64)
    words = np.zeros((pd.n_vars + pd.n_gates, n_bp), dtype=np.uint64) if want_words else None
    key = key or (382307844, 293830103)
    rc = lib().orc_run_keyed(*args, int(n_bp), int(first_iter), int(n_iters), tally.shape[0],
                             _p(tally), _p(ones), _p(words) if words is not None else None,
                             int(threads), int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF)
    if rc != 0:
        raise ValueError(lib().orc_last_error().decode())
    return ones, words


def point_estimates(ones: int, total: int):
    out = np.zeros(6, dtype=np.float64)
    lib().orc_point_estimates(int(ones), int(total), _p(out))
    return out


def bdd_probability(pd, max_nodes: int = 20_000_000):
    """Exact P(root) by a BDD -> (probability, BDD nodes); probability is
    None when the BDD would exceed max_nodes."""
    keep, args = _pdag_args(pd)
    nodes = C.c_uint64()
    p = lib().orc_bdd_probability(*args, int(max_nodes), C.byref(nodes))
    if p == -2.0:
        raise ValueError(lib().orc_bdd_last_error().decode())
    return (None if p < 0 else p), int(nodes.value)


def max_threads() -> int:
    return int(lib().orc_max_threads())
