"""ctypes mirror of the engine's C ABI (include/mcscram_amd.h).

Names follow the reference's Monte Carlo interface:
``LayerManager`` is queue::layer_manager (src/mc/queue/layer_manager.h:284),
``pass_wait_collect`` / ``collect_tallies`` its methods
(layer_manager.tpp:225-279), and ``probability_analysis`` is
ProbabilityAnalyzer<mc::DirectEval>::CalculateTotalProbability
(src/mc/core/probability_analysis.cc:88-209). Errors raise ``EngineError``
with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .pdag import Pdag

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MCS_REFERENCE_KEY = (382307844, 293830103)


class EngineError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class _Pdag(C.Structure):
    _fields_ = [("n_vars", C.c_int32), ("var_index", C.c_void_p), ("var_p", C.c_void_p),
                ("n_gates", C.c_int32), ("gate_index", C.c_void_p), ("gate_type", C.c_void_p),
                ("gate_k", C.c_void_p), ("arg_offsets", C.c_void_p), ("args", C.c_void_p),
                ("root", C.c_int32)]


class _Shape(C.Structure):
    _fields_ = [("batch_size", C.c_uint64), ("bitpacks_per_batch", C.c_uint64)]


class _Tally(C.Structure):
    _fields_ = [("index", C.c_int32), ("num_one_bits", C.c_uint64), ("total_bits", C.c_uint64),
                ("mean", C.c_double), ("std_err", C.c_double), ("ci", C.c_double * 4)]


class _Info(C.Structure):
    _fields_ = [("num_bitpacks", C.c_uint64), ("row_stride", C.c_uint64),
                ("iteration", C.c_uint32), ("num_events", C.c_int32), ("num_gates", C.c_int32),
                ("num_layers", C.c_int32), ("num_tallies", C.c_int32), ("num_args", C.c_uint64),
                ("device_bytes", C.c_uint64), ("event_buffers", C.c_int32),
                ("launches_per_pass", C.c_uint32), ("fused_gates", C.c_int32)]


class _BatchConfig(C.Structure):
    _fields_ = [("lds_bytes", C.c_uint32), ("max_tile_words", C.c_uint32),
                ("item_words", C.c_uint32), ("use_graph", C.c_uint32),
                ("key0", C.c_uint32), ("key1", C.c_uint32), ("workgroup_lanes", C.c_uint32),
                ("min_fused_tile_words", C.c_uint32), ("layered_streams", C.c_uint32)]


class _BatchInfo(C.Structure):
    _fields_ = [("num_trees", C.c_int32), ("num_fused_trees", C.c_int32),
                ("num_layered_trees", C.c_int32), ("layered_launches", C.c_uint32),
                ("num_items", C.c_uint32), ("grid", C.c_uint32),
                ("lds_bytes", C.c_uint32), ("workgroup_lanes", C.c_uint32),
                ("iteration", C.c_uint32), ("device_bytes", C.c_uint64)]


class _Settings(C.Structure):
    _fields_ = [("num_trials", C.c_uint64), ("burn_in_trials", C.c_uint64),
                ("ci_confidence", C.c_double), ("ci_rel_margin", C.c_double),
                ("max_passes", C.c_uint32), ("key0", C.c_uint32), ("key1", C.c_uint32),
                ("ci_policy", C.c_uint32), ("early_stop", C.c_uint32)]


class _Convergence(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("target_epsilon", C.c_double),
                ("target_trials", C.c_uint64), ("epsilon_log10", C.c_double),
                ("target_epsilon_log10", C.c_double), ("target_trials_log10", C.c_uint64)]


CI_POLICIES = {"bayes": 0, "wald": 1}  # src/settings.h:51-53


def library_path() -> str:
    return os.path.join(_HERE, "libmcscram_amd.so")


def load_library():
    """Load libmcscram_amd.so; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m mcscram_amd.build`")
    L = C.CDLL(path)
    vp, i32, u32, u64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64
    L.mcs_last_error.restype = C.c_char_p
    sigs = {
        "mcs_engine_create": [vp, vp, vp, i32, vp],
        "mcs_engine_create_ex": [vp, vp, vp, i32, u32, u32, i32, vp],
        "mcs_engine_destroy": [vp],
        "mcs_engine_pass": [vp, u32],
        "mcs_engine_wait": [vp],
        "mcs_engine_pass_wait_collect": [vp, u32, u32, vp, i32],
        "mcs_engine_collect": [vp, vp, i32],
        "mcs_engine_get_info": [vp, vp],
        "mcs_engine_event_order": [vp, vp, i32],
        "mcs_engine_read_words": [vp, i32, vp, u64],
        "mcs_engine_restore": [vp, u32, vp, i32],
        "mcs_engine_time_passes": [vp, u32, vp, vp, vp],
        "mcs_plan_passes": [u64, vp, vp, vp],
        "mcs_sample_shape_for": [u64, u64, u64, u64, u64, vp, vp, vp],
        "mcs_sample_shape_for_device": [u64, u64, u64, u64, u64, i32, C.c_double, u64, vp, vp, vp],
        "mcs_pdag_event_order": [vp, vp, i32, vp],
        "mcs_pdag_gate_depths": [vp, vp, i32],
        "mcs_probability_analysis": [vp, vp, vp, vp, i32, vp, i32, vp, vp],
        "mcs_probability_analysis_noalloc": [vp, vp, vp, vp, i32, vp, i32, vp, vp, vp],
        "mcs_batch_create": [vp, i32, vp, vp, vp, vp, vp],
        "mcs_batch_destroy": [vp],
        "mcs_batch_pass": [vp, u32],
        "mcs_batch_wait": [vp],
        "mcs_batch_collect": [vp, i32, vp, i32],
        "mcs_batch_get_info": [vp, vp],
        "mcs_batch_tile_words": [vp, i32, vp],
        "mcs_batch_time_passes": [vp, u32, vp],
        "mcs_ua_create": [vp, vp, i32, vp, vp, vp, i32, vp, vp],
        "mcs_ua_destroy": [vp],
        "mcs_ua_set_samples": [vp, i32, vp, vp],
        "mcs_ua_run": [vp, u32],
        "mcs_ua_collect": [vp, vp, C.c_int64, vp],
        "mcs_ua_time_run": [vp, u32, vp],
        "mcs_ua_info": [vp, vp, vp, vp, vp],
        "mcs_independence_logits": [vp, vp, C.c_int64, vp],
        "mcs_lognormal_samples": [vp, vp, i32, u64, C.c_int64, C.c_int64, vp, vp],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i32
    L.mcs_settings_default.argtypes = [vp]
    L.mcs_settings_default.restype = None
    L.mcs_batch_config_default.argtypes = [vp]
    L.mcs_batch_config_default.restype = None
    L.mcs_tally_point_estimates.argtypes = [vp]
    L.mcs_tally_point_estimates.restype = None
    L.mcs_ua_config_default.argtypes = [vp]
    L.mcs_ua_config_default.restype = None
    L.mcs_normal_quantile_two_sided.argtypes = [C.c_double]
    L.mcs_normal_quantile_two_sided.restype = C.c_double
    L.mcs_beta_quantile.argtypes = [C.c_double, C.c_double, C.c_double]
    L.mcs_beta_quantile.restype = C.c_double
    L.mcs_convergence_update.argtypes = [u32, u64, u64, C.c_double, C.c_double, vp]
    _LIB = L
    return L


def _check(status: int) -> None:
    if status != 0:
        raise EngineError(status, load_library().mcs_last_error().decode())


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class _PdagView:
    """Keeps contiguous copies alive while the C struct points at them."""

    def __init__(self, pd: Pdag):
        self.arrays = [np.ascontiguousarray(pd.var_index, np.int32),
                       np.ascontiguousarray(pd.var_p, np.float64),
                       np.ascontiguousarray(pd.gate_index, np.int32),
                       np.ascontiguousarray(pd.gate_type, np.uint8),
                       np.ascontiguousarray(pd.gate_k, np.int32),
                       np.ascontiguousarray(pd.arg_off, np.int64),
                       np.ascontiguousarray(pd.args, np.int32)]
        a = self.arrays
        self.struct = _Pdag(a[0].shape[0], _p(a[0]), _p(a[1]), a[2].shape[0], _p(a[2]),
                            _p(a[3]), _p(a[4]), _p(a[5]), _p(a[6]), int(pd.root))


@dataclass
class Tally:
    """Reference tally record (src/mc/stats/tally.h:18-33); ci = (lo95, hi95, lo99, hi99)."""

    index: int
    num_one_bits: int
    total_bits: int
    mean: float
    std_err: float
    ci: tuple

    @staticmethod
    def _from(t: _Tally) -> "Tally":
        return Tally(t.index, int(t.num_one_bits), int(t.total_bits), t.mean, t.std_err, tuple(t.ci))


def _is_shape(s) -> bool:
    ints = (int, np.integer)
    return isinstance(s, ints) or (isinstance(s, tuple) and len(s) == 2 and
                                   all(isinstance(x, ints) for x in s))


def _shape(n_bp: int | tuple) -> _Shape:
    if isinstance(n_bp, tuple):
        return _Shape(int(n_bp[0]), int(n_bp[1]))
    return _Shape(1, int(n_bp))


class LayerManager:
    """One compiled PDAG on one GPU (queue::layer_manager).

    ``shape`` is the words per pass (int) or (batch_size, bitpacks_per_batch).
    ``tally`` lists PDAG indices to tally (default: the root).
    """

    def __init__(self, pdag: Pdag, shape, tally=None, key=MCS_REFERENCE_KEY, device: int = -1):
        L = load_library()
        self._view = _PdagView(pdag)
        self._shape = _shape(shape)
        idx = np.ascontiguousarray([pdag.root] if tally is None else tally, dtype=np.int32)
        handle = C.c_void_p()
        _check(L.mcs_engine_create_ex(C.byref(self._view.struct), C.byref(self._shape), _p(idx),
                                      idx.shape[0], key[0] & 0xFFFFFFFF, key[1] & 0xFFFFFFFF,
                                      device, C.byref(handle)))
        self._h = handle
        self.info = self.get_info()

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().mcs_engine_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_info(self) -> dict:
        info = _Info()
        _check(load_library().mcs_engine_get_info(self._h, C.byref(info)))
        return {f: getattr(info, f) for f, _ in _Info._fields_}

    def pass_(self, n: int = 1) -> None:
        _check(load_library().mcs_engine_pass(self._h, n))

    def wait(self) -> None:
        _check(load_library().mcs_engine_wait(self._h))

    def collect_tallies(self) -> list[Tally]:
        n = self.get_info()["num_tallies"]
        out = (_Tally * n)()
        _check(load_library().mcs_engine_collect(self._h, out, n))
        return [Tally._from(t) for t in out]

    def pass_wait_collect(self, total_passes: int = 1, passes_between_waits: int = 0) -> list[Tally]:
        n = self.get_info()["num_tallies"]
        out = (_Tally * n)()
        _check(load_library().mcs_engine_pass_wait_collect(self._h, total_passes,
                                                           passes_between_waits, out, n))
        return [Tally._from(t) for t in out]

    def event_order(self) -> np.ndarray:
        n = self.info["num_events"]
        out = np.zeros(max(n, 1), dtype=np.int32)
        _check(load_library().mcs_engine_event_order(self._h, _p(out), out.shape[0]))
        return out[:n]

    def read_words(self, pdag_index: int) -> np.ndarray:
        n = self.info["num_bitpacks"]
        out = np.zeros(n, dtype=np.uint64)
        _check(load_library().mcs_engine_read_words(self._h, int(pdag_index), _p(out), n))
        return out

    def time_passes(self, n: int) -> dict:
        """Run n passes with HIP events around every launch on the engine's
        stream; returns summed kernel ms, launch count and first-to-last ms."""
        k, t = C.c_double(), C.c_double()
        launches = C.c_uint32()
        _check(load_library().mcs_engine_time_passes(self._h, n, C.byref(k), C.byref(launches),
                                                     C.byref(t)))
        return {"kernel_ms": k.value, "launches": launches.value, "pass_ms": t.value}

    def restore(self, iteration: int, num_one_bits) -> None:
        ones = np.ascontiguousarray(num_one_bits, dtype=np.uint64)
        _check(load_library().mcs_engine_restore(self._h, iteration, _p(ones), ones.shape[0]))


@dataclass
class Settings:
    """Probability-analysis settings (src/settings.h:439-464 defaults)."""

    num_trials: int = 0
    burn_in_trials: int = 1 << 20
    ci_confidence: float = 0.99
    ci_rel_margin: float = 0.001
    max_passes: int = 0
    key: tuple = MCS_REFERENCE_KEY
    ci_policy: str = "bayes"  # src/settings.h:426
    # Stop once converged (src/settings.h:453). num_trials == 0 implies it
    # (src/settings.cc:105-112); with num_trials > 0 the run stops at the
    # trial count or at convergence, whichever comes first. The CLI's "-N n"
    # without "--early-stop" is early_stop=False
    # (targets/scram/utils/ConstructSettings.h:103-109).
    early_stop: bool = True

    def _c(self) -> _Settings:
        if self.ci_policy not in CI_POLICIES:
            raise ValueError(f"ci_policy must be one of {sorted(CI_POLICIES)}")
        return _Settings(self.num_trials, self.burn_in_trials, self.ci_confidence,
                         self.ci_rel_margin, self.max_passes, self.key[0], self.key[1],
                         CI_POLICIES[self.ci_policy], int(bool(self.early_stop)))


def probability_analysis(pdag: Pdag, settings: Settings, shape, observe=(), *,
                         no_alloc: bool = False, stream: int | None = None):
    """-> (p_total, [root tally, observed tallies...], passes run).

    no_alloc runs mcs_probability_analysis_noalloc: the engine lives in the
    library's fixed per-device scratch (NOALLOC_SCRATCH_BYTES) and launches
    on ``stream`` (a hipStream_t as an int; None = the null stream)."""
    L = load_library()
    view = _PdagView(pdag)
    sh = _shape(shape)
    obs = np.ascontiguousarray(list(observe) or [0], dtype=np.int32)
    n_obs = len(observe)
    out = (_Tally * (1 + n_obs))()
    p = C.c_double()
    passes = C.c_uint32()
    st = settings._c()
    args = (C.byref(view.struct), C.byref(st), C.byref(sh), _p(obs), n_obs, out, 1 + n_obs,
            C.byref(p), C.byref(passes))
    if no_alloc:
        _check(L.mcs_probability_analysis_noalloc(*args, C.c_void_p(stream or 0)))
    else:
        _check(L.mcs_probability_analysis(*args))
    distinct = 1 + len(set(observe) - {pdag.root})
    return p.value, [Tally._from(t) for t in out[:distinct]], passes.value


NOALLOC_SCRATCH_BYTES = 64 << 20  # MCS_NOALLOC_SCRATCH_BYTES


def plan_passes(num_trials: int, shape) -> tuple[int, int]:
    sh = _shape(shape)
    t, n = C.c_uint64(), C.c_uint64()
    _check(load_library().mcs_plan_passes(num_trials, C.byref(sh), C.byref(t), C.byref(n)))
    return t.value, n.value


def sample_shape_for(num_trials: int, num_nodes: int, max_alloc_bytes: int,
                     max_work_item_sizes=(1024, 1024), nvidia_gpu: bool = False,
                     overhead_ratio: float = 0.05, cache_line_bytes: int = 128):
    """The (batch, bitpacks) shape the reference's sample_shaper picks on a
    device whose largest allocation is max_alloc_bytes and whose work-item
    size limits (y, z) are max_work_item_sizes, with the rounded trial count
    and the pass count (0 when num_trials is 0: no limit). nvidia_gpu applies
    the shaper's rule for NVIDIA GPUs (the largest allocation split over the
    nodes with overhead_ratio, in whole cache lines of cache_line_bytes;
    src/mc/queue/sample_shaper.h:120-134).
    -> ((batch, bitpacks), rounded_trials, passes)."""
    sh = _Shape()
    t, n = C.c_uint64(), C.c_uint64()
    _check(load_library().mcs_sample_shape_for_device(
        int(num_trials), int(num_nodes), int(max_alloc_bytes), int(max_work_item_sizes[0]),
        int(max_work_item_sizes[1]), int(bool(nvidia_gpu)), float(overhead_ratio),
        int(cache_line_bytes), C.byref(sh), C.byref(t), C.byref(n)))
    return (sh.batch_size, sh.bitpacks_per_batch), t.value, n.value


def pdag_event_order(pdag: Pdag) -> np.ndarray:
    """Host-only: PDAG indices of the basic events in event_idx order."""
    L = load_library()
    view = _PdagView(pdag)
    n = C.c_int32()
    _check(L.mcs_pdag_event_order(C.byref(view.struct), None, 0, C.byref(n)))
    out = np.zeros(max(n.value, 1), dtype=np.int32)
    _check(L.mcs_pdag_event_order(C.byref(view.struct), _p(out), out.shape[0], C.byref(n)))
    return out[:n.value]


def pdag_gate_depths(pdag: Pdag) -> np.ndarray:
    """Host-only: layer of each gate slot (-1 if unreachable)."""
    view = _PdagView(pdag)
    out = np.zeros(pdag.n_gates, dtype=np.int32)
    _check(load_library().mcs_pdag_gate_depths(C.byref(view.struct), _p(out), out.shape[0]))
    return out


def normal_quantile_two_sided(confidence: float) -> float:
    return load_library().mcs_normal_quantile_two_sided(confidence)


def beta_quantile(a: float, b: float, q: float) -> float:
    """Quantile of Beta(a, b) (host only)."""
    return load_library().mcs_beta_quantile(a, b, q)


def convergence_update(policy: str, ones: int, total: int, rel_margin: float, z: float) -> dict:
    """One tally's stopping-rule update (host only): current and target
    half-widths and trial counts in linear and log10 space."""
    out = _Convergence()
    _check(load_library().mcs_convergence_update(CI_POLICIES[policy], ones, total, rel_margin, z,
                                                 C.byref(out)))
    return {f: getattr(out, f) for f, _ in _Convergence._fields_}


class Batch:
    """Independent trees evaluated whole in LDS by one persistent launch per
    pass (no reference counterpart; each tree's tallies equal those of its
    own LayerManager with the same shape).

    ``shapes``: one shape per tree (int words or (batch, bitpacks)), or one
    shape for all. ``tallies``: one list of PDAG indices per tree (None or an
    empty list tallies the root). Config keywords: lds_bytes,
    max_tile_words, item_words, use_graph, key, workgroup_lanes,
    min_fused_tile_words, layered_streams.
    """

    def __init__(self, pdags, shapes, tallies=None, **config):
        L = load_library()
        self._views = [_PdagView(p) for p in pdags]
        n = len(pdags)
        if _is_shape(shapes):
            shapes = [shapes] * n
        self._structs = (_Pdag * n)(*[v.struct for v in self._views])
        self._shapes = (_Shape * n)(*[_shape(s) for s in shapes])
        tallies = tallies or [[] for _ in range(n)]
        off = np.zeros(n + 1, dtype=np.int32)
        flat = []
        for t, lst in enumerate(tallies):
            flat.extend(int(x) for x in lst)
            off[t + 1] = len(flat)
        self._off = off
        self._flat = np.ascontiguousarray(flat or [0], dtype=np.int32)
        self._n_tally = [max(len(set(l)), 1) if l else 1 for l in tallies]
        cfg = _BatchConfig()
        L.mcs_batch_config_default(C.byref(cfg))
        key = config.pop("key", None)
        if key is not None:
            cfg.key0, cfg.key1 = key[0] & 0xFFFFFFFF, key[1] & 0xFFFFFFFF
        known = {f[0] for f in _BatchConfig._fields_}
        for k, v in config.items():
            if k not in known:
                raise TypeError(f"unknown batch config keyword {k!r}")
            setattr(cfg, k, int(v))
        self._cfg = cfg
        handle = C.c_void_p()
        _check(L.mcs_batch_create(self._structs, n, self._shapes, _p(self._off), _p(self._flat),
                                  C.byref(cfg), C.byref(handle)))
        self._h = handle
        self.n_trees = n

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().mcs_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pass_(self, n: int = 1) -> None:
        _check(load_library().mcs_batch_pass(self._h, n))

    def wait(self) -> None:
        _check(load_library().mcs_batch_wait(self._h))

    def collect(self, tree: int) -> list[Tally]:
        n = self._n_tally[tree]
        out = (_Tally * n)()
        _check(load_library().mcs_batch_collect(self._h, tree, out, n))
        return [Tally._from(t) for t in out]

    def info(self) -> dict:
        i = _BatchInfo()
        _check(load_library().mcs_batch_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _BatchInfo._fields_}

    def tile_words(self, tree: int) -> int:
        w = C.c_uint32()
        _check(load_library().mcs_batch_tile_words(self._h, tree, C.byref(w)))
        return w.value

    def time_passes(self, n: int) -> float:
        ms = C.c_double()
        _check(load_library().mcs_batch_time_passes(self._h, n, C.byref(ms)))
        return ms.value
