"""Surrogate network for uncertainty analysis (ctypes mirror of mcs_mlp_*).

An MLP from the vector of sampled basic-event parameters (standard-normal
scores of their log probabilities) to the top event's logit, trained on
pairs the GPU Monte Carlo engine generates. A fixed input map turns each
score back into the event's probability on a per-event scale before the
first layer. BF16 weights and activations, FP32 accumulation on the matrix
cores, MSE loss, Adam with a cosine schedule. No counterpart in the
reference.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import _check, _p, load_library

_SIGS = {
    "mcs_mlp_create": [C.c_int32, C.c_int32, C.c_uint64, C.c_void_p],
    "mcs_mlp_destroy": [C.c_void_p],
    "mcs_mlp_num_params": [C.c_void_p, C.c_void_p],
    "mcs_mlp_get_params": [C.c_void_p, C.c_void_p, C.c_int64],
    "mcs_mlp_set_params": [C.c_void_p, C.c_void_p, C.c_int64],
    "mcs_mlp_set_input_map": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "mcs_mlp_set_weight_decay": [C.c_void_p, C.c_float],
    "mcs_mlp_predict": [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p],
    "mcs_mlp_loss_grad": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p],
    "mcs_mlp_fit": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                    C.c_float, C.c_uint64, C.c_void_p],
    "mcs_mlp_gemm_test": [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                          C.c_int32, C.c_void_p],
}


def _lib():
    L = load_library()
    if not getattr(L, "_mlp_bound", False):
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int32
        L._mlp_bound = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """n_in -> hidden x 3 (GELU) -> 1 on the GPU."""

    def __init__(self, n_in: int, hidden: int = 512, seed: int = 1):
        L = _lib()
        h = C.c_void_p()
        _check(L.mcs_mlp_create(n_in, hidden, seed, C.byref(h)))
        self._h = h
        self.n_in, self.hidden = n_in, hidden
        n = C.c_int64()
        _check(L.mcs_mlp_num_params(h, C.byref(n)))
        self.n_params = n.value

    def close(self):
        if getattr(self, "_h", None):
            _lib().mcs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(_lib().mcs_mlp_get_params(self._h, _p(out), out.shape[0]))
        return out

    def set_params(self, p) -> None:
        p = _f32(p)
        _check(_lib().mcs_mlp_set_params(self._h, _p(p), p.shape[0]))

    def set_input_map(self, scale=None, shift=None, cap=None) -> None:
        """Fixed map of each input column before the first layer:
        u = min(cap, exp(scale * x + shift)); no arguments clears it."""
        if scale is None and shift is None and cap is None:
            _check(_lib().mcs_mlp_set_input_map(self._h, None, None, None))
            return
        arrs = [_f32(a) for a in (scale, shift, cap)]
        if any(a.shape != (self.n_in,) for a in arrs):
            raise ValueError(f"input map arrays must have {self.n_in} entries")
        _check(_lib().mcs_mlp_set_input_map(self._h, *(_p(a) for a in arrs)))

    def set_weight_decay(self, weight_decay: float) -> None:
        """Decoupled weight decay of later fit steps (0: plain Adam)."""
        _check(_lib().mcs_mlp_set_weight_decay(self._h, float(weight_decay)))

    def predict(self, x) -> np.ndarray:
        x = _f32(x)
        out = np.zeros(x.shape[0], dtype=np.float32)
        _check(_lib().mcs_mlp_predict(self._h, _p(x), x.shape[0], _p(out)))
        return out

    def loss_grad(self, x, y):
        x, y = _f32(x), _f32(y)
        loss = C.c_float()
        g = np.zeros(self.n_params, dtype=np.float32)
        _check(_lib().mcs_mlp_loss_grad(self._h, _p(x), _p(y), x.shape[0], C.byref(loss), _p(g)))
        return loss.value, g

    def fit(self, x, y, epochs: int = 10, batch: int = 256, lr: float = 1e-3, seed: int = 7):
        """Adam, cosine learning-rate schedule from lr down to 0 over the epochs.
        -> mean training loss of each epoch."""
        x, y = _f32(x), _f32(y)
        losses = np.zeros(epochs, dtype=np.float32)
        _check(_lib().mcs_mlp_fit(self._h, _p(x), _p(y), x.shape[0], epochs, batch, lr, seed,
                                  _p(losses)))
        return losses


def gemm_test(a, b, trans_a=False, trans_b=False):
    """c = A @ B on the matrix cores (A, B rounded to BF16)."""
    a, b = _f32(a), _f32(b)
    M = a.shape[1] if trans_a else a.shape[0]
    K = a.shape[0] if trans_a else a.shape[1]
    N = b.shape[0] if trans_b else b.shape[1]
    c = np.zeros((M, N), dtype=np.float32)
    _check(_lib().mcs_mlp_gemm_test(_p(a), _p(b), M, N, K, int(trans_a), int(trans_b), _p(c)))
    return c


class Ensemble:
    """Members trained on the same pairs from different initial weights and
    shuffles; their disagreement routes a sample back to plain Monte Carlo.
    Targets are standardised for training (mean / std of the training
    logits) and predictions mapped back. ``input_map`` (scale, shift, cap)
    is set on every member (Mlp.set_input_map)."""

    def __init__(self, n_in: int, hidden: int = 512, members: int = 2, seed: int = 11,
                 input_map=None, weight_decay: float = 0.0, baseline: str | None = None,
                 baseline_clip: float = np.inf):
        self.members = [Mlp(n_in, hidden, seed + 101 * i) for i in range(members)]
        self.mu, self.sd = 0.0, 1.0
        # The closed-form logit the members' output is added to (None: the
        # members predict the logit itself), clipped to +-baseline_clip; the
        # caller supplies its values.
        self.baseline = baseline
        self.baseline_clip = float(baseline_clip)
        for m in self.members:
            if input_map is not None:
                m.set_input_map(*input_map)
            if weight_decay:
                m.set_weight_decay(weight_decay)

    def fit(self, x, y, epochs=10, batch=256, lr=1e-3, seed=5):
        y = np.asarray(y, dtype=np.float64)
        self.mu, self.sd = float(y.mean()), float(y.std() or 1.0)
        yn = (y - self.mu) / self.sd
        return [m.fit(x, yn, epochs, batch, lr, seed + 17 * i) for i, m in enumerate(self.members)]

    def predict(self, x) -> np.ndarray:
        """-> [members, rows] logits (with a baseline set: the residual to add
        to the baseline's logit of the same rows)."""
        return np.stack([m.predict(x).astype(np.float64) * self.sd + self.mu for m in self.members])

    def save(self, path: str, info: dict | None = None) -> None:
        """Write what inference uses: each member's weight matrices as their
        BF16 values (the network multiplies in BF16), biases in FP32, the
        target standardisation and ``info`` (JSON), to an .npz file."""
        import json
        arrs = {"mu": np.float64(self.mu), "sd": np.float64(self.sd),
                "shape": np.array([self.members[0].n_in, self.members[0].hidden,
                                   len(self.members)], dtype=np.int64),
                "baseline": np.array(self.baseline or ""),
                "baseline_clip": np.float64(self.baseline_clip),
                "info": np.array(json.dumps(info or {}))}
        for i, m in enumerate(self.members):
            p = m.params()
            for name, (a, b) in _param_ranges(m.n_in, m.hidden).items():
                v = p[a:b]
                arrs[f"m{i}_{name}"] = _to_bf16_bits(v) if name[0] in "Ww" else v
        np.savez(path, **arrs)

    @classmethod
    def load(cls, path: str, input_map=None):
        """An ensemble written by ``save``. -> (ensemble, info dict)."""
        import json
        z = np.load(path, allow_pickle=False)
        n_in, hidden, members = (int(v) for v in z["shape"])
        baseline = str(z["baseline"]) if "baseline" in z.files else ""
        clip = float(z["baseline_clip"]) if "baseline_clip" in z.files else np.inf
        ens = cls(n_in, hidden, members, input_map=input_map, baseline=baseline or None,
                  baseline_clip=clip)
        ens.mu, ens.sd = float(z["mu"]), float(z["sd"])
        for i, m in enumerate(ens.members):
            p = np.zeros(m.n_params, dtype=np.float32)
            for name, (a, b) in _param_ranges(n_in, hidden).items():
                v = z[f"m{i}_{name}"]
                p[a:b] = _from_bf16_bits(v) if name[0] in "Ww" else v
            m.set_params(p)
        return ens, json.loads(str(z["info"]))


def _param_ranges(n_in: int, h: int) -> dict:
    """Slices of the parameter vector: W1 [n_in x h], b1, W2 [h x h], b2, W3,
    b3, w4 [h], b4."""
    sizes = [("W1", n_in * h), ("b1", h), ("W2", h * h), ("b2", h), ("W3", h * h), ("b3", h),
             ("w4", h), ("b4", 1)]
    out, o = {}, 0
    for name, n in sizes:
        out[name] = (o, o + n)
        o += n
    return out


def _to_bf16_bits(x) -> np.ndarray:
    """FP32 -> BF16 bit patterns, round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _from_bf16_bits(b) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)
