"""Learned proposal for the tracklet stage (no counterpart in the reference).

A 24-64-64-64-1 MLP scores (inner segment, outer-inner mini-doublet) pairs
before the exact tracklet cuts.  This module wraps the C-ABI (include/sdl_amd.h:
sdl_mlp_*): parameters, GPU training (FP32 forward/backward kernels + Adam),
BF16-MFMA inference, and the training rows exported by a plain chain.
"""
import ctypes
import json
import os

import numpy as np

from . import hip, runtime

N_IN, N_H = 24, 64
# parameter layout (float32): W1[64][24], b1, W2[64][64], b2, W3[64][64], b3, w4[64], b4
SHAPES = (("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H, N_H)),
          ("b3", (N_H,)), ("w4", (N_H,)), ("b4", (1,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    """Flat parameter vector -> dict of arrays (views)."""
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = flat[k:k + n].reshape(shape)
        k += n
    return out


def init_params(seed=0):
    """He-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(N_PARAMS, np.float32)
    p = unpack(flat)
    for name, fan_in in (("W1", N_IN), ("W2", N_H), ("W3", N_H), ("w4", N_H)):
        lim = np.sqrt(6.0 / fan_in)
        p[name][...] = rng.uniform(-lim, lim, p[name].shape).astype(np.float32)
    return flat


class MLP:
    def __init__(self, params=None):
        L = runtime.lib()
        assert L.sdl_mlp_num_params() == N_PARAMS
        self.h = ctypes.c_void_p()
        runtime._check(L.sdl_mlp_create(ctypes.byref(self.h)))
        self.set_params(init_params() if params is None else params)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            runtime.lib().sdl_mlp_free(self.h)
            self.h = None

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == N_PARAMS
        runtime._check(runtime.lib().sdl_mlp_set_params(self.h, flat.ctypes.data, N_PARAMS))

    def params(self):
        out = np.zeros(N_PARAMS, np.float32)
        runtime._check(runtime.lib().sdl_mlp_get_params(self.h, out.ctypes.data, N_PARAMS))
        return out

    def gradient(self, dX, dY, n, pos_weight=1.0):
        """(gradient, loss) of the mean weighted BCE over n device rows."""
        g = np.zeros(N_PARAMS, np.float32)
        loss = ctypes.c_float()
        runtime._check(runtime.lib().sdl_mlp_gradient(self.h, dX, dY, n, pos_weight, g.ctypes.data,
                                                      ctypes.byref(loss)))
        return g, loss.value

    def train(self, dX, dY, dIdx, start, batch, steps, lr, pos_weight):
        loss = ctypes.c_float()
        runtime._check(runtime.lib().sdl_mlp_train(self.h, dX, dY, dIdx, start, batch, steps, lr, pos_weight,
                                                   ctypes.byref(loss)))
        return loss.value

    def forward(self, dX, n, dOut):
        runtime._check(runtime.lib().sdl_mlp_forward(self.h, dX, n, dOut))

    def save(self, path, **meta):
        with open(path, "w") as f:
            json.dump({"params": self.params().tolist(), **meta}, f)

    @staticmethod
    def load(path):
        with open(path) as f:
            d = json.load(f)
        return MLP(np.asarray(d["params"], np.float32)), d


def pair_rows(ev):
    """Training rows of the last plain chain of `ev`: (DeviceArray X (n x 24), DeviceArray Y (n), n)."""
    L = runtime.lib()
    n = ctypes.c_longlong()
    runtime._check(L.sdl_event_tl_pair_rows(ev.h, None, None, 0, ctypes.byref(n)))
    X = hip.DeviceArray.empty(max(n.value, 1) * N_IN, np.float32)
    Y = hip.DeviceArray.empty(max(n.value, 1), np.float32)
    runtime._check(L.sdl_event_tl_pair_rows(ev.h, X.address, Y.address, n.value, ctypes.byref(n)))
    return X, Y, n.value


def set_proposal(ev, mlp, threshold):
    runtime._check(runtime.lib().sdl_event_set_proposal(ev.h, mlp.h if mlp is not None else None,
                                                        float(threshold)))


# ---------------------------------------------------------------- references (numpy, for the tests)
def forward_np(flat, X, bf16=False):
    """Logits of rows X.  bf16=True rounds weights and the inputs of every layer
    to bfloat16 as the MFMA path does (accumulation stays in float)."""
    p = {k: v.astype(np.float64) for k, v in unpack(np.asarray(flat, np.float32)).items()}
    r = round_bf16 if bf16 else (lambda a: a)
    x = r(np.asarray(X, np.float64))
    h = np.maximum(x @ r(p["W1"]).T + r(p["b1"]), 0)
    h = np.maximum(r(h) @ r(p["W2"]).T + p["b2"], 0)
    h = np.maximum(r(h) @ r(p["W3"]).T + p["b3"], 0)
    return h @ p["w4"] + p["b4"][0]


def round_bf16(a):
    """float64/32 -> nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def loss_grad_np(flat, X, Y, pos_weight=1.0):
    """Mean weighted BCE on logits and its gradient (float64), by hand-written backprop."""
    p = {k: v.astype(np.float64) for k, v in unpack(np.asarray(flat, np.float64)).items()}
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    n = len(Y)
    a1 = X @ p["W1"].T + p["b1"]; h1 = np.maximum(a1, 0)
    a2 = h1 @ p["W2"].T + p["b2"]; h2 = np.maximum(a2, 0)
    a3 = h2 @ p["W3"].T + p["b3"]; h3 = np.maximum(a3, 0)
    z = h3 @ p["w4"] + p["b4"][0]
    w = np.where(Y > 0.5, pos_weight, 1.0) / n
    loss = np.sum(w * (np.maximum(z, 0) - z * Y + np.log1p(np.exp(-np.abs(z)))))
    d4 = w * (1 / (1 + np.exp(-z)) - Y)
    g = {"w4": h3.T @ d4, "b4": np.array([d4.sum()])}
    d3 = np.outer(d4, p["w4"]) * (a3 > 0)
    g["W3"] = d3.T @ h2; g["b3"] = d3.sum(0)
    d2 = (d3 @ p["W3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1; g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ X; g["b1"] = d1.sum(0)
    flatg = np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])
    return loss, flatg
