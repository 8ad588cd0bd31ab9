"""CPU oracle of the importance network (TEST INFRASTRUCTURE ONLY).

Architecture (no counterpart in the reference; the learned proposal is a
build extension): 8 inputs -> 64 -> 64 -> 64 -> 1, ReLU on the hidden
layers, output o; the importance is g = exp(clip(o, -4, 4)).  Flat FP32
parameter layout (the device library uses the same):
    W1[8][64]  b1[64]  W2[64][64]  b2[64]  W3[64][64]  b3[64]  W4[64]  b4[1]
(W[in][out], z = a @ W + b).

emulate=True mirrors the device arithmetic: inputs, W1..W3 and the hidden
activations are rounded to bfloat16 (round to nearest even) before each
matrix product, products accumulate in FP32, W4 and the last layer stay
FP32; in the backward pass the layer gradients dZ are rounded to bfloat16
before they enter a product, bias gradients use the unrounded dZ.
emulate=False is exact FP64 arithmetic; its gradient is pinned by finite
differences in tests/test_mlp_oracle.py.

Loss: mean over rows of (o - t)^2.
"""
import numpy as np

N_IN, N_HID = 8, 64
SHAPES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("W4", (N_HID,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, k = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = flat[k:k + size].reshape(shape)
        k += size
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def forward(flat, X, emulate=True):
    p = unpack(np.asarray(flat, np.float64))
    if emulate:
        f32 = np.float32
        x = bf16(X)
        z1 = (x @ bf16(p["W1"].astype(f32))).astype(f32) + p["b1"].astype(f32)
        a1 = bf16(np.maximum(z1, 0))
        z2 = (a1 @ bf16(p["W2"].astype(f32))).astype(f32) + p["b2"].astype(f32)
        a2 = bf16(np.maximum(z2, 0))
        z3 = (a2 @ bf16(p["W3"].astype(f32))).astype(f32) + p["b3"].astype(f32)
        a3 = bf16(np.maximum(z3, 0))
        o = (a3 @ p["W4"].astype(f32)).astype(f32) + f32(p["b4"][0])
        cache = dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3)
        return o.astype(np.float64), cache
    x = np.asarray(X, np.float64)
    z1 = x @ p["W1"] + p["b1"]
    a1 = np.maximum(z1, 0)
    z2 = a1 @ p["W2"] + p["b2"]
    a2 = np.maximum(z2, 0)
    z3 = a2 @ p["W3"] + p["b3"]
    a3 = np.maximum(z3, 0)
    o = a3 @ p["W4"] + p["b4"][0]
    return o, dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3)


def loss_grad(flat, X, T, emulate=True):
    p = unpack(np.asarray(flat, np.float64))
    o, c = forward(flat, X, emulate)
    B = len(T)
    r = o - np.asarray(T, np.float64)
    loss = float(np.mean(r * r))
    q = bf16 if emulate else (lambda v: v)
    W2 = bf16(p["W2"].astype(np.float32)) if emulate else p["W2"]
    W3 = bf16(p["W3"].astype(np.float32)) if emulate else p["W3"]
    dO = 2.0 * r / B
    g = {}
    g["W4"] = c["a3"].astype(np.float64).T @ dO
    g["b4"] = np.array([dO.sum()])
    dZ3 = np.outer(dO, p["W4"]) * (c["z3"] > 0)
    g["b3"] = dZ3.sum(0)
    dZ3q = q(dZ3).astype(np.float64)
    g["W3"] = c["a2"].astype(np.float64).T @ dZ3q
    dZ2 = (dZ3q @ np.asarray(W3, np.float64).T) * (c["z2"] > 0)
    g["b2"] = dZ2.sum(0)
    dZ2q = q(dZ2).astype(np.float64)
    g["W2"] = c["a1"].astype(np.float64).T @ dZ2q
    dZ1 = (dZ2q @ np.asarray(W2, np.float64).T) * (c["z1"] > 0)
    g["b1"] = dZ1.sum(0)
    dZ1q = q(dZ1).astype(np.float64)
    g["W1"] = c["x"].astype(np.float64).T @ dZ1q
    return loss, pack(g)


def init_params(seed):
    """He-normal weights, zero biases (the device initialiser draws its own)."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(size=(N_IN, N_HID)) * np.sqrt(2.0 / N_IN),
         "b1": np.zeros(N_HID),
         "W2": rng.normal(size=(N_HID, N_HID)) * np.sqrt(2.0 / N_HID),
         "b2": np.zeros(N_HID),
         "W3": rng.normal(size=(N_HID, N_HID)) * np.sqrt(2.0 / N_HID),
         "b3": np.zeros(N_HID),
         "W4": rng.normal(size=N_HID) * np.sqrt(1.0 / N_HID),
         "b4": np.zeros(1)}
    return pack(d)


def adam(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) in FP64; returns new params, m, v."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
