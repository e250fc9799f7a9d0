"""NumPy restatement of the engine's mesh message-passing network (forward,
loss and backward).  TEST INFRASTRUCTURE ONLY: the checker for the HIP
training and inference kernels; the engine never imports it.

Network on the triangle graph (neighbours = TriC, mean aggregation):
    h0 = relu(X We + be)                               X: [N, 16], We: [16, 64]
    h_l = h_{l-1} + relu([h_{l-1} | mean_nb(h_{l-1})] W_l + b_l),  l = 1..3, W_l: [128, 64]
    y  = h_3 Wd + bd                                   Wd: [64, 16] (first 3 outputs used)
    loss = sum_i ||y_i[:3] - t_i||^2 / N
With bf16=True the GEMM operands are rounded to bfloat16 (round to nearest
even) and accumulated in float32, as the MFMA kernels do.

The backward pass is pinned in the tests by central finite differences of the
float64 forward pass.
"""

from __future__ import annotations

import numpy as np

F_IN, HID, F_OUT, LAYERS, N_T = 16, 64, 16, 3, 3


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {"We": rng.normal(0, np.sqrt(2.0 / F_IN), (F_IN, HID)), "be": np.zeros(HID)}
    for l in range(LAYERS):
        p[f"W{l}"] = rng.normal(0, 0.5 * np.sqrt(2.0 / (2 * HID)), (2 * HID, HID))
        p[f"b{l}"] = np.zeros(HID)
    p["Wd"] = rng.normal(0, np.sqrt(1.0 / HID), (HID, F_OUT))
    p["Wd"][:, N_T:] = 0.0
    p["bd"] = np.zeros(F_OUT)
    return {k: v.astype(np.float32) for k, v in p.items()}


def neighbour_mean(h, nb):
    """mean over valid neighbours (nb: [N, 3], -1 = none)."""
    valid = nb >= 0
    deg = np.maximum(valid.sum(1), 1)
    acc = np.zeros_like(h)
    for c in range(nb.shape[1]):
        m = valid[:, c]
        acc[m] += h[nb[m, c]]
    return acc / deg[:, None]


def neighbour_mean_T(g, nb):
    """transpose of neighbour_mean: out[j] = sum_{i: j in nb(i)} g[i] / deg_i."""
    valid = nb >= 0
    deg = np.maximum(valid.sum(1), 1)
    out = np.zeros_like(g)
    for c in range(nb.shape[1]):
        m = valid[:, c]
        np.add.at(out, nb[m, c], g[m] / deg[m, None])
    return out


def _mm(a, b, bf):
    if bf:
        return bf16(a).astype(np.float32) @ bf16(b).astype(np.float32)
    return a @ b


def forward(p, X, nb, bf=False):
    cache = {}
    z = _mm(X, p["We"], bf) + p["be"]
    h = np.maximum(z, 0)
    cache["z_e"] = z
    for l in range(LAYERS):
        a = neighbour_mean(h, nb)
        cat = np.concatenate([h, a], axis=1)
        z = _mm(cat, p[f"W{l}"], bf) + p[f"b{l}"]
        cache[f"cat{l}"], cache[f"z{l}"] = cat, z
        h = h + np.maximum(z, 0)
    cache["h"] = h
    y = _mm(h, p["Wd"], bf) + p["bd"]
    return y, cache


def loss_and_grad(p, X, nb, T, bf=False):
    y, c = forward(p, X, nb, bf)
    N = X.shape[0]
    d = y[:, :N_T] - T
    loss = float((d * d).sum() / N)
    dy = np.zeros_like(y)
    dy[:, :N_T] = 2.0 * d / N
    g = {"Wd": _mm(c["h"].T, dy, bf), "bd": dy.sum(0)}
    dh = _mm(dy, p["Wd"].T, bf)
    for l in reversed(range(LAYERS)):
        dz = dh * (c[f"z{l}"] > 0)
        g[f"W{l}"] = _mm(c[f"cat{l}"].T, dz, bf)
        g[f"b{l}"] = dz.sum(0)
        dcat = _mm(dz, p[f"W{l}"].T, bf)
        dh = dh + dcat[:, :HID] + neighbour_mean_T(dcat[:, HID:], nb)
    dz = dh * (c["z_e"] > 0)
    g["We"] = _mm(X.T, dz, bf)
    g["be"] = dz.sum(0)
    return loss, g, y
