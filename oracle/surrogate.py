"""CPU oracle of the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

The surrogate has no counterpart in the reference (it is new in this build):
an MLP D -> 256 -> 256 -> 1 with SiLU, inputs scaled by 1/100, trained with
Adam on a ring archive of true evaluations y = log(1 + f), used to rank
candidate trials.  This module restates it in float64 numpy:

* forward / analytic backward, optionally with every GEMM operand rounded to
  bfloat16 (round-to-nearest-even) exactly where the HIP kernels round, so the
  GPU's BF16-operand / FP32-accumulate results can be compared tightly;
* Adam (Kingma & Ba, 2015) with bias correction;
* Spearman rank correlation (average ranks for ties).

Pinning (tests/test_surrogate_oracle.py): the analytic gradient against
central finite differences in float64, Adam's first step against its closed
form (step = -lr * sign(g) for eps -> 0), Spearman against scipy.stats.
"""
from __future__ import annotations

import numpy as np

HIDDEN = 256
X_SCALE = 0.01


def bf16(a):
    """Round float64/float32 values to bfloat16 (RNE), returned as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a32)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def sigmoid(a):
    return 1.0 / (1.0 + np.exp(-a))


def silu(a):
    return a * sigmoid(a)


def dsilu(a):
    s = sigmoid(a)
    return s * (1.0 + a * (1.0 - s))


def init_params(D, seed=0):
    """He-style uniform init (the GPU path takes its weights from the caller
    or from its own Philox init; tests pass the same arrays to both)."""
    rng = np.random.default_rng(seed)

    def u(fan_in, shape):
        lim = np.sqrt(6.0 / fan_in)
        return rng.uniform(-lim, lim, shape)
    return {
        "W1": u(D, (HIDDEN, D)), "b1": np.zeros(HIDDEN),
        "W2": u(HIDDEN, (HIDDEN, HIDDEN)), "b2": np.zeros(HIDDEN),
        "w3": u(HIDDEN, (HIDDEN,)) * 0.1, "b3": np.zeros(1),
    }


def forward(p, X, round_bf16=False):
    """Returns prediction y [B] and the cache for backward.  X: [B, D] raw."""
    r = bf16 if round_bf16 else (lambda v: np.asarray(v, np.float64))
    Xs = np.asarray(X, np.float64) * X_SCALE
    a1 = r(Xs) @ r(p["W1"]).T + p["b1"]
    h1 = silu(a1)
    a2 = r(h1) @ r(p["W2"]).T + p["b2"]
    h2 = silu(a2)
    y = h2 @ p["w3"] + p["b3"][0]
    return y, {"Xs": Xs, "a1": a1, "h1": h1, "a2": a2, "h2": h2}


def loss_and_grads(p, X, t, round_bf16=False):
    """L = mean((y - t)^2) / 2 and dL/dparams (analytic)."""
    r = bf16 if round_bf16 else (lambda v: np.asarray(v, np.float64))
    y, c = forward(p, X, round_bf16)
    B = y.shape[0]
    dy = (y - t) / B
    L = 0.5 * np.mean((y - t) ** 2)
    g = {}
    g["w3"] = c["h2"].T @ dy
    g["b3"] = np.array([dy.sum()])
    dh2 = np.outer(dy, p["w3"]) * dsilu(c["a2"])
    g["W2"] = r(dh2).T @ r(c["h1"])
    g["b2"] = dh2.sum(0)
    dh1 = (r(dh2) @ r(p["W2"])) * dsilu(c["a1"])
    g["W1"] = r(dh1).T @ r(c["Xs"])
    g["b1"] = dh1.sum(0)
    return L, g


def adam_step(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (t counts from 1); returns new (p, m, v)."""
    out_p, out_m, out_v = {}, {}, {}
    for k in p:
        mk = b1 * m[k] + (1 - b1) * g[k]
        vk = b2 * v[k] + (1 - b2) * g[k] ** 2
        mh = mk / (1 - b1 ** t)
        vh = vk / (1 - b2 ** t)
        out_p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
        out_m[k], out_v[k] = mk, vk
    return out_p, out_m, out_v


def rank_average(a):
    a = np.asarray(a, np.float64)
    order = np.argsort(a, kind="mergesort")
    ranks = np.empty(a.size, np.float64)
    sa = a[order]
    i = 0
    while i < a.size:
        j = i
        while j + 1 < a.size and sa[j + 1] == sa[i]:
            j += 1
        ranks[order[i:j + 1]] = 0.5 * (i + j)
        i = j + 1
    return ranks


def spearman(a, b):
    ra, rb = rank_average(a), rank_average(b)
    ra -= ra.mean()
    rb -= rb.mean()
    den = np.sqrt((ra * ra).sum() * (rb * rb).sum())
    return float((ra * rb).sum() / den) if den > 0 else 0.0
