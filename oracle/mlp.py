"""The learned preconditioner's network: forward, backward, Adam.  Test oracle only.

There is no counterpart in the reference (its preconditioners are ``dpr`` and
``olsen``, source/eigen/block_davidson.f90:1288-1419); this is the CPU
checker of the HIP training kernels.

Network (per vector element j of root i):
  features  g = H_jj - theta_i, theta_i, rhat_j = r_j/|r_i|, class c_j
            (0 = 1h1p, 1..5 = 2h2p type), the class entering through a
            learned embedding row E[c] added to the first layer;
  h1 = silu(W1 [g', theta', rhat'] + E[c] + b1)      (64)
  h2 = silu(W2 h1 + b2)                                (64)
  m  = w3 . h2 + b3                                    (scalar)
  correction t_j = d_j * (1 + m), d_j = r_j / (theta_i - H_jj) (the DPR step)
with the input transform g' = asinh(g), theta' = theta, rhat' = asinh(rhat).

Training pairs (per element): features, dn_j = d_j/rms(e), en_j = e_j/rms(e)
with e = x - x*/(x.x*) the exact correction towards the converged eigenvector
x* and rms(e) its root-mean-square element;
loss = mean_j (asinh(dn_j (1 + m_j)) - asinh(en_j))^2
(compared through asinh, so that the rare elements with a near-zero DPR
denominator, whose dn is orders of magnitude above the rest, do not dominate
the mean).

Weights are stored in bfloat16 and every product accumulates in fp32 on the
GPU; this oracle evaluates the same bf16-rounded weights in float64.

Pinned by: finite differences of ``loss`` (test_oracle_pins), which fix every
gradient term independently of the analytic backward.
"""
from __future__ import annotations

import numpy as np

H = 64
NCLASS = 6


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def init_params(seed: int = 0) -> dict:
    rng = np.random.Generator(np.random.PCG64(seed))
    p = {
        "W1": rng.normal(0, 1 / np.sqrt(3), size=(H, 3)),
        "E": rng.normal(0, 0.1, size=(NCLASS, H)),
        "b1": np.zeros(H),
        "W2": rng.normal(0, 1 / np.sqrt(H), size=(H, H)),
        "b2": np.zeros(H),
        "w3": np.zeros(H),
        "b3": np.zeros(1),
    }
    return {k: bf16_round(v).astype(np.float64) for k, v in p.items()}


def transform(gap, theta, rhat):
    return np.stack([np.arcsinh(gap), theta, np.arcsinh(rhat)], axis=-1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, cls):
    """x: (n,3) transformed features, cls: (n,) int. Returns m and a cache."""
    z1 = x @ p["W1"].T + p["E"][cls] + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = silu(z2)
    m = h2 @ p["w3"] + p["b3"][0]
    return m, (x, cls, z1, h1, z2, h2)


def loss(p, x, cls, dn, en):
    m, _ = forward(p, x, cls)
    r = np.arcsinh(dn * (1.0 + m)) - np.arcsinh(en)
    return np.mean(r * r)


def backward(p, x, cls, dn, en):
    """Gradients of ``loss`` w.r.t. every parameter."""
    m, (x, cls, z1, h1, z2, h2) = forward(p, x, cls)
    n = x.shape[0]
    t = dn * (1.0 + m)
    r = np.arcsinh(t) - np.arcsinh(en)
    gm = 2.0 * r * dn / np.sqrt(1.0 + t * t) / n
    g = {"w3": h2.T @ gm, "b3": np.array([gm.sum()])}
    gz2 = np.outer(gm, p["w3"]) * dsilu(z2)
    g["W2"] = gz2.T @ h1
    g["b2"] = gz2.sum(axis=0)
    gz1 = (gz2 @ p["W2"]) * dsilu(z1)
    g["W1"] = gz1.T @ x
    g["b1"] = gz1.sum(axis=0)
    gE = np.zeros_like(p["E"])
    np.add.at(gE, cls, gz1)
    g["E"] = gE
    return g


def adam_step(p, g, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update on fp32 master weights (state['master']); p is the
    bf16-rounded copy the network evaluates."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in p:
        mk = state.setdefault("m_" + k, np.zeros_like(p[k]))
        vk = state.setdefault("v_" + k, np.zeros_like(p[k]))
        wk = state.setdefault("w_" + k, p[k].copy())
        mk[...] = b1 * mk + (1 - b1) * g[k]
        vk[...] = b2 * vk + (1 - b2) * g[k] * g[k]
        mh = mk / (1 - b1 ** t)
        vh = vk / (1 - b2 ** t)
        wk[...] = wk - lr * mh / (np.sqrt(vh) + eps)
        p[k] = bf16_round(wk).astype(np.float64)
    return p
