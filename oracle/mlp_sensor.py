"""NumPy restatement of the learned hybrid sensor (features, MLP forward and
backward, the Adam update and the flag rule).

TEST INFRASTRUCTURE ONLY: the checker for the HIP sensor kernels (feature
extraction, BF16-MFMA inference, FP32 training kernels).  Nothing in the
product imports it.

The plain sensor it stands beside is the Jameson style-2 selector of
hybrid_auto (Set_Scheme_HybridAuto / Comput_Scheme_point_Jameson,
src/OCFD_Schemes_hybrid_auto.cu:21-68, 611-703): kp = 1 + (dp > i1) +
(dp > i2) with dp the sum over the three directions of
|p- - 2p + p+| / (p- + 2p + p+); kp = 1 selects UP7, 2 WENO7, 3 WENO5.

Features (24 = 8 per direction d in x, y, z, at node n, FP32), from
p = p00*rho*T, rho and the contravariant velocity U = (A . u)/|A| of the
direction (A = the direction's cofactor row) at the nodes n-3..n+4 of the
8-point stencil of face n+1/2:

  0..2  (p[n+m+1] - p[n+m]) / (p[n+m+1] + p[n+m]),  m = -1, 0, 1
  3..5  (rho[n+m+1] - rho[n+m]) / (rho[n+m+1] + rho[n+m]),  m = -1, 0, 1
  6     max over the outer intervals m = -3, -2, 2, 3 of |p jump|
  7     (U[n+1] - U[n]) / c[n],  c = sqrt(T)/Ma

A jump whose denominator is 0 (ghost layers that are never written hold 0)
is 0, and so is feature 7 where c = 0.

Network: 24 -> 32 -> 32 -> 1, GELU (tanh form) on the hidden layers,
sigmoid output = probability that the point needs WENO (plain kp > 1).
Flag rule: prob < lo -> 1 (UP7); prob > hi -> max(kp, 2); otherwise the
plain kp (the low-confidence band falls back to the plain sensor).
"""
from __future__ import annotations

import numpy as np

NF, NH = 24, 32
GELU_C = 0.7978845608028654  # sqrt(2/pi)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3 * 0.044715 * x * x)


def init_params(seed=1, dtype=np.float32):
    """He-style seeded initial weights: W1 (32,24), b1 (32), W2 (32,32),
    b2 (32), W3 (32), b3 ()."""
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.standard_normal((NH, NF)) * np.sqrt(2.0 / NF),
        "b1": np.zeros(NH),
        "W2": rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH),
        "b2": np.zeros(NH),
        "W3": rng.standard_normal(NH) * np.sqrt(1.0 / NH),
        "b3": np.zeros(1),
    }
    return {k: v.astype(dtype) for k, v in p.items()}


def forward(p, X):
    """X (n,24) -> (prob (n,), cache)."""
    z1 = X @ p["W1"].T + p["b1"]
    a1 = gelu(z1)
    z2 = a1 @ p["W2"].T + p["b2"]
    a2 = gelu(z2)
    z3 = a2 @ p["W3"] + p["b3"][0]
    prob = 1.0 / (1.0 + np.exp(-z3))
    return prob, (X, z1, a1, z2, a2, z3)


def loss_and_grads(p, X, y, wpos=1.0):
    """Mean weighted binary cross-entropy and its gradients (y in {0,1},
    positives weighted by wpos)."""
    prob, (X, z1, a1, z2, a2, z3) = forward(p, X)
    n = X.shape[0]
    w = np.where(y > 0.5, wpos, 1.0)
    eps = 1e-7
    loss = -np.mean(w * (y * np.log(prob + eps) + (1 - y) * np.log(1 - prob + eps)))
    dz3 = w * (prob - y) / n
    g = {"W3": a2.T @ dz3, "b3": np.array([dz3.sum()])}
    da2 = np.outer(dz3, p["W3"])
    dz2 = da2 * gelu_grad(z2)
    g["W2"] = dz2.T @ a1
    g["b2"] = dz2.sum(0)
    da1 = dz2 @ p["W2"]
    dz1 = da1 * gelu_grad(z1)
    g["W1"] = dz1.T @ X
    g["b1"] = dz1.sum(0)
    return loss, g


def adam_step(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update in place (t = 1-based step)."""
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)


def jameson_kp(P, i1, i2):
    """Plain flags on the interior of a padded pressure field P (nz+2, ny+2,
    nx+2) -> (nz, ny, nx) int8."""
    c = P[1:-1, 1:-1, 1:-1]
    dp = 0.0
    for ax in (2, 1, 0):  # x, y, z: the summation order of the sensor
        sl_m = [slice(1, -1)] * 3
        sl_p = [slice(1, -1)] * 3
        sl_m[ax] = slice(0, -2)
        sl_p[ax] = slice(2, None)
        pm, pp = P[tuple(sl_m)], P[tuple(sl_p)]
        dp = dp + np.abs(-pm + 2 * c - pp) / (pm + 2 * c + pp)
    return (1 + (dp > i1).astype(np.int8) + (dp > i2).astype(np.int8)).astype(np.int8)


def features_line(p, rho, U, c, n):
    """The 8 features of one direction at node n of 1-D arrays indexed by
    node + 4 (n-3..n+4 must exist)."""
    q = n + 4

    def jump(a, m):
        s = a[q + m + 1] + a[q + m]
        return (a[q + m + 1] - a[q + m]) / s if s != 0 else 0.0

    f = [jump(p, -1), jump(p, 0), jump(p, 1), jump(rho, -1), jump(rho, 0), jump(rho, 1)]
    f.append(max(abs(jump(p, m)) for m in (-3, -2, 2, 3)))
    f.append((U[q + 1] - U[q]) / c[q] if c[q] != 0 else 0.0)
    return np.array(f)


def features_field(d, u, v, w, T, A, p00, rama, nodes):
    """Features (len(nodes), 24) from LAP-box fields (nz+8, ny+8, nx+8) of the
    primitives and the metric cofactors A (9, ...) in the order Akx Aky Akz
    Aix Aiy Aiz Asx Asy Asz; nodes = (i, j, k) interior indices (>= -1)."""
    P = p00 * d * T
    c = np.sqrt(np.maximum(T, 0.0)) * rama
    out = np.zeros((len(nodes), 24))
    for r, (i, j, k) in enumerate(nodes):
        for dd in range(3):
            a1, a2, a3 = A[3 * dd], A[3 * dd + 1], A[3 * dd + 2]
            idx = []
            for o in range(-4, 5):
                ii, jj, kk = i + 4 + (o if dd == 0 else 0), j + 4 + (o if dd == 1 else 0), k + 4 + (o if dd == 2 else 0)
                idx.append((kk, jj, ii))
            pl = np.array([P[t] for t in idx] + [0.0] * 3)
            rl = np.array([d[t] for t in idx] + [0.0] * 3)
            cl = np.array([c[t] for t in idx] + [0.0] * 3)
            Ul = np.zeros(12)
            for m in (4, 5):
                t = idx[m]
                nrm = np.sqrt(a1[t] ** 2 + a2[t] ** 2 + a3[t] ** 2)
                Ul[m] = (a1[t] * u[t] + a2[t] * v[t] + a3[t] * w[t]) / nrm
            # line arrays indexed by node + 4 with node 0 = this node
            out[r, 8 * dd:8 * dd + 8] = features_line(pl, rl, Ul, cl, 0)
    return out


def flag_rule(prob, kp, lo=0.2, hi=0.8):
    out = kp.copy()
    out = np.where(prob < lo, 1, out)
    out = np.where(prob > hi, np.maximum(kp, 2), out)
    return out.astype(np.int8)
