"""CPU restatement of the rank MLP (test infrastructure only).

y(u) = b2 + sum_h w2[h] * relu(w1[h] * u + b1[h]); loss = mean((y - t)^2).
Forward, analytic gradient (float64) and a central finite-difference gradient
that pins the analytic one.  Parameter order: w1[32], b1[32], w2[32], b2.
"""

from __future__ import annotations

import numpy as np

H = 32


def loss(params, u, t) -> float:
    p = np.asarray(params, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)[:, None]
    z = p[:H] * u + p[H:2 * H]
    y = p[3 * H] + (p[2 * H:3 * H] * np.maximum(z, 0.0)).sum(axis=1)
    r = y - np.asarray(t, dtype=np.float64)
    return float(np.mean(r * r))


def grad(params, u, t) -> np.ndarray:
    p = np.asarray(params, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    z = p[:H] * u[:, None] + p[H:2 * H]
    a = np.maximum(z, 0.0)
    y = p[3 * H] + (p[2 * H:3 * H] * a).sum(axis=1)
    dy = 2.0 * (y - np.asarray(t, dtype=np.float64)) / len(u)
    g = np.empty(3 * H + 1)
    g[3 * H] = dy.sum()
    g[2 * H:3 * H] = (dy[:, None] * a).sum(axis=0)
    dz = dy[:, None] * p[2 * H:3 * H] * (z > 0)
    g[:H] = (dz * u[:, None]).sum(axis=0)
    g[H:2 * H] = dz.sum(axis=0)
    return g


def grad_fd(params, u, t, h: float = 1e-6) -> np.ndarray:
    p = np.asarray(params, dtype=np.float64)
    g = np.empty_like(p)
    for k in range(len(p)):
        e = np.zeros_like(p)
        e[k] = h * max(1.0, abs(p[k]))
        g[k] = (loss(p + e, u, t) - loss(p - e, u, t)) / (2 * e[k])
    return g
