"""CPU reference of the learned surrogate's MLP -- TEST INFRASTRUCTURE ONLY.

The surrogate has no counterpart in the reference; this module is the textbook
forward pass and back-propagation of the network that the library trains, in
numpy, used only by ``tests/`` to check the GPU kernels' outputs and gradients:

    x = f / f_scale,  z_l = h_l W_l^T + b_l,  h_{l+1} = GELU(z_l),  y = z_{L-1},
    loss = mean over the nrows x N entries of (y - q / q_scale)^2.

With ``bf16=True`` the values are rounded to BF16 (round to nearest even) at the
same points as on the GPU -- the input, the weights, every activation and every
back-propagated dz that feeds a matrix product -- while sums run in float64, so
the GPU result differs only by its FP32 accumulation.  With ``bf16=False``
everything is float64; ``tests/test_surrogate_ref.py`` pins that mode against
central finite differences of the loss.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

_SQRT1_2 = 0.70710678118654752
_INV_SQRT_2PI = 0.39894228040143268


def bf16(x) -> np.ndarray:
    """float32 -> nearest BF16 (ties to even), returned as float32 (finite inputs)."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * _SQRT1_2))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z * _SQRT1_2)) + z * _INV_SQRT_2PI * np.exp(-0.5 * z * z)


def forward_backward(Ws, bs, f_rows, q_rows, f_scale=1.0, q_scale=1.0, bf16_path=True):
    """Returns (loss, dWs, dbs, y) for weights Ws[l] of shape [out, in]."""
    L = len(Ws)
    rnd = (lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)) if bf16_path else \
          (lambda a: np.asarray(a, np.float64))
    f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if bf16_path else \
          (lambda a: np.asarray(a, np.float64))
    nrows, N = f_rows.shape
    Wr = [rnd(W) for W in Ws]
    h = [rnd(f32(np.asarray(f_rows, np.float64) * (1.0 / f_scale)))]
    z = []
    for l in range(L):
        zl = h[l] @ Wr[l].T + np.asarray(bs[l], np.float64)
        z.append(zl)
        if l + 1 < L:
            h.append(rnd(gelu(zl)))
    y = z[-1]
    t = f32(np.asarray(q_rows, np.float64) * (1.0 / q_scale))
    d = y - t
    loss = float(np.sum(d * d) / (nrows * N))
    g = (2.0 / (nrows * N)) * d
    dWs, dbs = [None] * L, [None] * L
    dz = g
    for l in range(L - 1, -1, -1):
        dbs[l] = dz.sum(axis=0)
        dzr = rnd(dz)
        dWs[l] = dzr.T @ h[l]
        if l > 0:
            dh = dzr @ Wr[l]
            dz = dh * gelu_d(z[l - 1])
    return loss, dWs, dbs, y


def adam(W, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    W = W - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return W, m, v


def project_invariants(q, u, v, w, gw):
    """Orthogonal projection of each row of q onto the complement of
    psi = 1, u, v, w, |u|^2 in the gw-weighted inner product (float64)."""
    psi = np.stack([np.ones_like(u), u, v, w, u * u + v * v + w * w])
    G = (psi * gw) @ psi.T
    m = (q * gw) @ psi.T
    lam = np.linalg.solve(G, m.T).T
    return q - lam @ psi
