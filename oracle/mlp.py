"""CPU restatement of the GJK/EPA warm-start network (rbps_amd/csrc/rbgpu_mlp.hpp):
forward, loss and backward in float64 with BF16 rounding at the points where the GPU
rounds (network input, hidden activations, output gradient, hidden-layer gradients).

TEST INFRASTRUCTURE ONLY: the checker of the GPU kernels' gradients and outputs, never
the product.  The network has no counterpart in the reference, so this file is pinned
by finite differences of its own float64 path (tests/test_oracle_mlp.py), not by
reference vectors.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 32, 128, 16
W1, B1 = 0, H * IN
W2, B2 = B1 + H, B1 + H + H * H
W3, B3 = B2 + H, B2 + H + H * H
W4, B4 = B3 + H, B3 + H + OUT * H
NPARAM = B4 + OUT
LAYERS = [(W1, B1, IN, H), (W2, B2, H, H), (W3, B3, H, H), (W4, B4, H, OUT)]


def bf16(x):
    """Round to bfloat16 (round to nearest even), returned as float32."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32)
    b = ((b + np.uint32(0x7FFF) + ((b >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)).astype(np.uint32)
    return b.view(np.float32)


def _weights(p, rounded):
    out = []
    for wo, bo, k, n in LAYERS:
        w = p[wo:wo + n * k].reshape(n, k).astype(np.float64)
        out.append((bf16(w).astype(np.float64) if rounded else w, p[bo:bo + n].astype(np.float64)))
    return out


def forward(p, feat, rounded=True):
    """Activations a0..a3 and the output (n x 16)."""
    r = (lambda x: bf16(x).astype(np.float64)) if rounded else (lambda x: x)
    acts = [r(np.asarray(feat, np.float64))]
    ws = _weights(p, rounded)
    for li, (w, b) in enumerate(ws):
        z = acts[-1] @ w.T + b
        if li < 3:
            acts.append(r(np.maximum(z, 0.0)))
        else:
            out = z
    return acts, out


def loss_and_grad(p, feat, tgt, rounded=True):
    """Mean loss over the samples and its gradient w.r.t. the flat parameters.
    tgt: n x 8 = (t_gjk xyz, valid, t_epa xyz, valid)."""
    r = (lambda x: bf16(x).astype(np.float64)) if rounded else (lambda x: x)
    n = len(feat)
    acts, out = forward(p, feat, rounded)
    scale = 1.0 / n
    dout = np.zeros_like(out)
    loss = 0.0
    for h in range(2):
        pr = out[:, 3 * h:3 * h + 3]
        t = np.asarray(tgt, np.float64)[:, 4 * h:4 * h + 3]
        valid = np.asarray(tgt)[:, 4 * h + 3] != 0
        nrm = np.linalg.norm(pr, axis=1)
        ok = valid & (nrm * nrm >= 1e-24)
        c = np.where(ok, np.einsum("ij,ij->i", pr, t) / np.where(ok, nrm, 1.0), 0.0)
        loss += scale * np.sum(np.where(valid, 1.0 - c, 0.0))
        g = (pr * (c / np.where(ok, nrm, 1.0))[:, None] - t) / np.where(ok, nrm, 1.0)[:, None]
        dout[:, 3 * h:3 * h + 3] = np.where(ok[:, None], g * scale, 0.0)
    grad = np.zeros(NPARAM)
    ws = _weights(p, rounded)
    d = r(dout)
    for li in range(3, -1, -1):
        wo, bo, k, nn = LAYERS[li]
        a = acts[li]
        grad[wo:wo + nn * k] = (d.T @ a).ravel()
        grad[bo:bo + nn] = d.sum(0)
        if li > 0:
            da = d @ ws[li][0]
            d = r(np.where(a > 0, da, 0.0))
    return loss, grad


def predict(p, feat, rounded=True):
    """Normalised (GJK direction, EPA seed normal) per sample (zero when degenerate)."""
    _, out = forward(p, feat, rounded)
    res = np.zeros((len(feat), 6))
    for h in range(2):
        v = out[:, 3 * h:3 * h + 3]
        nv = np.linalg.norm(v, axis=1)
        res[:, 3 * h:3 * h + 3] = np.where(nv[:, None] > 1e-12, v / np.maximum(nv, 1e-30)[:, None], 0.0)
    return res


def init_params(seed=1):
    """He-uniform weights and zero biases (a fresh network for the tests)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for wo, bo, k, n in LAYERS:
        lim = np.sqrt(6.0 / k)
        p[wo:wo + n * k] = rng.uniform(-lim, lim, n * k)
        p[bo:bo + n] = rng.uniform(-0.05, 0.05, n)
    return p
