"""CPU oracle of the spike-proposal network -- TEST INFRASTRUCTURE ONLY.

The proposal network has no counterpart in the reference; it is the learned
gate in front of the matcher.  This module restates it in float64 numpy, with
the forward pass, the loss and a hand-derived backward pass, so the GPU
kernels can be checked against it (forward outputs and gradients).  The
backward pass here is itself pinned by central finite differences in
``tests/test_proposal_oracle.py``.

Network (per 16-channel group g of the probe, weights shared across groups):
  input  x[g]  = whitened channels 16g-4 .. 16g+19 (24 rows, zero off-probe)
  layer  l=1..3: a_l = relu(conv_l(a_{l-1}) + b_l), 32 channels, kernel 9, 'same'
  head   z = conv_4(a_3) + b_4 (1 channel), p = sigmoid(z)
conv(x)[co][w] = sum_{ci,k} W[co][ci][k] * x[ci][w + k - 4] (zero outside).
Loss: weighted binary cross-entropy with logits, mean over all (g, w):
  L = mean( pw * y * softplus(-z) + (1 - y) * softplus(z) ).
"""
from __future__ import annotations

import numpy as np

GROUP = 16
HALO = 4
CIN = GROUP + 2 * HALO
CH = 32
KW = 9
PAD = KW // 2

# parameter blocks in their packed order: (name, shape)
LAYOUT = (("W1", (CH, CIN, KW)), ("b1", (CH,)), ("W2", (CH, CH, KW)), ("b2", (CH,)),
          ("W3", (CH, CH, KW)), ("b3", (CH,)), ("W4", (1, CH, KW)), ("b4", (1,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(flat: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in LAYOUT])


def group_input(xw: np.ndarray) -> np.ndarray:
    """[C][W] whitened data -> [G][24][W] group inputs (zero channels off the probe)."""
    C, W = xw.shape
    G = (C + GROUP - 1) // GROUP
    out = np.zeros((G, CIN, W))
    for g in range(G):
        for ci in range(CIN):
            ch = g * GROUP - HALO + ci
            if 0 <= ch < C:
                out[g, ci] = xw[ch]
    return out


def conv(x: np.ndarray, Wt: np.ndarray, b: np.ndarray) -> np.ndarray:
    """x [G][CI][W], Wt [CO][CI][9] -> [G][CO][W] ('same', zero padded)."""
    G, CI, W = x.shape
    xp = np.zeros((G, CI, W + 2 * PAD))
    xp[:, :, PAD:PAD + W] = x
    out = np.zeros((G, Wt.shape[0], W))
    for k in range(KW):
        out += np.einsum("oc,gcw->gow", Wt[:, :, k], xp[:, :, k:k + W])
    return out + b[None, :, None]


def conv_back(dy: np.ndarray, x: np.ndarray, Wt: np.ndarray):
    """Gradients of conv: (dx, dW, db) for dy [G][CO][W]."""
    G, CI, W = x.shape
    xp = np.zeros((G, CI, W + 2 * PAD))
    xp[:, :, PAD:PAD + W] = x
    dW = np.zeros_like(Wt)
    for k in range(KW):
        dW[:, :, k] = np.einsum("gow,gcw->oc", dy, xp[:, :, k:k + W])
    db = dy.sum(axis=(0, 2))
    dxp = np.zeros_like(xp)
    for k in range(KW):
        dxp[:, :, k:k + W] += np.einsum("oc,gow->gcw", Wt[:, :, k], dy)
    return dxp[:, :, PAD:PAD + W], dW, db


def softplus(z):
    return np.logaddexp(0.0, z)


def bf16(a) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params: dict, x0: np.ndarray, bf16_storage: bool = False):
    """Returns (logits [G][W], activations [a0, a1, a2, a3]).  With
    ``bf16_storage`` the input, the weights of layers 1-3 and the activations
    are rounded to bfloat16 where the GPU stores them (accumulation stays in
    high precision; the head's weights stay FP32, as on the GPU)."""
    r = bf16 if bf16_storage else (lambda v: v)
    x0 = r(x0)
    acts = [x0]
    a = x0
    for l in (1, 2, 3):
        a = r(np.maximum(conv(a, r(params[f"W{l}"]), params[f"b{l}"]), 0.0))
        acts.append(a)
    z = conv(a, params["W4"], params["b4"])[:, 0, :]
    return z, acts


def loss_and_grad(params: dict, x0: np.ndarray, y: np.ndarray, pos_weight: float, bf16_storage: bool = False):
    """Weighted BCE loss and its gradient (dict of the same shapes as params)."""
    if bf16_storage:
        params = dict(params)
        for l in (1, 2, 3):
            params[f"W{l}"] = bf16(params[f"W{l}"])
    z, acts = forward(params, x0, bf16_storage)
    n = z.size
    loss = float(np.mean(pos_weight * y * softplus(-z) + (1.0 - y) * softplus(z)))
    s = 1.0 / (1.0 + np.exp(-z))
    wgt = np.where(y > 0.5, pos_weight, 1.0)
    dz = wgt * (s - y) / n        # d/dz of w*(y*sp(-z)) + (1-y)*sp(z) for y in {0, 1}
    grads = {}
    dy = dz[:, None, :]
    for l in (4, 3, 2, 1):
        dx, dW, db = conv_back(dy, acts[l - 1], params[f"W{l}"])
        grads[f"W{l}"], grads[f"b{l}"] = dW, db
        if l > 1:
            dy = dx * (acts[l - 1] > 0)
    return loss, grads


def labels(centres_cols: np.ndarray, groups: np.ndarray, G: int, W: int, tol: int = 1) -> np.ndarray:
    """y[g][w] = 1 within +-tol columns of a spike whose main channel is in group g."""
    y = np.zeros((G, W))
    for c, g in zip(centres_cols, groups):
        y[g, max(0, c - tol):min(W, c + tol + 1)] = 1.0
    return y
