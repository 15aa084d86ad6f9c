"""FP64 CPU multi-layer perceptron: the oracle of the chi_B surrogate's device kernels.

Test infrastructure only (tests/ import it as the checker; the library never does). The
surrogate has no counterpart in the reference; what is checked against this restatement is the
arithmetic of qa_mlp_forward / qa_mlp_loss_grad: an ensemble member is
    a_0 = (x - mean) / scale
    a_{l+1} = tanh(W_l a_l + b_l)          l < depth
    y = sigmoid(W_depth a_depth + b_depth)
with loss = mean over batch and outputs of (y - t)^2, and its exact gradient by backpropagation.
Parameters of one member are flat: for each layer W_l [out][in] (row-major) then b_l [out].
This module's gradient is itself pinned by central finite differences in FP64
(tests/test_oracle_mlp.py), which a dropped term, a transposed W or a wrong activation
derivative fails.
"""
from __future__ import annotations

import numpy as np


def layer_dims(n_in: int, width: int, depth: int, n_out: int):
    return [(n_in if l == 0 else width, n_out if l == depth else width) for l in range(depth + 1)]


def unpack(params, n_in: int, width: int, depth: int, n_out: int):
    """[(W_l, b_l)] views of one member's flat parameter vector."""
    p = np.asarray(params, dtype=np.float64).reshape(-1)
    out, off = [], 0
    for din, dout in layer_dims(n_in, width, depth, n_out):
        W = p[off:off + din * dout].reshape(dout, din)
        off += din * dout
        b = p[off:off + dout]
        off += dout
        out.append((W, b))
    if off != p.size:
        raise ValueError(f"parameter vector has {p.size} entries, the layout needs {off}")
    return out


def n_params(n_in: int, width: int, depth: int, n_out: int) -> int:
    return sum(din * dout + dout for din, dout in layer_dims(n_in, width, depth, n_out))


def forward(params, x, n_in, width, depth, n_out, mean=None, scale=None):
    """Activations [a_0, a_1, ..., y] of one member for inputs x [batch][n_in]."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, n_in)
    mean = np.zeros(n_in) if mean is None else np.asarray(mean, np.float64)
    scale = np.ones(n_in) if scale is None else np.asarray(scale, np.float64)
    acts = [(x - mean) / scale]
    layers = unpack(params, n_in, width, depth, n_out)
    for l, (W, b) in enumerate(layers):
        z = acts[-1] @ W.T + b
        acts.append(np.tanh(z) if l < depth else 1.0 / (1.0 + np.exp(-z)))
    return acts


def loss_grad(params, x, t, n_in, width, depth, n_out, mean=None, scale=None):
    """(loss, flat gradient) of one member: loss = mean (y - t)^2."""
    acts = forward(params, x, n_in, width, depth, n_out, mean, scale)
    y = acts[-1]
    t = np.asarray(t, dtype=np.float64).reshape(y.shape)
    count = y.size
    loss = float(np.sum((y - t) ** 2) / count)
    layers = unpack(params, n_in, width, depth, n_out)
    grads = [None] * len(layers)
    dA = 2.0 * (y - t) / count
    for l in range(depth, -1, -1):
        a_out = acts[l + 1]
        dZ = dA * (a_out * (1.0 - a_out) if l == depth else 1.0 - a_out * a_out)
        W, _ = layers[l]
        grads[l] = (dZ.T @ acts[l], dZ.sum(axis=0))
        dA = dZ @ W
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, flat
