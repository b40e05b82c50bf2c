"""CPU restatement of the warm-start perceptron (test infrastructure: the
checker of the GPU network kernels, never part of the product path).

Network: x -> tanh(x W1 + b1) -> tanh(. W2 + b2) -> tanh(. W3 + b3) -> . W4 + b4,
parameters in one fp32 vector [W1 (n_in x h1, row-major), b1, ..., W4, b4];
loss = mean of (y - t)^2 over all outputs of the batch.

`terms` follows the device's product precision: 1 rounds every
matrix-product operand to bfloat16 (round to nearest, ties to even); 3
splits each operand a into hi = bf16(a) and lo = bf16(a - hi) and forms
hi*hi + hi*lo + lo*hi; products and sums run in float64 here, and the
activations and their gradients are stored as float32 as on the device.
With terms=None the arithmetic is plain float64, and the gradients are
pinned against central finite differences of the loss
(tests/test_mlp_oracle.py).
"""
import numpy as np


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def layout(dims):
    """[(W offset, W shape, b offset, b size)] per layer and the total count."""
    out = []
    off = 0
    for l in range(4):
        wo, ws = off, (dims[l], dims[l + 1])
        off += dims[l] * dims[l + 1]
        bo = off
        off += dims[l + 1]
        out.append((wo, ws, bo, dims[l + 1]))
    return out, off


def _unpack(p, dims):
    lay, n = layout(dims)
    assert p.size == n
    return [(p[wo:wo + ws[0] * ws[1]].reshape(ws).astype(np.float64), p[bo:bo + bn].astype(np.float64))
            for wo, ws, bo, bn in lay]


def matmul(a, b, terms):
    if terms is None:
        return np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64)
    a32 = np.asarray(a, dtype=np.float32)
    b32 = np.asarray(b, dtype=np.float32)
    ah, bh = bf16_round(a32), bf16_round(b32)
    out = ah.astype(np.float64) @ bh.astype(np.float64)
    if terms == 3:
        al = bf16_round(a32 - ah).astype(np.float64)
        bl = bf16_round(b32 - bh).astype(np.float64)
        out = out + ah.astype(np.float64) @ bl + al @ bh.astype(np.float64)
    return out


def forward(p, dims, x, terms=1):
    """Activations [x0, x1, x2, x3, y] (float64 values; stored as float32
    unless terms is None)."""
    layers = _unpack(np.asarray(p, dtype=np.float32), dims)
    acts = [np.asarray(x, dtype=np.float64)]
    for l, (W, b) in enumerate(layers):
        z = matmul(acts[-1], W, terms) + b
        a = np.tanh(z) if l < 3 else z
        if terms is not None:
            a = a.astype(np.float32).astype(np.float64)
        acts.append(a)
    return acts


def loss_and_gradients(p, dims, x, t, terms=1):
    p = np.asarray(p, dtype=np.float32)
    layers = _unpack(p, dims)
    acts = forward(p, dims, x, terms)
    y = acts[4]
    t = np.asarray(t, dtype=np.float64)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    lay, total = layout(dims)
    g = np.zeros(total)
    dz = 2.0 * (y - t) / n
    for l in range(3, -1, -1):
        W, _ = layers[l]
        if l < 3:
            dz = dz * (1.0 - acts[l + 1] ** 2)
        if terms is not None:
            dz = dz.astype(np.float32).astype(np.float64)
        wo, ws, bo, bn = lay[l]
        g[wo:wo + ws[0] * ws[1]] = matmul(acts[l].T, dz, terms).ravel()
        g[bo:bo + bn] = dz.sum(axis=0)
        if l > 0:
            dz = matmul(dz, W.T, terms)
    return loss, g
