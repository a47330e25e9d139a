"""CPU reference of the warm-start MLP (TEST INFRASTRUCTURE: only tests/
use it, as the checker of the library's MLP kernels).

The network: x -> [W_l, b_l, GELU] ... -> W_L x + b_L (no activation on the
last layer), loss = mean squared error.  `bf16=True` rounds every GEMM
operand to bfloat16 (round to nearest even) exactly as the library's matrix-
core kernels do, and accumulates in FP64, so the GPU result differs from it
only by its FP32 accumulation.  Parameters are laid out layer by layer: W
(width[l] x width[l+1], row major) then b.
"""
import numpy as np
from scipy.special import erf


def bf16(a):
    """round to nearest even to bfloat16, returned as float64"""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(p, width):
    out, off = [], 0
    for l in range(len(width) - 1):
        n = width[l] * width[l + 1]
        W = p[off:off + n].reshape(width[l], width[l + 1])
        off += n
        b = p[off:off + width[l + 1]]
        off += width[l + 1]
        out.append((W, b))
    return out


def _mm(a, b, rnd):
    return (bf16(a) @ bf16(b)) if rnd else (np.asarray(a, np.float64) @ np.asarray(b, np.float64))


def forward(p, width, x, bf16_ops=True):
    layers = unpack(np.asarray(p, np.float64), width)
    acts, pres = [np.asarray(x, np.float64)], [None]
    for l, (W, b) in enumerate(layers):
        z = _mm(acts[-1], W, bf16_ops) + b
        pres.append(z)
        if l < len(layers) - 1:
            z = gelu(z)
            if bf16_ops:
                z = z.astype(np.float32).astype(np.float64)   # activations are stored in FP32
        acts.append(z)
    return acts, pres


def loss_and_grad(p, width, x, t, bf16_ops=True):
    acts, pres = forward(p, width, x, bf16_ops)
    layers = unpack(np.asarray(p, np.float64), width)
    y = acts[-1]
    n = y.size
    loss = np.sum((y - t) ** 2) / n
    dA = 2.0 * (y - t) / n
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dZ = dA if l == len(layers) - 1 else dA * gelu_grad(pres[l + 1])
        if bf16_ops:
            dZ = dZ.astype(np.float32).astype(np.float64)
        dW = _mm(acts[l].T, dZ, bf16_ops)
        db = dZ.sum(axis=0)
        grads.append((dW, db))
        if l > 0:
            dA = _mm(dZ, W.T, bf16_ops)
    g = []
    for dW, db in reversed(grads):
        g += [dW.reshape(-1), db]
    return loss, np.concatenate(g)
