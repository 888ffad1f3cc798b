"""CPU restatement (numpy, FP64 arithmetic) of the warm-start MLP of
loadsg_amd/csrc/lsg_mlp.hip, with BF16 rounding at exactly the points where the GPU
path rounds (inputs, weights and activations as MFMA operands, dY and dZ as GEMM
operands).  Test infrastructure only.  Its backward pass is pinned against central
finite differences of its own loss (with the BF16 rounding switched off, where the
loss is smooth) in tests/test_mlp.py.
"""
import math

import numpy as np


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unflatten(p, shapes):
    out, o = [], 0
    for (i, n) in shapes:
        W = p[o:o + i * n].reshape(i, n).astype(np.float64)
        o += i * n
        b = p[o:o + n].astype(np.float64)
        o += n
        out.append((W, b))
    return out


def forward(layers, x, round_bf16=True):
    rnd = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    a = rnd(x)
    zs, acts = [], [a]
    for li, (W, b) in enumerate(layers):
        z = a @ rnd(W) + b
        if li + 1 < len(layers):
            zs.append(z)
            a = rnd(gelu(z))
            acts.append(a)
        else:
            return z, zs, acts


def loss_and_grads(layers, x, t, round_bf16=True):
    rnd = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    y, zs, acts = forward(layers, x, round_bf16)
    n = y.size
    e = y - t
    loss = float(np.sum(e * e) / n)
    dz = 2.0 * e / n
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        dzb = rnd(dz)
        grads[li] = (acts[li].T @ dzb, dz.sum(axis=0))
        if li == 0:
            break
        da = dzb @ rnd(W).T
        dz = da * gelu_grad(zs[li - 1])
    return loss, grads


def flatten_grads(grads):
    return np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in grads])
