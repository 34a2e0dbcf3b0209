"""Test infrastructure (oracle): NOT part of the shipped product.

CPU restatement of the surrogate network's forward pass, MSE loss and backward pass
(unicorn_amd/surrogate.py, unicorn_amd/csrc/mlp_kernels.hip), in float64 numpy.

With ``bf16=True`` it rounds where the GPU stores BF16: the inputs, the weights,
each hidden activation H and each back-propagated dZ.  Everything else (GEMM
accumulation, bias, GELU, the loss) is exact here, so the GPU's FP32 accumulation
differs only by rounding; the GPU's GELU evaluates erf by Abramowitz & Stegun 7.1.28
(within 2e-6 of erf in float arithmetic), far below the bf16 rounding.  With ``bf16=False`` it is the plain float64 network, whose
gradients tests/test_surrogate.py pins against central finite differences.
"""
from __future__ import annotations

import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(math.erf)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even) and back, as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def forward(Ws, bs, X, round_bf16=True):
    """Ws[l]: [out, in], bs[l]: [out]; X: [B, in0].  Returns (out [B, out_last], cache)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    h = r(X)
    cache = {"h": [h], "z": []}
    L = len(Ws)
    for li in range(L):
        z = h @ r(Ws[li]).T + bs[li]
        cache["z"].append(z)
        if li < L - 1:
            h = r(gelu(z))
            cache["h"].append(h)
        else:
            out = z
    return out, cache


def loss_and_grads(Ws, bs, X, T, nused, round_bf16=True):
    """MSE over the first nused output columns (mean over B*nused) and its gradients."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    out, cache = forward(Ws, bs, X, round_bf16)
    B = X.shape[0]
    d = np.zeros_like(out)
    d[:, :nused] = out[:, :nused] - T[:, :nused]
    loss = float((d[:, :nused] ** 2).sum()) / (B * nused)
    dz = r(2.0 * d / (B * nused))
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for li in range(len(Ws) - 1, -1, -1):
        h = cache["h"][li]
        gW[li] = dz.T @ h
        gb[li] = dz.sum(axis=0)
        if li == 0:
            break
        dh = dz @ r(Ws[li])
        dz = r(dh * gelu_grad(cache["z"][li - 1]))
    return loss, gW, gb
