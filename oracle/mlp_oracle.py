"""CPU oracle of the warm-start MLP (test infrastructure only).

The network: 16 -> 64 -> 64 -> 16, tanh, tanh, linear; two outputs trained.
Parameter layout (FP32): W1t[64][16], b1[64], W2t[64][64], b2[64],
W3t[16][64], b3[16]; weights stored [out][in].

mode="bf16" rounds where the gfx950 kernels round (round to nearest even):
inputs, weights, the activations fed to the next layer, and the
back-propagated dZ used as matrix-core operands; sums, tanh, the output and
bias gradients stay unrounded.  mode="exact" rounds nowhere (float64): its
gradient is pinned to central finite differences of its loss in the tests.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT, USED = 16, 64, 16, 2
W1, B1 = 0, HID * IN
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + OUT * HID
NPARAMS = B3 + OUT


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> BF16 bit patterns, round to nearest even (inf/nan kept)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    s = ((u >> 16) | np.where((u & 0xFFFF) != 0, 0x40, 0).astype(np.uint64)) & 0xFFFF
    return np.where(special, s, r).astype(np.uint16)


def bf16_value(bits: np.ndarray) -> np.ndarray:
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def round_bf16(x: np.ndarray) -> np.ndarray:
    return bf16_value(bf16_bits(np.asarray(x, np.float32)))


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    return (p[W1:B1].reshape(HID, IN), p[B1:W2], p[W2:B2].reshape(HID, HID), p[B2:W3],
            p[W3:B3].reshape(OUT, HID), p[B3:NPARAMS])


def forward(params, x, mode="bf16"):
    """x: (n, 16) float (already BF16 values in bf16 mode).  Returns the
    outputs (n, 2) and the intermediates for backward."""
    rnd = round_bf16 if mode == "bf16" else (lambda a: np.asarray(a, np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(params)
    w1r, w2r, w3r = rnd(w1), rnd(w2), rnd(w3)
    xt = np.asarray(x, np.float64).T                      # (16, n)
    z1 = w1r @ xt + b1[:, None]
    h1 = np.tanh(z1)
    h1r = rnd(h1)
    z2 = w2r @ h1r + b2[:, None]
    h2 = np.tanh(z2)
    h2r = rnd(h2)
    z3 = w3r @ h2r + b3[:, None]
    cache = dict(xt=xt, h1=h1, h1r=h1r, h2=h2, h2r=h2r, z3=z3, w2r=w2r, w3r=w3r, rnd=rnd)
    return z3[:USED].T.copy(), cache


def loss_and_grad(params, x, t, mode="bf16"):
    """Loss sum_nodes sum_{o<2} (y - t)^2 / n and its gradient (parameter layout)."""
    n = x.shape[0]
    y, c = forward(params, x, mode)
    rnd = c["rnd"]
    d = y - np.asarray(t, np.float64)
    loss = float((d * d).sum() / n)
    dz3 = np.zeros((OUT, n))
    dz3[:USED] = 2.0 * d.T / n
    dz3r = rnd(dz3)
    dz2 = (c["w3r"].T @ dz3r) * (1.0 - c["h2"] ** 2)
    dz2r = rnd(dz2)
    dz1 = (c["w2r"].T @ dz2r) * (1.0 - c["h1"] ** 2)
    dz1r = rnd(dz1)
    g = np.zeros(NPARAMS)
    g[W1:B1] = (dz1r @ c["xt"].T).reshape(-1)
    g[B1:W2] = dz1.sum(axis=1)
    g[W2:B2] = (dz2r @ c["h1r"].T).reshape(-1)
    g[B2:W3] = dz2.sum(axis=1)
    g[W3:B3] = (dz3r @ c["h2r"].T).reshape(-1)
    g[B3:NPARAMS] = dz3.sum(axis=1)
    return loss, g
