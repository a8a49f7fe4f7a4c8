"""CPU oracle of the learned nonlinear column step (test infrastructure only).

The reference has no learned component; this is the checker for the gfx950
network kernels.  The network stands in for one dZ of the KT update of a real
t-column u (ASS:833-839 is what it imitates):

    A = max_t |u|,  x = u / A,  features X0[t] = (x, x^2, 0, ..., 0)  (32 channels)
    X_j = relu(conv(X_{j-1}, W_j) + b_j),  j = 1, 2, 3     (32 channels, kernel 9)
    y   = conv(X_3, W_4) + b_4  (channel 0)
    u_new = u + c A^2 y,  c = dZ N / dT

conv is the zero-padded cross-correlation out[t] = sum_k W[:, :, k] in[t + k - 4]
(torch.nn.Conv1d with padding 4).  The gfx950 kernels store X_0..X_3 and the
back-propagated dZ_j in bfloat16 and accumulate in float32; ``quant=True``
rounds to bfloat16 at exactly those points, so the oracle's gradients are the
ones the kernels should produce up to float32 summation order.  Layer 1 uses
input channels 0-1 only and layer 4 output channel 0 only; the other entries
of W_1, W_4, b_4 are held at zero.

Loss: mean over the masked samples of (y - y*)^2.
"""
from __future__ import annotations

import numpy as np

C = 32      # channels
KW = 9      # kernel width
HALF = KW // 2
NLAYER = 4
W_SIZE = C * C * KW
LAYER_SIZE = W_SIZE + C
PARAM_COUNT = NLAYER * LAYER_SIZE


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    bias = 0x7FFF + ((b >> 16) & 1)
    r = ((b + bias) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).copy()
    nan = np.isnan(a)
    out[nan] = a[nan]
    return out


def unpack(params: np.ndarray):
    """Flat parameter vector -> [(W (C, C, KW), b (C,))] * 4 (views)."""
    p = np.asarray(params)
    layers = []
    for j in range(NLAYER):
        o = j * LAYER_SIZE
        layers.append((p[o:o + W_SIZE].reshape(C, C, KW), p[o + W_SIZE:o + LAYER_SIZE]))
    return layers


def trainable_mask() -> np.ndarray:
    """1 where a parameter is trained, 0 where it is held at zero."""
    m = np.ones(PARAM_COUNT, dtype=np.float32)
    (W1, _), _, _, (W4, b4) = unpack(m)
    W1[:, 2:, :] = 0.0
    W4[1:, :, :] = 0.0
    b4[1:] = 0.0
    return m


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation of the trainable entries (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAM_COUNT, dtype=np.float32)
    fan = [2 * KW, C * KW, C * KW, C * KW]
    for j, (W, b) in enumerate(unpack(p)):
        W[...] = rng.normal(0.0, np.sqrt(2.0 / fan[j]), size=W.shape).astype(np.float32)
        b[...] = 0.0
    p *= trainable_mask()
    (_, _), _, _, (W4, _) = unpack(p)
    W4 *= 0.1
    return p


def features(u: np.ndarray, quant: bool = True):
    """u (B, T) float32 -> (X0 (B, T, C), A (B,)): x = u / A and x^2 in channels 0, 1."""
    u = np.asarray(u, dtype=np.float32)
    A = np.max(np.abs(u), axis=1).astype(np.float32)
    As = np.where(A > 0, A, np.float32(1.0)).astype(np.float32)
    x = (u / As[:, None]).astype(np.float32)
    X = np.zeros(u.shape + (C,), dtype=np.float32)
    X[..., 0] = x
    X[..., 1] = (x * x).astype(np.float32)
    return (bf16(X) if quant else X), A


def conv(X: np.ndarray, W: np.ndarray) -> np.ndarray:
    """Zero-padded cross-correlation: X (B, T, Ci), W (Co, Ci, KW) -> (B, T, Co), float64 sums."""
    B, T, _ = X.shape
    Xp = np.zeros((B, T + 2 * HALF, X.shape[2]), dtype=np.float64)
    Xp[:, HALF:HALF + T] = X
    out = np.zeros((B, T, W.shape[0]), dtype=np.float64)
    for k in range(KW):
        out += Xp[:, k:k + T, :] @ W[:, :, k].T.astype(np.float64)
    return out


def conv_t(D: np.ndarray, W: np.ndarray) -> np.ndarray:
    """Adjoint of ``conv`` in its input: D (B, T, Co) -> (B, T, Ci)."""
    B, T, _ = D.shape
    Dp = np.zeros((B, T + 2 * HALF, D.shape[2]), dtype=np.float64)
    Dp[:, HALF:HALF + T] = D
    out = np.zeros((B, T, W.shape[1]), dtype=np.float64)
    for k in range(KW):
        # out[t] += W[:, :, k]^T D[t - k + HALF]
        out += Dp[:, 2 * HALF - k:2 * HALF - k + T, :] @ W[:, :, k].astype(np.float64)
    return out


def conv_w(D: np.ndarray, X: np.ndarray) -> np.ndarray:
    """Gradient of sum(D * conv(X, W)) in W: (Co, Ci, KW)."""
    B, T, _ = X.shape
    Xp = np.zeros((B, T + 2 * HALF, X.shape[2]), dtype=np.float64)
    Xp[:, HALF:HALF + T] = X
    g = np.zeros((D.shape[2], X.shape[2], KW), dtype=np.float64)
    Dm = D.reshape(-1, D.shape[2]).astype(np.float64)
    for k in range(KW):
        g[:, :, k] = Dm.T @ Xp[:, k:k + T, :].reshape(-1, X.shape[2])
    return g


def _weights(layers, quant: bool):
    """Weights as the kernels use them: bfloat16-rounded (quant) or exact float64."""
    if quant:
        return [bf16(W.astype(np.float32)) for W, _ in layers]
    return [W.astype(np.float64) for W, _ in layers]


def forward(params: np.ndarray, X0: np.ndarray, quant: bool = True):
    """Returns ([X0, X1, X2, X3], y (B, T) float64)."""
    layers = unpack(params)
    wq = _weights(layers, quant)
    acts = [X0]
    X = X0
    for j in range(3):
        Z = conv(X, wq[j]) + layers[j][1].astype(np.float64)
        X = bf16(np.maximum(Z, 0.0).astype(np.float32)) if quant else np.maximum(Z, 0.0)
        acts.append(X)
    y = (conv(X, wq[3]) + layers[3][1].astype(np.float64))[..., 0]
    return acts, y


def loss_and_grads(params: np.ndarray, acts, y: np.ndarray, ystar: np.ndarray, mask: np.ndarray,
                   quant: bool = True):
    """MSE loss over ``mask`` and its gradient in the flat parameters (float64)."""
    layers = unpack(params)
    wq = _weights(layers, quant)
    m = np.asarray(mask, dtype=np.float64)
    cnt = max(float(m.sum()), 1.0)
    r = (y - ystar) * m
    loss = float((r * r).sum() / cnt)
    dy = (2.0 / cnt) * r
    grad = np.zeros(PARAM_COUNT, dtype=np.float64)
    gl = unpack(grad)
    D = np.zeros(y.shape + (C,), dtype=np.float64)
    D[..., 0] = dy
    if quant:
        D = bf16(D.astype(np.float32))
    for j in (3, 2, 1, 0):
        gW, gb = gl[j]
        gW[...] = conv_w(D, acts[j])
        gb[...] = D.astype(np.float64).sum(axis=(0, 1))
        if j > 0:
            dX = conv_t(D, wq[j]) * (acts[j] > 0)
            D = bf16(dX.astype(np.float32)) if quant else dX
    grad *= trainable_mask()
    return loss, grad


def apply(params: np.ndarray, u: np.ndarray, c: float, quant: bool = True, chunk: int = 0, halo: int = 16,
          quiet: float = 1e-4, negligible: float = 1e-5) -> np.ndarray:
    """u_new = u + c A^2 y for every column of u (B, T).

    As the fused kernel: a column whose largest step D = max |u[t+1] - u[t]|
    has c D < ``negligible`` is unchanged (its Burgers increment, about c A D,
    is below that fraction of A), and an output chunk [t0, t0 + chunk) whose
    window [t0 - halo, t0 + chunk + halo) has max |u / A| below ``quiet`` is
    unchanged.  The kernel's segment is 128 samples (``chunk=0``)."""
    u = np.asarray(u, dtype=np.float32)
    if chunk <= 0:
        chunk = min(u.shape[1], 128)
    X0, A = features(u, quant)
    _, y = forward(params, X0, quant)
    As = np.where(A > 0, A, np.float32(1.0)).astype(np.float32)
    x = np.abs(u / As[:, None])
    keep = np.zeros(u.shape, dtype=bool)
    T = u.shape[1]
    for t0 in range(0, T, chunk):
        lo, hi = max(t0 - halo, 0), min(t0 + chunk + halo, T)
        keep[:, t0:t0 + chunk] = (np.max(x[:, lo:hi], axis=1) >= quiet)[:, None]
    D = np.max(np.abs(np.diff(u, axis=1)), axis=1)
    keep &= (np.float32(c) * D >= np.float32(negligible))[:, None]
    inc = c * (A.astype(np.float64) ** 2)[:, None] * y * keep
    return (u.astype(np.float64) + inc).astype(np.float32)


def max_dpdt(u: np.ndarray, dT: float) -> np.ndarray:
    """Fallback indicator max_t |u[t+1] - u[t]| / dT per column (Pa/s)."""
    u = np.asarray(u, dtype=np.float32)
    return np.max(np.abs(np.diff(u, axis=1)), axis=1).astype(np.float64) / dT
