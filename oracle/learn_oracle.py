"""CPU restatement (numpy, FP64) of the learned warm start's network: feature
layout, 2-layer MLP forward pass with BF16 weights and activations, the
phase-invariant subspace loss and its gradient.

TEST INFRASTRUCTURE ONLY: tests, the smoke check and the benchmark's CPU
baseline may call this module, the product never does.

The reference has no learned component (it calls cusolverDnXsyevBatched,
lambda_adaptive_beamformed_spectra_pipeline.hpp:679-686); the features restate
its covariance extraction and whitening (spatial.cuh:672-713, 1065-1093).
Pinning: the analytic gradient is checked against central finite
differences (tests/test_learn_oracle.py), the feature layout against the
covariance built by the C oracle, and BF16 rounding against bit patterns.
"""
from __future__ import annotations

import numpy as np


def round_up(x, m):
    return (x + m - 1) // m * m


def dims(n, KW=4, H=256):
    Dx = n * (n - 1) + 2 * KW * n
    Dy = 2 * KW * n
    d = dict(KW=KW, n=n, Dx=Dx, Dp=round_up(Dx, 32), H=H, Dy=Dy, Dyp=round_up(Dy, 32))
    d["off_b1"] = H * d["Dp"]
    d["off_W2"] = d["off_b1"] + H
    d["off_b2"] = d["off_W2"] + d["Dyp"] * H
    d["P"] = d["off_b2"] + d["Dyp"]
    return d


# ---- BF16 (round to nearest even, as the kernels' conversion) ----

def bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)
    return u.astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << np.uint32(16)).view(np.float32)


def bf16_round(x):
    return bf16_value(bf16_bits(x))


# ---- features ----

def features(A_white, basis, d):
    """BF16 feature bits of one problem: A_white (n x n complex, whitened,
    A[i, j] = V[bl(i, j)] / (d_i d_j) for i < j), basis (KW x n complex)."""
    n, KW = d["n"], d["KW"]
    x = np.zeros(d["Dp"], np.float32)
    e = 0
    for j in range(1, n):
        for i in range(j):
            x[2 * e] = A_white[i, j].real
            x[2 * e + 1] = A_white[i, j].imag
            e += 1
    off = n * (n - 1)
    # componentwise (a complex product would turn a -0 imaginary part into +0)
    b = np.asarray(basis).reshape(KW * n)
    x[off:off + 2 * KW * n:2] = np.sqrt(n) * b.real
    x[off + 1:off + 2 * KW * n:2] = np.sqrt(n) * b.imag
    return bf16_bits(x)


def whiten(A):
    """diagonalWhiten (spatial.cuh:1065-1093) on a Hermitian covariance."""
    dd = np.sqrt(np.maximum(A.diagonal().real, 1e-10))
    return A / np.outer(dd, dd), dd


# ---- network ----

def unpack(params, d):
    H, Dp, Dyp = d["H"], d["Dp"], d["Dyp"]
    W1 = params[:d["off_b1"]].reshape(H, Dp)
    b1 = params[d["off_b1"]:d["off_W2"]]
    W2 = params[d["off_W2"]:d["off_b2"]].reshape(Dyp, H)
    b2 = params[d["off_b2"]:d["P"]]
    return W1, b1, W2, b2


def forward(params, d, xbits, bf16=True):
    """Z1 = X W1^T + b1, H = bf16(relu(Z1)), Y = H W2^T + b2 + skip, where the
    skip adds the previous block's vectors held in X back (FP64 sums of
    BF16 operands; bf16=False keeps everything in FP64 for gradient checks)."""
    W1, b1, W2, b2 = unpack(np.asarray(params, np.float64), d)
    if bf16:
        X = bf16_value(xbits).astype(np.float64)
        W1 = bf16_round(W1).astype(np.float64)
        W2 = bf16_round(W2).astype(np.float64)
    else:
        X = np.asarray(xbits, np.float64)
    Z1 = X @ W1.T + b1
    Hh = np.maximum(Z1, 0.0)
    if bf16:
        Hh = bf16_round(Hh).astype(np.float64)
    Y = Hh @ W2.T + b2
    # skip connection: the previous block's vectors held in the features
    off, Dy = d["n"] * (d["n"] - 1), d["Dy"]
    Y[:, :Dy] += X[:, off:off + Dy] / np.sqrt(d["n"])
    return X, W1, W2, Z1, Hh, Y


def loss_and_grad(params, d, xbits, labels, K, bf16=True, eps=1e-12):
    """Loss (1/M) sum_m sum_{k<K_m} (1 - |u_k^H y_k|^2 / (|y_k|^2 + eps)) and
    its gradient with respect to the parameters (straight-through the BF16
    rounding of weights and activations)."""
    n, KW = d["n"], d["KW"]
    X, W1, W2, Z1, Hh, Y = forward(params, d, xbits, bf16)
    M = X.shape[0]
    y = Y[:, :2 * KW * n].reshape(M, KW, n, 2)
    y = y[..., 0] + 1j * y[..., 1]
    u = np.asarray(labels, np.float64).reshape(M, KW, n, 2)
    u = u[..., 0] + 1j * u[..., 1]
    on = np.arange(KW)[None, :] < np.minimum(np.asarray(K), KW)[:, None]
    s = np.sum(np.conj(u) * y, axis=2)                    # [M][KW]
    den = np.sum(np.abs(y) ** 2, axis=2) + eps
    f = np.abs(s) ** 2 / den
    loss = np.sum(np.where(on, 1.0 - f, 0.0)) / M
    g = -(2.0 / M) * (s[..., None] * u - f[..., None] * y) / den[..., None]
    g = np.where(on[..., None], g, 0.0)
    dY = np.zeros_like(Y)
    dY[:, :2 * KW * n] = np.stack([g.real, g.imag], -1).reshape(M, 2 * KW * n)
    dW2 = dY.T @ Hh
    db2 = dY.sum(0)
    dH = (dY @ W2) * (Z1 > 0)
    dW1 = dH.T @ X
    db1 = dH.sum(0)
    grad = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2])
    return loss, grad
