"""CPU oracle of the learned-closure convolutions -- TEST INFRASTRUCTURE ONLY.

NumPy restatement of the three 3x3 convolution passes that
aeromhd-cuda_amd/csrc/amhd_cnn.hip runs on BF16 matrix cores:

* conv_fwd:         y[p, co] = act(b[co] + sum_{tap, ci} x[p + off(tap), ci] W[co, tap*Cin + ci])
* conv_bwd_data:    dx[p, ci] = mask * sum_{tap, co} dy[p - off(tap), co] W[co, tap*Cin + ci]
* conv_bwd_weight:  dW[co, tap*Cin + ci] = sum_p dy[p, co] x[p + off(tap), ci];  db = sum_p dy

on NHWC arrays (n, nx, ny, C), zero padding in x and periodic in y.  With
``bf16=True`` the operands are rounded to BF16 exactly where the kernels
round them (activations and weights are BF16 tensors; FP32 gradients are
rounded when loaded into the MFMA); sums are in float64.

Pins (tests/test_closure_oracle.py): the forward equals
torch.nn.functional.conv2d on the CPU (float64) with the same padding; the
backward passes equal torch autograd of that conv2d; the adjoint identity
<conv(x), y> = <x, conv^T(y)> holds; a finite-difference check of the whole
closure loss.
"""
from __future__ import annotations

import numpy as np

TAPS = [(di, dj) for di in (-1, 0, 1) for dj in (-1, 0, 1)]   # tap = (di+1)*3 + (dj+1)


def bf16_round(a) -> np.ndarray:
    """float32 -> nearest-even BF16 -> float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, (u >> 16) << 16, r).astype(np.uint32)
    return r.view(np.float32)


def kpad(cin: int) -> int:
    return (9 * cin + 31) // 32 * 32


def shift(x: np.ndarray, di: int, dj: int) -> np.ndarray:
    """out[n, i, j] = x[n, i + di, (j + dj) mod ny], zero where i + di is
    outside [0, nx)."""
    n, nx, ny, c = x.shape
    out = np.zeros_like(x)
    src = np.roll(x, -dj, axis=2)
    lo, hi = max(0, -di), min(nx, nx - di)
    out[:, lo:hi] = src[:, lo + di:hi + di]
    return out


def conv_fwd(x, W, b, relu: bool, bf16: bool = True):
    x = np.asarray(x, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    if bf16:
        x = bf16_round(x).astype(np.float64)
        W = bf16_round(W).astype(np.float64)
    cin = x.shape[3]
    y = np.zeros(x.shape[:3] + (W.shape[0],))
    for t, (di, dj) in enumerate(TAPS):
        y += shift(x, di, dj) @ W[:, t * cin:(t + 1) * cin].T
    y += np.asarray(b, dtype=np.float64)
    return np.maximum(y, 0.0) if relu else y


def conv_bwd_data(dy, W, cin: int, mask=None, bf16: bool = True):
    dy = np.asarray(dy, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    if bf16:
        dy = bf16_round(dy).astype(np.float64)
        W = bf16_round(W).astype(np.float64)
    dx = np.zeros(dy.shape[:3] + (cin,))
    for t, (di, dj) in enumerate(TAPS):
        dx += shift(dy, -di, -dj) @ W[:, t * cin:(t + 1) * cin]
    if mask is not None:
        dx = np.where(np.asarray(mask) > 0, dx, 0.0)
    return dx


def conv_bwd_weight(x, dy, bf16: bool = True):
    x = np.asarray(x, dtype=np.float64)
    dy64 = np.asarray(dy, dtype=np.float64)
    dyr = bf16_round(dy64).astype(np.float64) if bf16 else dy64
    if bf16:
        x = bf16_round(x).astype(np.float64)
    cin, cout = x.shape[3], dy64.shape[3]
    dW = np.zeros((cout, kpad(cin)))
    a = dyr.reshape(-1, cout)
    for t, (di, dj) in enumerate(TAPS):
        dW[:, t * cin:(t + 1) * cin] = a.T @ shift(x, di, dj).reshape(-1, cin)
    return dW, dy64.reshape(-1, cout).sum(axis=0)


# ------------------------------------------------------------------ network
LAYERS = [(16, 32, True), (32, 32, True), (32, 32, True), (32, 32, False)]
FEAT = [0, 1, 2, 3, 4, 5, 6, 7, 9]


def net_forward(x, Ws, bs, bf16: bool = True):
    """Layer inputs and the FP32-accumulated output of the 4-layer net;
    hidden activations are stored in BF16 (as the kernels store them)."""
    acts = [np.asarray(x, dtype=np.float64)]
    h = acts[0]
    for li, (cin, cout, relu) in enumerate(LAYERS):
        y = conv_fwd(h, Ws[li], bs[li], relu, bf16=bf16)
        if li < len(LAYERS) - 1:
            h = bf16_round(y).astype(np.float64) if bf16 else y
            acts.append(h)
        else:
            h = y
    return h, acts


def net_backward(acts, dout, Ws, bf16: bool = True):
    gW, gb = [None] * len(LAYERS), [None] * len(LAYERS)
    dy = np.asarray(dout, dtype=np.float64)
    for li in range(len(LAYERS) - 1, -1, -1):
        cin, cout, _ = LAYERS[li]
        gW[li], gb[li] = conv_bwd_weight(acts[li], dy, bf16=bf16)
        if li > 0:
            dy = conv_bwd_data(dy, Ws[li], cin, mask=acts[li], bf16=bf16)
    return gW, gb


def source_from_flux(dF, dx, dy):
    """S = -div(dF): channels 0..8 at x faces i-1/2 (zero on both x
    boundaries), 9..17 at y faces j-1/2 (periodic)."""
    n9 = len(FEAT)
    fx = dF[..., :n9]
    z = np.zeros_like(fx[:, :1])
    fx = np.concatenate([z, fx[:, 1:], z], axis=1)
    fy = dF[..., n9:2 * n9]
    return -(fx[:, 1:] - fx[:, :-1]) / dx - (np.roll(fy, -1, axis=2) - fy) / dy


def loss(out, R, fscale, rstd, dx, dy):
    """Normalised MSE of S(out * fscale) against R (n, 9, nx, ny)."""
    S = source_from_flux(out[..., :18] * np.concatenate([fscale, fscale]), dx, dy)
    e = (S - np.transpose(R, (0, 2, 3, 1))) / rstd
    return float(np.mean(e * e))


def loss_grad_out(out, R, fscale, rstd, dx, dy):
    """d loss / d out (adjoint of the divergence, by hand)."""
    n9 = len(FEAT)
    fs = np.concatenate([fscale, fscale])
    S = source_from_flux(out[..., :18] * fs, dx, dy)
    g = 2.0 * (S - np.transpose(R, (0, 2, 3, 1))) / rstd ** 2 / S.size
    gx = np.zeros_like(g)                    # d/d fx_i: S_i has +fx_i/dx, S_{i-1} has -fx_i/dx
    gx[:, 1:] = (g[:, 1:] - g[:, :-1]) / dx
    gy = (g - np.roll(g, 1, axis=2)) / dy
    d = np.zeros_like(out)
    d[..., :n9] = gx * fscale
    d[..., n9:2 * n9] = gy * fscale
    return d
