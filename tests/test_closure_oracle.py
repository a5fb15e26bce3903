"""Pins of the closure convolution oracle against an independent library
(torch.nn.functional.conv2d and autograd on the CPU, float64) and against
the adjoint identity."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

import closure_oracle as co


def torch_conv(x, W, b, cin, cout):
    """NHWC x; W[co, tap*cin+ci] -> conv2d with zero pad in x, circular in y."""
    xt = torch.as_tensor(x).permute(0, 3, 1, 2)                      # N C H W
    xt = F.pad(xt, (1, 1, 0, 0), mode="circular")                     # y (W axis) periodic
    xt = F.pad(xt, (0, 0, 1, 1))                                      # x (H axis) zeros
    w = torch.as_tensor(W[:, :9 * cin]).reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
    return F.conv2d(xt, w, torch.as_tensor(b)).permute(0, 2, 3, 1)


@pytest.mark.parametrize("cin,cout", [(16, 32), (32, 32), (32, 16)])
def test_fwd_and_backward_match_torch(cin, cout):
    rng = np.random.default_rng(cin + cout)
    x = rng.standard_normal((2, 7, 6, cin))
    W = np.zeros((cout, co.kpad(cin)))
    W[:, :9 * cin] = rng.standard_normal((cout, 9 * cin)) * 0.2
    b = rng.standard_normal(cout)
    y = co.conv_fwd(x, W, b, relu=False, bf16=False)
    xt = torch.tensor(x, requires_grad=True)
    Wt = torch.tensor(W, requires_grad=True)
    bt = torch.tensor(b, requires_grad=True)
    yt = torch_conv(xt, Wt, bt, cin, cout)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    g = rng.standard_normal(y.shape)
    yt.backward(torch.tensor(g))
    np.testing.assert_allclose(co.conv_bwd_data(g, W, cin, bf16=False), xt.grad.numpy(), rtol=1e-12, atol=1e-12)
    dW, db = co.conv_bwd_weight(x, g, bf16=False)
    np.testing.assert_allclose(dW, Wt.grad.numpy(), rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_adjoint_identity():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 5, 9, 16))
    W = rng.standard_normal((32, co.kpad(16)))
    y = rng.standard_normal((1, 5, 9, 32))
    lhs = np.sum(co.conv_fwd(x, W, np.zeros(32), relu=False, bf16=False) * y)
    rhs = np.sum(x * co.conv_bwd_data(y, W, 16, bf16=False))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.0e38, 1e-40, np.inf], dtype=np.float32)
    r = co.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125   # ties to even, then up
    assert np.isinf(r[5]) and np.all(np.isfinite(r[:5]))
    t = torch.tensor(v).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(r, t)


def test_network_gradient_finite_differences():
    """The oracle's network backward (float64, no rounding) against central
    differences of the closure loss."""
    rng = np.random.default_rng(11)
    n, nx, ny = 1, 6, 5
    x = np.zeros((n, nx, ny, 16))
    x[..., :9] = rng.standard_normal((n, nx, ny, 9))
    Ws, bs = [], []
    for cin, cout, _ in co.LAYERS:
        W = np.zeros((cout, co.kpad(cin)))
        W[:, :9 * cin] = rng.standard_normal((cout, 9 * cin)) * 0.15
        Ws.append(W)
        bs.append(rng.standard_normal(cout) * 0.1)
    R = rng.standard_normal((n, 9, nx, ny))
    fscale, rstd, dx, dy = rng.uniform(0.5, 2, 9), rng.uniform(0.5, 2, 9), 0.3, 0.2
    out, acts = co.net_forward(x, Ws, bs, bf16=False)
    gW, gb = co.net_backward(acts, co.loss_grad_out(out, R, fscale, rstd, dx, dy), Ws, bf16=False)

    def L():
        return co.loss(co.net_forward(x, Ws, bs, bf16=False)[0], R, fscale, rstd, dx, dy)

    h = 1e-6
    for li, (cin, cout, _) in enumerate(co.LAYERS):
        for (a, b) in ((0, 0), (cout - 1, 9 * cin - 1), (3, 40)):
            old = Ws[li][a, b]
            Ws[li][a, b] = old + h
            lp = L()
            Ws[li][a, b] = old - h
            lm = L()
            Ws[li][a, b] = old
            assert gW[li][a, b] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-9), (li, a, b)
        old = bs[li][1]
        bs[li][1] = old + h
        lp = L()
        bs[li][1] = old - h
        lm = L()
        bs[li][1] = old
        assert gb[li][1] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-9)
