"""Learned closure on the GPU: MFMA lane layout, every convolution pass
against the CPU oracle, whole-network gradients against the oracle chain,
and (with the committed trained weights) the gain over the plain coarse
solver as judged by the fine-grid solver."""
import ctypes as C
import os

import numpy as np
import pytest

import closure_oracle as co

pytestmark = pytest.mark.gpu
PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "aeromhd-cuda_amd")


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_mfma_lane_layout(torch_gpu):
    import amhd_lib

    torch = torch_gpu
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 32)).astype(np.float32)
    B = rng.standard_normal((32, 16)).astype(np.float32)
    At, Bt = torch.tensor(A, device="cuda"), torch.tensor(B, device="cuda")
    D = torch.empty(16, 16, device="cuda")
    amhd_lib.call("amhd_mfma_probe", _p(At), _p(Bt), _p(D), amhd_lib.stream_handle(torch))
    ref = co.bf16_round(A).astype(np.float64) @ co.bf16_round(B).astype(np.float64)
    np.testing.assert_allclose(D.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("cin,cout,relu,last", [(16, 32, True, False), (32, 32, True, False), (32, 32, False, True)])
def test_conv_passes_vs_oracle(torch_gpu, cin, cout, relu, last):
    import amhd_lib

    torch = torch_gpu
    rng = np.random.default_rng(cin + 3 * cout)
    n, nx, ny = 2, 9, 7
    x = co.bf16_round(np.abs(rng.standard_normal((n, nx, ny, cin))) * (rng.random((n, nx, ny, cin)) > 0.3))
    W = np.zeros((cout, co.kpad(cin)), dtype=np.float32)
    W[:, :9 * cin] = rng.standard_normal((cout, 9 * cin)) * 0.2
    b = rng.standard_normal(cout).astype(np.float32)
    xt = torch.tensor(x, device="cuda").to(torch.bfloat16)
    Wt = torch.tensor(W, device="cuda").to(torch.bfloat16)
    bt = torch.tensor(b, device="cuda")
    y = torch.empty(n, nx, ny, cout, device="cuda", dtype=torch.float32 if last else torch.bfloat16)
    s = amhd_lib.stream_handle(torch)
    amhd_lib.call("amhd_conv3x3_fwd", _p(xt), n, nx, ny, cin, _p(Wt), _p(bt), cout, int(relu), _p(y), int(last), s)
    ref = co.conv_fwd(x, W, b, relu)
    got = y.float().cpu().numpy()
    if last:
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5)
    else:   # stored in BF16: at most one BF16 ulp from the rounded oracle
        assert np.all(np.abs(got - co.bf16_round(ref)) <= 2.0 ** -7 * np.abs(ref) + 1e-6)
    # backward data with the ReLU mask of the layer input, and weights
    dy = rng.standard_normal((n, nx, ny, cout)).astype(np.float32)
    dyt = torch.tensor(dy, device="cuda")
    WT = W[:, :9 * cin].reshape(cout, 9, cin).transpose(2, 1, 0).reshape(cin, 9 * cout)
    WTp = np.zeros((cin, co.kpad(cout)), dtype=np.float32)
    WTp[:, :9 * cout] = WT
    WTt = torch.tensor(WTp, device="cuda").to(torch.bfloat16)
    dx = torch.empty(n, nx, ny, cin, device="cuda")
    amhd_lib.call("amhd_conv3x3_bwd_data", _p(dyt), n, nx, ny, cout, _p(WTt), cin, _p(xt), _p(dx), s)
    np.testing.assert_allclose(dx.cpu().numpy(), co.conv_bwd_data(dy, W, cin, mask=x), rtol=1e-4, atol=1e-4)
    dW = torch.empty(cout, co.kpad(cin), device="cuda")
    db = torch.empty(cout, device="cuda")
    ws = torch.empty(5 * cout * co.kpad(cin), device="cuda")
    amhd_lib.call("amhd_conv3x3_bwd_weight", _p(xt), _p(dyt), n, nx, ny, cin, cout, _p(dW), _p(db), _p(ws), 5, s)
    rW, rb = co.conv_bwd_weight(x, dy)
    np.testing.assert_allclose(dW.cpu().numpy(), rW, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(db.cpu().numpy(), rb, rtol=1e-5, atol=1e-4)


def test_network_gradients_vs_oracle(torch_gpu):
    """The network through the HIP kernels vs the CPU oracle chain, one piece
    at a time: the forward (hidden activations are stored in BF16, so a
    value can land one BF16 ulp away: norm-wise 1e-2), the loss gradient at
    the GPU's own output (1e-4), and every weight / bias gradient from the
    GPU's own activations and output gradient (norm-wise 1e-3)."""
    torch = torch_gpu
    import amhd_closure as cl

    rng = np.random.default_rng(5)
    n, nx, ny = 2, 16, 12
    model = cl.FluxCNN(seed=3)
    X = torch.tensor(rng.standard_normal((n, 9, nx, ny)), device="cuda", dtype=torch.float32)
    R = torch.tensor(rng.standard_normal((n, 9, nx, ny)), device="cuda", dtype=torch.float32)
    model.stats = {"in_mean": np.zeros(9, np.float32), "in_std": np.ones(9, np.float32),
                   "r_std": rng.uniform(0.5, 2, 9).astype(np.float32),
                   "f_scale": rng.uniform(0.5, 2, 9).astype(np.float32)}
    st = model.stats
    dx, dy = 0.25, 0.2
    loss, dout, acts = cl.loss_and_grad(model, X, R, st, dx, dy)
    gW, gb = model.backward(acts, dout, n_split=3)
    Ws = [W.cpu().numpy() for W in model.W]
    bs = [b.cpu().numpy() for b in model.b]
    gacts = [a.float().cpu().numpy().astype(np.float64) for a in acts]
    out_gpu = model.forward(acts[0]).cpu().numpy().astype(np.float64)
    out_ref, _ = co.net_forward(gacts[0], Ws, bs)
    assert np.linalg.norm(out_gpu - out_ref) <= 1e-2 * np.linalg.norm(out_ref)
    Rn = R.cpu().numpy().astype(np.float64)
    assert loss == pytest.approx(co.loss(out_gpu, Rn, st["f_scale"], st["r_std"], dx, dy), rel=1e-5)
    dref = co.loss_grad_out(out_gpu, Rn, st["f_scale"].astype(np.float64), st["r_std"].astype(np.float64), dx, dy)
    np.testing.assert_allclose(dout.cpu().numpy(), dref, rtol=1e-4, atol=1e-9)
    rW, rb = co.net_backward(gacts, dout.cpu().numpy().astype(np.float64), Ws)
    for li in range(len(co.LAYERS)):
        g, r = gW[li].cpu().numpy(), rW[li] * model.wmask[li].cpu().numpy()
        assert np.linalg.norm(g - r) <= 1e-3 * np.linalg.norm(r), li
        assert np.linalg.norm(gb[li].cpu().numpy() - rb[li]) <= 1e-3 * np.linalg.norm(rb[li]) + 1e-9, li


def test_closure_gain_over_plain_coarse(torch_gpu):
    """The trained network (closure_weights.npz) against the plain coarse
    solver on held-out seeds, judged by the fine 2048x1024 solver: lower L1
    density error; the mixing width is reported."""
    import amhd_closure as cl

    path = os.path.join(PKG, "closure_weights.npz")
    assert os.path.exists(path), "closure_weights.npz (trained weights) missing"
    model = cl.FluxCNN.load(path)
    rows = cl.judge([210, 217], model, log=print)
    for r in rows:
        assert r["l1rho_closure"] < r["l1rho_plain"], r
        assert r["fallbacks_closure"] < r["steps_closure"], r
